"""`manta` — the reference's Python scene API over the MI355X grid path.

Scene scripts written for mantaflow (``from manta import *``) run unchanged
for the part of the API this package covers: the solver, grid classes, flag
setup, Box/Sphere/Cylinder shapes and the per-step plugins advectSemiLagrange,
setWallBcs, addBuoyancy, setOpenBound, resetOutflow and solvePressure (with
computePressureRhs / solvePressureSystem / correctVelocity / releaseMG).
Plugin names, keyword names, defaults, positional order and error messages
follow the reference's wrapper rules (pwrapper/pconvert.cpp:459-505): keyword
arguments by C++ parameter name, else by position; an unknown keyword raises
"Argument 'x' unknown"; a missing required one "Argument 'x' is not defined.";
``notiming`` is accepted everywhere.  Grids live in GPU memory; every call
goes through libmfx's C ABI (include/mfx.h).

Run a scene with ``python -m mantaflow_amd scene.py`` (this module is then
importable as ``manta``).
"""
import ctypes as C
import math
import os
import sys
import time

import numpy as np

from . import _abi, gridio
from ._abi import check
from ._abi import lib as _lib_now

_lib_now()  # the GPU path is the only path: fail at import if libmfx.so is missing

# Held-back calls.  setWallBcs (without obvel, fractions or phiObs), the
# velocity correction that ends solvePressure and a Real grid's mult are
# kept for one call, so that the call the plume scenes make next runs them
# fused with its own pass over the grid (setWallBcs + addBuoyancy;
# correctVelocity + setWallBcs; mult + add of the same grid:
# mfx_set_wall_bcs_buoyancy, mfx_correct_velocity_wall_bcs,
# mfx_grid_mult_add, each bit-identical to the separate calls).  Every other call into the library,
# grid methods included, goes through lib(), which first runs the held-back
# call on its own, so no grid is ever read or written out of order.
_pending = None


def _flush():
    """Run the held-back call, if any."""
    global _pending
    p, _pending = _pending, None
    if p is None:
        return
    if p[0] == "wall":
        check(_lib_now().mfx_set_wall_bcs(p[1]._h, p[2]._h, None))
    elif p[0] == "mult":
        check(_lib_now().mfx_grid_mult(p[1]._h, p[2]._h))
    else:
        check(_lib_now().mfx_correct_velocity(p[2]._h, p[3]._h, p[1]._h))


def lib():
    """The library, after any held-back plugin call has run."""
    if _pending is not None:
        _flush()
    return _lib_now()

# ---------------------------------------------------------------------------
# constants (source/python/defines.py:15-50, pwrapper/registry.cpp:398-420)
# ---------------------------------------------------------------------------
Real = float
false = False
true = True
FlagFluid, FlagObstacle, FlagEmpty, FlagInflow, FlagOutflow, FlagOpen, FlagStick, FlagReserved = 1, 2, 4, 8, 16, 32, 64, 256
TypeNone, TypeFluid, TypeObstacle, TypeEmpty, TypeInflow, TypeOutflow, TypeOpen, TypeStick, TypeReserved = 0, 1, 2, 4, 8, 16, 32, 64, 256
IntEuler, IntRK2, IntRK4 = 0, 1, 2
PcNone, PcMIC, PcMGDynamic, PcMGStatic = 0, 1, 2, 3
GUI = False
DEBUG = False
MT = True
DOUBLEPRECISION = False
CUDA = False
HIP = True

_debug_level = 1


def _err(msg):
    raise RuntimeError(msg)


# ---------------------------------------------------------------------------
# vec3 (pwrapper/pvec3.cpp)
# ---------------------------------------------------------------------------
class vec3:
    __slots__ = ("x", "y", "z")

    def __init__(self, x=0.0, y=None, z=None):
        if isinstance(x, vec3):
            self.x, self.y, self.z = x.x, x.y, x.z
        elif y is None and z is None:
            self.x = self.y = self.z = float(x)
        else:
            self.x, self.y, self.z = float(x), float(y or 0.0), float(z or 0.0)

    def _o(self, o):
        return o if isinstance(o, vec3) else vec3(o)

    def __add__(self, o):
        o = self._o(o)
        return vec3(self.x + o.x, self.y + o.y, self.z + o.z)

    __radd__ = __add__

    def __sub__(self, o):
        o = self._o(o)
        return vec3(self.x - o.x, self.y - o.y, self.z - o.z)

    def __rsub__(self, o):
        return self._o(o) - self

    def __mul__(self, o):
        o = self._o(o)
        return vec3(self.x * o.x, self.y * o.y, self.z * o.z)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = self._o(o)
        return vec3(self.x / o.x, self.y / o.y, self.z / o.z)

    def __neg__(self):
        return vec3(-self.x, -self.y, -self.z)

    def __getitem__(self, i):
        return (self.x, self.y, self.z)[i]

    def __setitem__(self, i, v):
        setattr(self, "xyz"[i], float(v))

    def __len__(self):
        return 3

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __eq__(self, o):
        return isinstance(o, vec3) and (self.x, self.y, self.z) == (o.x, o.y, o.z)

    def __repr__(self):
        return "[%+4.6f, %+4.6f, %+4.6f]" % (self.x, self.y, self.z)


Vec3 = vec3


class vec4(vec3):
    pass


Vec4 = vec4


def _f32(v):
    return float(np.float32(v))


def _v3(v, name="value"):
    if isinstance(v, vec3):
        return (v.x, v.y, v.z)
    if isinstance(v, (tuple, list, np.ndarray)) and len(v) == 3:
        return tuple(float(a) for a in v)
    if isinstance(v, (int, float, np.floating, np.integer)):
        return (float(v),) * 3
    _err("Argument '%s' is not a vec3" % name)


# ---------------------------------------------------------------------------
# argument handling (pwrapper/pconvert.cpp:459-505)
# ---------------------------------------------------------------------------
_REQ = object()


def _parse(sig, args, kwargs):
    kwargs = dict(kwargs)
    kwargs.pop("notiming", None)
    out = {}
    names = [n for n, _ in sig]
    if len(args) > len(sig):
        _err("Function does not read argument number #%d" % len(sig))
    for i, (n, default) in enumerate(sig):
        if n in kwargs:
            out[n] = kwargs.pop(n)
        elif i < len(args):
            out[n] = args[i]
        elif default is _REQ:
            _err("Argument '%s' is not defined." % n)
        else:
            out[n] = default
    for k in kwargs:
        if k not in names:
            _err("Argument '%s' unknown" % k)
    return out


# ---------------------------------------------------------------------------
# timing (timing.h:22-56)
# ---------------------------------------------------------------------------
class _TimingData:
    def __init__(self):
        self.data = {}
        self.last = None

    def add(self, name, dt):
        n, t = self.data.get(name, (0, 0.0))
        self.data[name] = (n + 1, t + dt)


_timing = _TimingData()


class Timings:
    def __init__(self, **kw):
        pass

    def display(self):
        print("\n-- STEP TIMINGS --")
        for k, (n, t) in sorted(self.items()):
            print("  %-28s %8.3f ms   (%d calls, %.3f ms avg)" % (k, t * 1e3, n, t * 1e3 / max(n, 1)))

    def items(self):
        return list(_timing.data.items())

    def saveMean(self, file):
        with open(file, "w") as f:
            for k, (n, t) in sorted(self.items()):
                f.write("%s %g\n" % (k, t / max(n, 1)))


def _plugin(fn):
    name = fn.__name__

    def wrapper(*args, **kwargs):
        notiming = kwargs.get("notiming", False)
        t0 = time.perf_counter()
        try:
            return fn(*args, **kwargs)
        except RuntimeError as e:
            if _debug_level >= 1:
                print("Error in %s" % name, file=sys.stderr)
            raise RuntimeError(str(e)) from None
        finally:
            if not notiming:
                _timing.add(name, time.perf_counter() - t0)

    wrapper.__name__ = name
    wrapper.__doc__ = fn.__doc__
    return wrapper


# ---------------------------------------------------------------------------
# solver (fluidsolver.h:26-122, fluidsolver.cpp:107-215)
# ---------------------------------------------------------------------------
class FluidSolver:
    """Python class `Solver` (alias `FluidSolver`).  Owns one libmfx scene."""

    def __init__(self, *args, **kwargs):
        kwargs = dict(kwargs)
        name = kwargs.pop("name", "")
        a = _parse([("gridSize", _REQ), ("dim", 3), ("fourthDim", -1)], args, kwargs)
        gs = _v3(a["gridSize"], "gridSize")
        self.mGridSize = (int(gs[0]), int(gs[1]), int(gs[2]))
        dim = int(a["dim"])
        if dim == 4 and a["fourthDim"] > 0:
            _err("Don't create 4D solvers, use 3D with fourth-dim parameter >0 instead.")
        if not (dim == 2 or dim == 3):
            _err("Only 2D and 3D solvers allowed.")
        if dim == 2 and self.mGridSize[2] != 1:
            _err("Trying to create 2D solver with size.z != 1")
        self.mDim = dim
        self.name = name
        self._timestep = 1.0
        self.timeTotal = 0.0
        self.frame = 0
        self.cfl = 1000.0
        self.timestepMin = 1.0
        self.timestepMax = 1.0
        self.frameLength = 1.0
        self.timePerFrame = 0.0
        self._lockDt = False
        h = C.c_void_p()
        check(lib().mfx_scene_create(self.mGridSize[0], self.mGridSize[1], self.mGridSize[2], dim, C.byref(h)))
        self._h = h
        check(lib().mfx_scene_set_dt(self._h, C.c_float(_f32(1.0))))
        self._grids = []
        self._solve_stats = None
        # pressure-solve extensions (not kwargs of the reference): relative
        # stopping criterion, precision mode (1 = BF16 multigrid smoother
        # coefficients) and FP64 residual replacement every k iterations
        self.mfxRelTol = 0.0
        self.mfxPrecision = 0
        self.mfxTrueResidualEvery = 0

    # timestep property keeps the device copy in sync (Real = float32)
    @property
    def timestep(self):
        return self._timestep

    @timestep.setter
    def timestep(self, v):
        self._timestep = _f32(v)
        check(lib().mfx_scene_set_dt(self._h, C.c_float(self._timestep)))

    def is2D(self):
        return self.mDim == 2

    def is3D(self):
        return self.mDim == 3

    def getGridSize(self):
        return vec3(*self.mGridSize)

    def getDt(self):
        return self._timestep

    def getDx(self):
        return _f32(1.0 / max(self.mGridSize))

    def step(self):
        """FluidSolver::step (fluidsolver.cpp:142-158): time bookkeeping only."""
        self.timePerFrame = _f32(self.timePerFrame + self._timestep)
        self.timeTotal = _f32(self.timeTotal + self._timestep)
        if (self.timePerFrame + 1e-6) > self.frameLength:
            self.frame += 1
            self.timeTotal = _f32(float(self.frame) * self.frameLength)
            self.timePerFrame = 0.0
            self._lockDt = False

    def adaptTimestep(self, maxVel):
        """FluidSolver::adaptTimestep (fluidsolver.cpp:176-196)."""
        dt = self._timestep
        mvt = _f32(maxVel * dt)
        if not self._lockDt:
            dt = max(min(_f32(dt * _f32(self.cfl / (mvt + 1e-05))), self.timestepMax), self.timestepMin)
            if (self.timePerFrame + dt * 1.05) > self.frameLength:
                dt = _f32((self.frameLength - self.timePerFrame) + 1e-04)
            elif (self.timePerFrame + dt + self.timestepMin) > self.frameLength or \
                    (self.timePerFrame + (dt * 1.25)) > self.frameLength:
                dt = _f32((self.frameLength - self.timePerFrame + 1e-04) * 0.5)
                self._lockDt = True
        self.timestep = dt
        if not (self._timestep > (self.timestepMin / 2.)):
            _err("Invalid dt encountered! Shouldnt happen...")

    def create(self, *args, **kwargs):
        kwargs = dict(kwargs)
        name = kwargs.pop("name", "")
        kwargs.pop("show", None)
        if args:
            t, args = args[0], args[1:]
        elif "type" in kwargs:
            t = kwargs.pop("type")
        else:
            _err("Need to specify object type. Use e.g. Solver.create(FlagGrid, ...) or Solver.create(type=FlagGrid, ...)")
        kwargs.pop("T", None)
        if isinstance(t, type) and issubclass(t, GridBase):
            if args or kwargs:
                for k in kwargs:
                    _err("Argument '%s' unknown" % k)
            g = t(self)
            g.name = name
            return g
        if isinstance(t, type) and issubclass(t, Shape):
            return t(self, *args, **kwargs)
        _err("Solver.create: unsupported type %r" % (t,))

    def printMemInfo(self):
        b = C.c_size_t()
        check(lib().mfx_scene_memory(self._h, C.byref(b)))
        print("Allocated device memory: %.1f MiB in %d grids" % (b.value / 2**20, len(self._grids)))

    def sync(self):
        check(lib().mfx_scene_sync(self._h))

    def __del__(self):
        try:
            for g in list(self._grids):
                g._release()
            if getattr(self, "_h", None):
                lib().mfx_scene_destroy(self._h)
                self._h = None
        except Exception:
            pass


Solver = FluidSolver


# ---------------------------------------------------------------------------
# grids (grid.h:27-365)
# ---------------------------------------------------------------------------
class GridBase:
    TypeNone, TypeReal, TypeInt, TypeVec3, TypeMAC, TypeLevelset, TypeFlags = 0, 1, 2, 4, 8, 16, 32
    _mfx_type = None

    def __init__(self, parent, show=True):
        if not isinstance(parent, FluidSolver):
            _err("grids must be created with a Solver as parent")
        self.parent = parent
        self.name = ""
        h = C.c_void_p()
        check(lib().mfx_grid_create(parent._h, self._mfx_type, C.byref(h)))
        self._h = h
        parent._grids.append(self)

    def _release(self):
        if getattr(self, "_h", None):
            lib().mfx_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and self.parent._h:
                self.parent._grids.remove(self)
                self._release()
        except Exception:
            pass

    # size / type
    def getParent(self):
        return self.parent

    def getSizeX(self):
        return self.parent.mGridSize[0]

    def getSizeY(self):
        return self.parent.mGridSize[1]

    def getSizeZ(self):
        return self.parent.mGridSize[2]

    def getSize(self):
        return vec3(*self.parent.mGridSize)

    def is3D(self):
        return self.parent.mDim == 3

    def is4D(self):
        return False

    def getSizeT(self):
        return 1

    def getStrideT(self):
        return 0

    def getGridType(self):
        t = C.c_int()
        check(lib().mfx_grid_type(self._h, C.byref(t)))
        return t.value

    def getDataPointer(self):
        return hex(self._h.value)

    @property
    def _n(self):
        gs = self.parent.mGridSize
        return gs[0] * gs[1] * gs[2]

    def _check_same(self, other):
        if not isinstance(other, GridBase) or other.parent.mGridSize != self.parent.mGridSize:
            _err("different grid resolutions")


class _ScalarOps:
    """Operations shared by Real/Int/Vec3 grids (grid.cpp:205-280)."""
    _ncomp = 1
    _np = np.float32

    def _val(self, v):
        if self._ncomp == 3:
            return _abi.floats(*_v3(v))
        return _abi.floats(float(v), 0, 0)

    def clear(self):
        check(lib().mfx_grid_clear(self._h))

    def copyFrom(self, a, copyType=True):
        self._check_same(a)
        check(lib().mfx_grid_copy(self._h, a._h))
        return self

    def setConst(self, s):
        check(lib().mfx_grid_set_const(self._h, self._val(s)))

    def addConst(self, s):
        check(lib().mfx_grid_add_const(self._h, self._val(s)))

    def multConst(self, s):
        check(lib().mfx_grid_mult_const(self._h, self._val(s)))

    def add(self, a):
        global _pending
        self._check_same(a)
        p = _pending
        if p is not None and p[0] == "mult" and p[1] is self and isinstance(a, RealGrid) and a is not self:
            # the held-back mult of this grid, fused with this add
            _pending = None
            check(_lib_now().mfx_grid_mult_add(self._h, p[2]._h, a._h))
            return
        check(lib().mfx_grid_add(self._h, a._h))

    def sub(self, a):
        self._check_same(a)
        check(lib().mfx_grid_sub(self._h, a._h))

    def mult(self, a):
        global _pending
        self._check_same(a)
        if isinstance(self, RealGrid) and isinstance(a, RealGrid):
            _flush()
            _pending = ("mult", self, a)  # held back: fused with a following add
            return
        check(lib().mfx_grid_mult(self._h, a._h))

    def addScaled(self, a, factor):
        self._check_same(a)
        check(lib().mfx_grid_add_scaled(self._h, a._h, self._val(factor)))

    def clamp(self, min, max):
        check(lib().mfx_grid_clamp(self._h, C.c_float(min), C.c_float(max)))

    def _reduce(self, op, bnd=0):
        # the reference returns Real (grid.cpp:315-389): FP64 sums, rounded once
        out = C.c_double()
        check(lib().mfx_grid_reduce(self._h, op, bnd, C.byref(out)))
        return float(np.float32(out.value))

    def getMax(self):
        return self._reduce(_abi.RED_MAX)

    def getMin(self):
        return self._reduce(_abi.RED_MIN)

    def getMaxAbs(self):
        return self._reduce(_abi.RED_MAXABS)

    def getL1(self, bnd=0):
        return self._reduce(_abi.RED_L1, bnd)

    def getL2(self, bnd=0):
        return self._reduce(_abi.RED_L2, bnd)

    def setBound(self, value, boundaryWidth=1):
        check(lib().mfx_grid_set_bound(self._h, self._val(value), int(boundaryWidth)))

    def setBoundNeumann(self, boundaryWidth=1):
        check(lib().mfx_grid_set_bound_neumann(self._h, int(boundaryWidth)))

    # host <-> device (reference host layout, x fastest)
    def toNumpy(self):
        shape = (self._n, 3) if self._ncomp == 3 else (self._n,)
        a = np.empty(shape, self._np)
        check(lib().mfx_grid_download(self._h, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return a

    def fromNumpy(self, a):
        a = np.ascontiguousarray(a, dtype=self._np).reshape(-1)
        if a.size != self._n * self._ncomp:
            _err("array size %d does not match grid size %d" % (a.size, self._n * self._ncomp))
        check(lib().mfx_grid_upload(self._h, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def _dims(self):
        return (self.getSizeX(), self.getSizeY(), self.getSizeZ())

    def save(self, name):
        """Grid<T>::save (grid.cpp:110-130): .uni and .raw in the reference's
        formats (gridio), .npz / .npy as numpy containers (x fastest)."""
        try:
            if name.endswith(".uni"):
                gridio.write_uni(name, self.toNumpy(), self._dims(), self.getGridType())
            elif name.endswith(".raw"):
                gridio.write_raw(name, self.toNumpy(), self._np == np.int32)
            elif name.endswith(".npz"):
                np.savez_compressed(name, arr_0=self.toNumpy())
            elif name.endswith(".npy"):
                np.save(name, self.toNumpy())
            else:
                _err("Grid::save: only .uni, .raw, .npz and .npy are supported by this build")
        except gridio.GridIOError as e:
            _err(str(e))
        return 1

    def load(self, name):
        """Grid<T>::load (grid.cpp:132-152), the formats of save."""
        try:
            if name.endswith(".uni"):
                _, a = gridio.read_uni(name, self._dims(), self.getGridType())
                self.fromNumpy(a)
            elif name.endswith(".raw"):
                self.fromNumpy(gridio.read_raw(name, self._n * self._ncomp, self._np == np.int32))
            elif name.endswith(".npz"):
                with np.load(name, allow_pickle=False) as z:
                    self.fromNumpy(z[z.files[0]])
            elif name.endswith(".npy"):
                self.fromNumpy(np.load(name, allow_pickle=False))
            else:
                _err("Grid::load: only .uni, .raw, .npz and .npy are supported by this build")
        except gridio.GridIOError as e:
            _err(str(e))
        return 1


class RealGrid(_ScalarOps, GridBase):
    _mfx_type = _abi.GRID_REAL


class LevelsetGrid(RealGrid):
    _mfx_type = _abi.GRID_LEVELSET


class IntGrid(_ScalarOps, GridBase):
    _mfx_type = _abi.GRID_INT
    _np = np.int32


class VecGrid(_ScalarOps, GridBase):
    _mfx_type = _abi.GRID_VEC3
    _ncomp = 3


Vec3Grid = VecGrid


class MACGrid(VecGrid):
    _mfx_type = _abi.GRID_MAC

    def setBoundMAC(self, value, boundaryWidth, normalOnly=False):
        check(lib().mfx_grid_set_bound_mac(self._h, _abi.floats(*_v3(value)), int(boundaryWidth), int(bool(normalOnly))))


class FlagGrid(IntGrid):
    _mfx_type = _abi.GRID_FLAGS
    TypeNone, TypeFluid, TypeObstacle, TypeEmpty, TypeInflow, TypeOutflow, TypeOpen, TypeStick, TypeReserved = \
        0, 1, 2, 4, 8, 16, 32, 64, 256

    def __init__(self, parent, dim=3, show=True):
        super().__init__(parent, show)

    def initDomain(self, *args, **kwargs):
        a = _parse([("boundaryWidth", 0), ("wall", "xXyYzZ"), ("open", "      "), ("inflow", "      "),
                    ("outflow", "      "), ("phiWalls", None)], args, kwargs)
        if a["phiWalls"] is not None:
            _err("initDomain: phiWalls is not supported by this build")
        check(lib().mfx_flags_init_domain(self._h, int(a["boundaryWidth"]), a["wall"].encode(), a["open"].encode(),
                                          a["inflow"].encode(), a["outflow"].encode()))

    def fillGrid(self, type=FlagFluid):
        check(lib().mfx_flags_fill_grid(self._h, int(type)))

    def countCells(self, flag, bnd=0, mask=None):
        out = C.c_longlong()
        check(lib().mfx_flags_count_cells(self._h, int(flag), int(bnd), mask._h if mask is not None else None,
                                          C.byref(out)))
        return int(out.value)


# ---------------------------------------------------------------------------
# shapes (shapes.h:25-130)
# ---------------------------------------------------------------------------
class Shape:
    _kind = 0

    def __init__(self, parent):
        self.parent = parent

    def _params(self):
        raise NotImplementedError

    def applyToGrid(self, *args, **kwargs):
        a = _parse([("grid", _REQ), ("value", _REQ), ("respectFlags", None)], args, kwargs)
        g = a["grid"]
        if isinstance(g, VecGrid):
            val = _v3(a["value"])
        else:
            val = (float(a["value"]), 0.0, 0.0)
        pa, pb, r = self._params()
        rf = a["respectFlags"]
        check(lib().mfx_apply_shape(g._h, self._kind, _abi.floats(*pa), _abi.floats(*pb), C.c_float(r),
                                    _abi.floats(*val), rf._h if rf is not None else None))


class Box(Shape):
    _kind = _abi.SHAPE_BOX

    def __init__(self, parent, center=None, p0=None, p1=None, size=None):
        super().__init__(parent)
        if center is not None and size is not None:
            c, s = vec3(center), vec3(size)
            self.p0, self.p1 = c - s, c + s
        elif p0 is not None and p1 is not None:
            self.p0, self.p1 = vec3(p0), vec3(p1)
        else:
            _err("Box: specify either p0,p1 or size,center")

    def _params(self):
        return tuple(self.p0), tuple(self.p1), 0.0

    def getCenter(self):
        return (self.p0 + self.p1) * 0.5

    def getSize(self):
        return self.p1 - self.p0


class Sphere(Shape):
    _kind = _abi.SHAPE_SPHERE

    def __init__(self, parent, center, radius, scale=None):
        super().__init__(parent)
        self.center, self.radius = vec3(center), float(radius)
        self.scale = vec3(scale) if scale is not None else vec3(1, 1, 1)

    def _params(self):
        return tuple(self.center), tuple(self.scale), self.radius

    def getCenter(self):
        return self.center


class Cylinder(Shape):
    _kind = _abi.SHAPE_CYLINDER

    def __init__(self, parent, center, radius, z):
        super().__init__(parent)
        self.center, self.radius, self.z = vec3(center), float(radius), vec3(z)

    def _params(self):
        return tuple(self.center), tuple(self.z), self.radius

    def setRadius(self, r):
        self.radius = float(r)

    def setZ(self, z):
        self.z = vec3(z)

    def getCenter(self):
        return self.center


# ---------------------------------------------------------------------------
# plugins
# ---------------------------------------------------------------------------
def _h(g):
    return g._h if g is not None and g != 0 else None


@_plugin
def advectSemiLagrange(*args, **kwargs):
    """advection.cpp:442-460"""
    a = _parse([("flags", _REQ), ("vel", _REQ), ("grid", _REQ), ("order", 1), ("strength", 1.0),
                ("orderSpace", 1), ("openBounds", False), ("boundaryWidth", -1), ("clampMode", 2),
                ("orderTrace", 1)], args, kwargs)
    if a["boundaryWidth"] != -1 or a["openBounds"]:
        if _debug_level >= 0:
            print("Warning: boundaryWidth and openBounds parameters in AdvectSemiLagrange plugin are deprecated "
                  "(and have no more effect), please remove.")
    if isinstance(a["grid"], MACGrid) and a["orderSpace"] != 1 and _debug_level >= 1:
        # advection.cpp:415, printed after the forward step there
        print("Warning higher order for MAC grids not yet implemented...")
    check(lib().mfx_advect_semi_lagrange(_h(a["flags"]), _h(a["vel"]), _h(a["grid"]), int(a["order"]),
                                         C.c_float(a["strength"]), int(a["orderSpace"]), int(a["clampMode"]),
                                         int(a["orderTrace"])))


@_plugin
def setWallBcs(*args, **kwargs):
    """extforces.cpp:307-315"""
    a = _parse([("flags", _REQ), ("vel", _REQ), ("obvel", None), ("fractions", None), ("phiObs", None),
                ("boundaryWidth", 0)], args, kwargs)
    global _pending
    fl, vel = a["flags"], a["vel"]
    if _h(a["fractions"]) is not None and _h(a["phiObs"]) is not None:
        check(lib().mfx_set_wall_bcs_frac(_h(fl), _h(vel), _h(a["obvel"]), _h(a["fractions"]),
                                          _h(a["phiObs"]), int(a["boundaryWidth"])))
    elif _h(a["obvel"]) is None and isinstance(fl, FlagGrid) and isinstance(vel, MACGrid) and \
            fl.parent is vel.parent:
        p = _pending
        if p is not None and p[0] == "corr" and p[1] is fl and p[2] is vel:
            # the held-back correction of solvePressure, fused with this call
            _pending = None
            check(_lib_now().mfx_correct_velocity_wall_bcs(vel._h, p[3]._h, fl._h))
        else:
            _flush()
            _pending = ("wall", fl, vel)
    else:
        check(lib().mfx_set_wall_bcs(_h(fl), _h(vel), _h(a["obvel"])))


@_plugin
def addBuoyancy(*args, **kwargs):
    """extforces.cpp:83-87"""
    a = _parse([("flags", _REQ), ("density", _REQ), ("vel", _REQ), ("gravity", _REQ), ("coefficient", 1.0),
                ("scale", True)], args, kwargs)
    global _pending
    g = _v3(a["gravity"], "gravity")
    p = _pending
    if p is not None and p[0] == "wall" and p[1] is a["flags"] and p[2] is a["vel"]:
        # the held-back setWallBcs on the same grids, fused with this call
        _pending = None
        try:
            check(_lib_now().mfx_set_wall_bcs_buoyancy(_h(a["flags"]), _h(a["vel"]), _h(a["density"]),
                                                       _abi.floats(*g), C.c_float(a["coefficient"]),
                                                       int(bool(a["scale"]))))
        except RuntimeError:
            # the arguments were refused before any work: setWallBcs alone
            check(_lib_now().mfx_set_wall_bcs(p[1]._h, p[2]._h, None))
            raise
        return
    check(lib().mfx_add_buoyancy(_h(a["flags"]), _h(a["density"]), _h(a["vel"]), _abi.floats(*g),
                                 C.c_float(a["coefficient"]), int(bool(a["scale"]))))


@_plugin
def setOpenBound(*args, **kwargs):
    """extforces.cpp:105-130"""
    a = _parse([("flags", _REQ), ("bWidth", _REQ), ("openBound", ""), ("type", FlagOutflow | FlagEmpty)],
               args, kwargs)
    check(lib().mfx_set_open_bound(_h(a["flags"]), int(a["bWidth"]), a["openBound"].encode(), int(a["type"])))


@_plugin
def resetOutflow(*args, **kwargs):
    """extforces.cpp:133-160"""
    a = _parse([("flags", _REQ), ("phi", None), ("parts", None), ("real", None), ("index", None),
                ("indexSys", None)], args, kwargs)
    if a["parts"] is not None:
        _err("resetOutflow: particle systems are not supported by this build")
    check(lib().mfx_reset_outflow(_h(a["flags"]), _h(a["phi"]), _h(a["real"])))


_PRESSURE_SIG = [("cgAccuracy", 1e-3), ("phi", None), ("perCellCorr", None), ("fractions", None),
                 ("obvel", None), ("gfClamp", 1e-04), ("cgMaxIterFac", 1.5), ("precondition", True),
                 ("preconditioner", PcMIC), ("enforceCompatibility", False), ("useL2Norm", False),
                 ("zeroPressureFixing", False), ("curv", None), ("surfTens", 0.)]


_BND_GRIDS = ("phi", "perCellCorr", "fractions", "obvel", "curv")


def _bnd(a):
    """The optional boundary inputs (pressure.cpp:480-495) as the C struct,
    or None when none of the grids is given (None or 0 both mean absent)."""
    if all(_h(a.get(k)) is None for k in _BND_GRIDS):
        return None
    b = _abi.PressureBnd()
    lib().mfx_pressure_bnd_default(C.byref(b))
    for k in _BND_GRIDS:
        setattr(b, k, _h(a.get(k)))
    b.gfClamp = _f32(a["gfClamp"])
    b.surfTens = _f32(a["surfTens"])
    return b


def _opts(a, solver):
    o = _abi.PressureOpts()
    lib().mfx_pressure_opts_default(C.byref(o))
    o.cgAccuracy = a["cgAccuracy"]
    o.cgMaxIterFac = a["cgMaxIterFac"]
    pc = int(a["preconditioner"])
    if not a["precondition"]:
        pc = PcNone  # backwards compatibility (pressure.cpp:327)
    o.preconditioner = pc
    o.enforceCompatibility = int(bool(a["enforceCompatibility"]))
    o.useL2Norm = int(bool(a["useL2Norm"]))
    o.zeroPressureFixing = int(bool(a["zeroPressureFixing"]))
    o.relTol = float(getattr(solver, "mfxRelTol", 0.0) or 0.0)
    o.precision = int(getattr(solver, "mfxPrecision", 0) or 0)
    o.trueResidualEvery = int(getattr(solver, "mfxTrueResidualEvery", 0) or 0)
    return o


def _report(solver, st, what="solvePressure"):
    solver._solve_stats = {"iterations": st.iterations, "resNorm": st.resNorm, "relResidual": st.relResidual,
                           "fixPidx": st.fixPidx, "maxIter": st.maxIter, "preconditioner": st.preconditionerUsed,
                           "converged": bool(st.converged), "seconds": st.seconds}
    if what == "solvePressure" and _debug_level >= 2:
        print("FluidSolver::solvePressure done. Iterations:%d, residual:%g" % (st.iterations, st.resNorm))
    elif what == "cgSolveWaveEq" and _debug_level >= 1:
        # waves.cpp:137 (debug level 1)
        print("cgSolveWaveEq iterations:%d, res:%g" % (st.iterations, st.resNorm))


@_plugin
def solvePressure(*args, **kwargs):
    """pressure.cpp:480-521"""
    global _pending
    sig = [("vel", _REQ), ("pressure", _REQ), ("flags", _REQ)] + _PRESSURE_SIG + [("retRhs", None)]
    a = _parse(sig, args, kwargs)
    solver = a["vel"].parent
    o = _opts(a, solver)
    st = _abi.SolveStats()
    b = _bnd(a)
    if b is None and _h(a["retRhs"]) is None:
        # the velocity correction is held back (run by the next call, fused
        # with it when that is setWallBcs on the same grids)
        o.deferCorrection = 1
        check(lib().mfx_solve_pressure(_h(a["vel"]), _h(a["pressure"]), _h(a["flags"]), C.byref(o), C.byref(st),
                                       None))
        _pending = ("corr", a["flags"], a["vel"], a["pressure"])
    elif b is None:
        check(lib().mfx_solve_pressure(_h(a["vel"]), _h(a["pressure"]), _h(a["flags"]), C.byref(o), C.byref(st),
                                       _h(a["retRhs"])))
    else:
        check(lib().mfx_solve_pressure_bnd(_h(a["vel"]), _h(a["pressure"]), _h(a["flags"]), C.byref(o),
                                           C.byref(b), C.byref(st), _h(a["retRhs"])))
    _report(solver, st)


def solvePressureBatch(problems, **kwargs):
    """Extension (no reference counterpart): solvePressure for many
    independent (vel, pressure, flags) problems at once, each on its own
    Solver; the keyword arguments are solvePressure's and apply to all.  The
    first problem's solver supplies mfxRelTol / mfxPrecision.  Sets every
    solver's lastSolveStats."""
    problems = list(problems)
    if not problems:
        return
    a = _parse(_PRESSURE_SIG, (), kwargs)
    if _bnd(a) is not None:
        _err("solvePressureBatch: the second-order boundary inputs (%s) are taken by solvePressure only"
             % ", ".join(_BND_GRIDS))
    solver = problems[0][0].parent
    o = _opts(a, solver)
    n = len(problems)
    V = (C.c_void_p * n)(*[_h(v) for v, _, _ in problems])
    Pp = (C.c_void_p * n)(*[_h(p) for _, p, _ in problems])
    F = (C.c_void_p * n)(*[_h(f) for _, _, f in problems])
    st = (_abi.SolveStats * n)()
    check(lib().mfx_solve_pressure_batch(n, V, Pp, F, C.byref(o), st))
    for i, (v, _, _) in enumerate(problems):
        _report(v.parent, st[i])


@_plugin
def computePressureRhs(*args, **kwargs):
    """pressure.cpp:277-299"""
    sig = [("rhs", _REQ), ("vel", _REQ), ("pressure", _REQ), ("flags", _REQ)] + \
        [x for x in _PRESSURE_SIG]
    a = _parse(sig, args, kwargs)
    o = _opts(a, a["vel"].parent)
    b = _bnd(a)
    if b is None:
        check(lib().mfx_compute_pressure_rhs(_h(a["rhs"]), _h(a["vel"]), _h(a["flags"]), C.byref(o)))
    else:
        check(lib().mfx_compute_pressure_rhs_bnd(_h(a["rhs"]), _h(a["vel"]), _h(a["flags"]), C.byref(o),
                                                 C.byref(b)))


@_plugin
def solvePressureSystem(*args, **kwargs):
    """pressure.cpp:312-452"""
    sig = [("rhs", _REQ), ("vel", _REQ), ("pressure", _REQ), ("flags", _REQ)] + \
        [x for x in _PRESSURE_SIG if x[0] != "obvel"]
    a = _parse(sig, args, kwargs)
    solver = a["vel"].parent
    o = _opts(dict(a, obvel=None), solver)
    st = _abi.SolveStats()
    b = _bnd(dict(a, obvel=None))
    if b is None:
        check(lib().mfx_solve_pressure_system(_h(a["rhs"]), _h(a["pressure"]), _h(a["flags"]), C.byref(o),
                                              C.byref(st)))
    else:
        check(lib().mfx_solve_pressure_system_bnd(_h(a["rhs"]), _h(a["pressure"]), _h(a["flags"]), C.byref(o),
                                                  C.byref(b), C.byref(st)))
    _report(solver, st)


@_plugin
def correctVelocity(*args, **kwargs):
    """pressure.cpp:455-476"""
    sig = [("vel", _REQ), ("pressure", _REQ), ("flags", _REQ)] + [x for x in _PRESSURE_SIG if x[0] != "obvel"]
    a = _parse(sig, args, kwargs)
    _opts(dict(a, obvel=None), a["vel"].parent)
    b = _bnd(dict(a, obvel=None))
    if b is None:
        check(lib().mfx_correct_velocity(_h(a["vel"]), _h(a["pressure"]), _h(a["flags"])))
    else:
        check(lib().mfx_correct_velocity_bnd(_h(a["vel"]), _h(a["pressure"]), _h(a["flags"]), C.byref(b)))


@_plugin
def cgSolveDiffusion(*args, **kwargs):
    """conjugategrad.cpp:347-423: implicit diffusion of a Real, Vec3 or MAC
    grid (every component, two in 2D) by unpreconditioned CG on the GPU"""
    a = _parse([("flags", _REQ), ("grid", _REQ), ("alpha", 0.25), ("cgMaxIterFac", 1.0), ("cgAccuracy", 1e-4)],
               args, kwargs)
    st = _abi.SolveStats()
    check(lib().mfx_cg_solve_diffusion(_h(a["flags"]), _h(a["grid"]), C.c_float(_f32(a["alpha"])),
                                       C.c_float(_f32(a["cgMaxIterFac"])), C.c_float(_f32(a["cgAccuracy"])),
                                       C.byref(st)))
    solver = a["flags"].parent
    _report(solver, st, "solveDiffusion")


@_plugin
def cgSolveWE(*args, **kwargs):
    """plugin/waves.cpp:87-141: one implicit wave-equation step by
    unpreconditioned CG on the GPU; afterwards utm1 holds the old ut and ut
    the new heights (= out)"""
    a = _parse([("flags", _REQ), ("ut", _REQ), ("utm1", _REQ), ("out", _REQ), ("crankNic", False),
                ("cSqr", 0.25), ("cgMaxIterFac", 1.5), ("cgAccuracy", 1e-5)], args, kwargs)
    st = _abi.SolveStats()
    check(lib().mfx_cg_solve_we(_h(a["flags"]), _h(a["ut"]), _h(a["utm1"]), _h(a["out"]),
                                int(bool(a["crankNic"])), C.c_float(_f32(a["cSqr"])),
                                C.c_float(_f32(a["cgMaxIterFac"])), C.c_float(_f32(a["cgAccuracy"])), C.byref(st)))
    solver = a["flags"].parent
    _report(solver, st, "cgSolveWaveEq")


@_plugin
def releaseMG(*args, **kwargs):
    """pressure.cpp:252-266"""
    a = _parse([("solver", None)], args, kwargs)
    s = a["solver"]
    check(lib().mfx_release_mg(s._h if s is not None else None))


def _maxdiff(g1, g2):
    # FP64 maximum returned as Real (grid.cpp:400-430)
    out = C.c_double()
    check(lib().mfx_grid_max_diff(g1._h, g2._h, C.byref(out)))
    return float(np.float32(out.value))


@_plugin
def gridMaxDiff(g1, g2):
    """grid.cpp:400-407"""
    return _maxdiff(g1, g2)


@_plugin
def gridMaxDiffInt(g1, g2):
    """grid.cpp:408-414"""
    return _maxdiff(g1, g2)


@_plugin
def gridMaxDiffVec3(g1, g2):
    """grid.cpp:415-430"""
    return _maxdiff(g1, g2)


@_plugin
def copyMacToVec3(source, target):
    check(lib().mfx_copy_mac_to_vec3(source._h, target._h))


@_plugin
def getComponent(source, target, component):
    check(lib().mfx_get_component(source._h, target._h, int(component)))


@_plugin
def setComponent(source, target, component):
    check(lib().mfx_set_component(source._h, target._h, int(component)))


# numpy bridge names of plugin/numpyconvert.cpp:145-183 (x-fastest arrays)
def copyArrayToGridReal(source, target):
    target.fromNumpy(source)


def copyGridToArrayReal(source, target):
    np.copyto(target.reshape(-1), source.toNumpy().reshape(-1))


copyArrayToGridLevelset = copyArrayToGridReal
copyGridToArrayLevelset = copyGridToArrayReal


def copyArrayToGridFlag(source, target):
    target.fromNumpy(source)


def copyGridToArrayFlag(source, target):
    np.copyto(target.reshape(-1), source.toNumpy().reshape(-1))


def copyArrayToGridVec3(source, target):
    target.fromNumpy(source)


def copyGridToArrayVec3(source, target):
    np.copyto(target.reshape(-1), source.toNumpy().reshape(-1))


copyArrayToGridMAC = copyArrayToGridVec3
copyGridToArrayMAC = copyGridToArrayVec3


def mantaMsg(out, level=1):
    if _debug_level >= level:
        print(out)


def setDebugLevel(level=1):
    global _debug_level
    _debug_level = int(level)


def printBuildInfo():
    info = "mantaflow_amd libmfx gfx950 fp1 hip"
    if _debug_level >= 1:
        print("Build info: %s " % info)
    return info


def lastSolveStats(solver):
    """Statistics of the last pressure solve on `solver` (extension)."""
    return solver._solve_stats
