"""CPU oracle for the per-step grid path — TEST INFRASTRUCTURE ONLY.

A ctypes view of liboracle.so / liboracle_f64.so (built from mfx_oracle.c and
mfx_oracle_mg.c by the Makefile in this directory).  Only the tests, the smoke
entry point and the bench's CPU-baseline leg use it, and only as the checker:
the product library never routes through it.

Arrays use the reference's host layout: Real grids float32[n], Vec3/MAC grids
float32[n, 3] (AoS), flag grids int32[n], flattened x-fastest
(idx = i + nx*j + nx*ny*k).  ``f64=True`` selects the double build.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

FLUID, OBSTACLE, EMPTY, INFLOW, OUTFLOW, OPEN, STICK = 1, 2, 4, 8, 16, 32, 64
TYPE_REAL, TYPE_INT, TYPE_VEC3, TYPE_MAC = 1, 2, 4, 8
PC_NONE, PC_MIC, PC_MGDYN, PC_MGSTATIC = 0, 1, 2, 3


class Dims(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("is3d", C.c_int)]


def _opts_struct(real):
    class Opts(C.Structure):
        _fields_ = [("cgAccuracy", real), ("cgMaxIterFac", real), ("preconditioner", C.c_int),
                    ("enforceCompatibility", C.c_int), ("useL2Norm", C.c_int),
                    ("zeroPressureFixing", C.c_int), ("relTol", C.c_double),
                    ("maxIterOverride", C.c_int)]
    return Opts


def _bnd_struct(real):
    class Bnd(C.Structure):
        _fields_ = [("phi", C.c_void_p), ("perCellCorr", C.c_void_p), ("fractions", C.c_void_p),
                    ("obvel", C.c_void_p), ("curv", C.c_void_p), ("gfClamp", real), ("surfTens", real)]
    return Bnd


def _stats_struct(real):
    class Stats(C.Structure):
        _fields_ = [("iterations", C.c_int), ("resNorm", real), ("relResidual", C.c_double),
                    ("fixPidx", C.c_int), ("diverged", C.c_int), ("maxIter", C.c_int)]
    return Stats


class Oracle:
    def __init__(self, f64=False, path=None):
        """``path`` loads another build of the float oracle (the planted-mistake
        builds of the oracle's own tests)."""
        name = "liboracle_f64.so" if f64 else "liboracle.so"
        path = path or os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle library %s is not built (run make in %s)" % (name, _HERE))
        self.lib = C.CDLL(path)
        self.f64 = f64
        self.dtype = np.float64 if f64 else np.float32
        self.creal = C.c_double if f64 else C.c_float
        self.Opts = _opts_struct(self.creal)
        self.Stats = _stats_struct(self.creal)
        self.Bnd = _bnd_struct(self.creal)
        L, R, P = self.lib, self.creal, C.c_void_p
        D, I = Dims, C.c_int
        sig = {
            "og_set_threads": (None, [I]),
            "og_init_domain": (None, [D, P, I, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]),
            "og_fill_grid": (None, [D, P, I]),
            "og_set_open_bound": (None, [D, P, I, C.c_char_p, I]),
            "og_reset_outflow": (None, [D, P, P, P]),
            "og_interpol": (R, [D, P, R, R, R]),
            "og_interpol_component": (R, [D, P, I, R, R, R]),
            "og_advect_semi_lagrange": (I, [D, R, P, P, P, I, I, R, I]),
            "og_advect_semi_lagrange_hi": (I, [D, R, P, P, P, I, I, R, I, I, I]),
            "og_interpol_cubic": (R, [D, P, R, R, R]),
            "og_interpol_cubic_component": (R, [D, P, I, R, R, R]),
            "og_interpol_mac": (None, [D, P, R, R, R, P]),
            "og_sl_scalar": (None, [D, P, P, P, R, I]),
            "og_sl_mac": (None, [D, P, P, P, R]),
            "og_apply_outflow_bc": (None, [D, P, P, P, R]),
            "og_set_wall_bcs": (None, [D, P, P, P]),
            "og_add_buoyancy": (None, [D, R, P, P, P, R, R, R, R, I]),
            "og_make_rhs": (I, [D, P, P, P, C.POINTER(C.c_double)]),
            "og_make_laplace": (None, [D, P, P, P, P, P]),
            "og_apply_matrix": (None, [D, P, P, P, P, P, P, P]),
            "og_find_fix_cell": (I, [D, P]),
            "og_correct_velocity": (None, [D, P, P, P]),
            "og_mic_precondition": (None, [D, P, P, P, P, P, P, P]),
            "og_solve_pressure": (I, [D, P, P, P, P, P, P]),
            "og_solve_pressure_system": (I, [D, P, P, P, P, P]),
            "og_pressure_opts_default": (None, [P]),
            "og_mg_create": (P, [D]),
            "og_mg_destroy": (None, [P]),
            "og_mg_set_a": (None, [P, P, P, P, P]),
            "og_mg_set_coarsest_accuracy": (None, [P, R]),
            "og_mg_vcycle": (R, [P, P, P]),
            "og_mg_num_levels": (I, [P]),
            "og_mg_level_size": (None, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
            "og_mg_level_types": (I, [P, I, P]),
            "og_mg_level_matrix": (I, [P, I, P]),
            "og_apply_shape": (None, [D, I, P, P, R, P, I, P, P]),
            "og_diffusion_matrix": (None, [D, P, R, P, P, P, P]),
            "og_cg_solve_diffusion": (None, [D, P, P, I, R, R, R, P]),
            "og_make_rhs_we": (None, [D, P, P, P, P, R, I]),
            "og_pressure_bnd_default": (None, [P]),
            "og_ghost_fluid_factor": (R, [R, R, R]),
            "og_make_rhs_bnd": (I, [D, P, P, P, P, C.POINTER(C.c_double)]),
            "og_make_laplace_bnd": (None, [D, P, P, P, P, P, P]),
            "og_correct_velocity_bnd": (None, [D, P, P, P, P]),
            "og_solve_pressure_system_bnd": (I, [D, P, P, P, P, P, P]),
            "og_solve_pressure_bnd": (I, [D, P, P, P, P, P, P, P]),
            "og_set_wall_bcs_frac": (None, [D, P, P, P]),
            "og_cg_solve_we": (None, [D, P, P, P, P, I, R, R, R, R, P]),
        }
        for k, (res, args) in sig.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args

    # -- helpers --------------------------------------------------------
    @staticmethod
    def dims(shape):
        nx, ny, nz = shape
        return Dims(nx, ny, nz, 1 if nz > 1 else 0)

    def _arr(self, a, dtype=None, ncomp=1):
        dt = dtype or self.dtype
        if a is None:
            return None
        if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
            raise TypeError("oracle expects contiguous %s arrays" % np.dtype(dt).name)
        return a.ctypes.data_as(C.c_void_p)

    def set_threads(self, n):
        self.lib.og_set_threads(int(n))

    # -- flags ------------------------------------------------------------
    def init_domain(self, shape, boundaryWidth=0, wall="xXyYzZ", open="      ",
                    inflow="      ", outflow="      "):
        n = int(np.prod(shape))
        f = np.zeros(n, np.int32)
        self.lib.og_init_domain(self.dims(shape), self._arr(f, np.int32), boundaryWidth,
                                wall.encode(), open.encode(), inflow.encode(), outflow.encode())
        return f

    def fill_grid(self, shape, flags, type=FLUID):
        self.lib.og_fill_grid(self.dims(shape), self._arr(flags, np.int32), type)

    def set_open_bound(self, shape, flags, bWidth, openBound="", type=OUTFLOW | EMPTY):
        self.lib.og_set_open_bound(self.dims(shape), self._arr(flags, np.int32), bWidth,
                                   openBound.encode(), type)

    def reset_outflow(self, shape, flags, phi=None, real=None):
        self.lib.og_reset_outflow(self.dims(shape), self._arr(flags, np.int32), self._arr(phi),
                                  self._arr(real))

    # -- advection --------------------------------------------------------
    def interpol(self, shape, data, pos):
        return self.lib.og_interpol(self.dims(shape), self._arr(data), *[self.creal(p) for p in pos])

    def advect(self, shape, dt, flags, vel, grid, gridType, order=1, strength=1.0, clampMode=2):
        rc = self.lib.og_advect_semi_lagrange(self.dims(shape), dt, self._arr(flags, np.int32),
                                              self._arr(vel), self._arr(grid), gridType, order,
                                              strength, clampMode)
        if rc != 0:
            raise RuntimeError("AdvectSemiLagrange: Only order 1 (regular SL) and 2 (MacCormack) supported")

    def advect_hi(self, shape, dt, flags, vel, grid, gridType, order=1, strength=1.0, clampMode=2,
                  orderSpace=1, orderTrace=1):
        rc = self.lib.og_advect_semi_lagrange_hi(self.dims(shape), dt, self._arr(flags, np.int32),
                                                 self._arr(vel), self._arr(grid), gridType, order,
                                                 strength, clampMode, orderSpace, orderTrace)
        if rc != 0:
            raise RuntimeError("AdvectSemiLagrange: unsupported order (%d)" % rc)

    def interpol_cubic(self, shape, data, pos):
        return self.lib.og_interpol_cubic(self.dims(shape), self._arr(data), *[self.creal(p) for p in pos])

    def interpol_cubic_component(self, shape, vec, c, pos):
        return self.lib.og_interpol_cubic_component(self.dims(shape), self._arr(vec), c,
                                                    *[self.creal(p) for p in pos])

    def interpol_mac(self, shape, vec, pos):
        out = np.zeros(3, self.dtype)
        self.lib.og_interpol_mac(self.dims(shape), self._arr(vec), *[self.creal(p) for p in pos],
                                 self._arr(out))
        return out

    def apply_outflow_bc(self, shape, flags, vel, velPrev, dt):
        self.lib.og_apply_outflow_bc(self.dims(shape), self._arr(flags, np.int32), self._arr(vel),
                                     self._arr(velPrev), dt)

    # -- forces -----------------------------------------------------------
    def set_wall_bcs(self, shape, flags, vel, obvel=None):
        self.lib.og_set_wall_bcs(self.dims(shape), self._arr(flags, np.int32), self._arr(vel),
                                 self._arr(obvel))

    def add_buoyancy(self, shape, dt, flags, density, vel, gravity, coefficient=1.0, scale=True):
        self.lib.og_add_buoyancy(self.dims(shape), dt, self._arr(flags, np.int32),
                                 self._arr(density), self._arr(vel), gravity[0], gravity[1],
                                 gravity[2], coefficient, 1 if scale else 0)

    # -- pressure -----------------------------------------------------------
    def make_rhs(self, shape, flags, vel):
        n = int(np.prod(shape))
        rhs = np.zeros(n, self.dtype)
        s = C.c_double(0)
        cnt = self.lib.og_make_rhs(self.dims(shape), self._arr(flags, np.int32), self._arr(rhs),
                                   self._arr(vel), C.byref(s))
        return rhs, cnt, s.value

    def make_laplace(self, shape, flags):
        n = int(np.prod(shape))
        A = [np.zeros(n, self.dtype) for _ in range(4)]
        self.lib.og_make_laplace(self.dims(shape), self._arr(flags, np.int32),
                                 *[self._arr(a) for a in A])
        return A

    def apply_matrix(self, shape, flags, src, A):
        dst = np.zeros_like(src)
        self.lib.og_apply_matrix(self.dims(shape), self._arr(flags, np.int32), self._arr(dst),
                                 self._arr(src), *[self._arr(a) for a in A])
        return dst

    def mic_precondition(self, shape, flags, A, r):
        z = np.zeros_like(r)
        self.lib.og_mic_precondition(self.dims(shape), self._arr(flags, np.int32),
                                     *[self._arr(a) for a in A], self._arr(r), self._arr(z))
        return z

    def find_fix_cell(self, shape, flags):
        return self.lib.og_find_fix_cell(self.dims(shape), self._arr(flags, np.int32))

    def correct_velocity(self, shape, flags, vel, pressure):
        self.lib.og_correct_velocity(self.dims(shape), self._arr(flags, np.int32), self._arr(vel),
                                     self._arr(pressure))

    def opts(self, cgAccuracy=1e-3, cgMaxIterFac=1.5, preconditioner=PC_MIC,
             enforceCompatibility=False, useL2Norm=False, zeroPressureFixing=False,
             relTol=0.0, maxIterOverride=0):
        return self.Opts(cgAccuracy, cgMaxIterFac, preconditioner, int(enforceCompatibility),
                         int(useL2Norm), int(zeroPressureFixing), relTol, maxIterOverride)

    def solve_pressure(self, shape, vel, pressure, flags, **kw):
        o = self.opts(**kw)
        st = self.Stats()
        n = int(np.prod(shape))
        rhs = np.zeros(n, self.dtype)
        rc = self.lib.og_solve_pressure(self.dims(shape), self._arr(vel), self._arr(pressure),
                                        self._arr(flags, np.int32), C.byref(o), C.byref(st),
                                        self._arr(rhs))
        if rc != 0:
            raise RuntimeError("GridCg::iterate: The CG solver diverged, residual norm > 1e30, stopping.")
        return {"iterations": st.iterations, "resNorm": st.resNorm, "relResidual": st.relResidual,
                "fixPidx": st.fixPidx, "maxIter": st.maxIter, "rhs": rhs}

    # -- second-order obstacle / free-surface boundaries ------------------
    def bnd(self, phi=None, perCellCorr=None, fractions=None, obvel=None, curv=None, gfClamp=1e-4,
            surfTens=0.0):
        """The optional solvePressure inputs (pressure.cpp:480-495); the arrays
        are kept referenced by the returned struct's ``_keep``."""
        b = self.Bnd()
        keep = []
        for name, a, nc in (("phi", phi, 1), ("perCellCorr", perCellCorr, 1), ("fractions", fractions, 3),
                            ("obvel", obvel, 3), ("curv", curv, 1)):
            if a is not None:
                ptr = self._arr(a, ncomp=nc)
                keep.append(a)
                setattr(b, name, ptr)
        b.gfClamp = gfClamp
        b.surfTens = surfTens
        b._keep = keep
        return b

    def ghost_fluid_factor(self, phi_in, phi_out, gfClamp=1e-4):
        return self.lib.og_ghost_fluid_factor(phi_in, phi_out, gfClamp)

    def make_rhs_bnd(self, shape, flags, vel, b):
        n = int(np.prod(shape))
        rhs = np.zeros(n, self.dtype)
        s = C.c_double(0)
        cnt = self.lib.og_make_rhs_bnd(self.dims(shape), self._arr(flags, np.int32), self._arr(rhs),
                                       self._arr(vel), C.byref(b), C.byref(s))
        return rhs, cnt, s.value

    def make_laplace_bnd(self, shape, flags, b):
        n = int(np.prod(shape))
        A = [np.zeros(n, self.dtype) for _ in range(4)]
        self.lib.og_make_laplace_bnd(self.dims(shape), self._arr(flags, np.int32), *[self._arr(a) for a in A],
                                     C.byref(b))
        return A

    def correct_velocity_bnd(self, shape, flags, vel, pressure, b):
        self.lib.og_correct_velocity_bnd(self.dims(shape), self._arr(flags, np.int32), self._arr(vel),
                                         self._arr(pressure), C.byref(b))

    def solve_pressure_bnd(self, shape, vel, pressure, flags, b, **kw):
        o = self.opts(**kw)
        st = self.Stats()
        n = int(np.prod(shape))
        rhs = np.zeros(n, self.dtype)
        rc = self.lib.og_solve_pressure_bnd(self.dims(shape), self._arr(vel), self._arr(pressure),
                                            self._arr(flags, np.int32), C.byref(o), C.byref(b), C.byref(st),
                                            self._arr(rhs))
        if rc != 0:
            raise RuntimeError("GridCg::iterate: The CG solver diverged, residual norm > 1e30, stopping.")
        return {"iterations": st.iterations, "resNorm": st.resNorm, "relResidual": st.relResidual,
                "fixPidx": st.fixPidx, "maxIter": st.maxIter, "rhs": rhs}

    def set_wall_bcs_frac(self, shape, flags, vel, phiObs):
        self.lib.og_set_wall_bcs_frac(self.dims(shape), self._arr(flags, np.int32), self._arr(vel),
                                      self._arr(phiObs))

    def solve_pressure_system(self, shape, rhs, pressure, flags, **kw):
        o = self.opts(**kw)
        st = self.Stats()
        rc = self.lib.og_solve_pressure_system(self.dims(shape), self._arr(rhs), self._arr(pressure),
                                               self._arr(flags, np.int32), C.byref(o), C.byref(st))
        return rc, {"iterations": st.iterations, "resNorm": st.resNorm,
                    "relResidual": st.relResidual, "fixPidx": st.fixPidx, "maxIter": st.maxIter}

    # -- other GridCg clients ------------------------------------------------
    def _stats(self, st):
        return {"iterations": st.iterations, "resNorm": st.resNorm, "relResidual": st.relResidual,
                "maxIter": st.maxIter, "diverged": bool(st.diverged)}

    def diffusion_matrix(self, shape, flags, alpha):
        n = int(np.prod(shape))
        A = [np.zeros(n, self.dtype) for _ in range(4)]
        self.lib.og_diffusion_matrix(self.dims(shape), self._arr(flags, np.int32), alpha,
                                     *[self._arr(a) for a in A])
        return A

    def cg_solve_diffusion(self, shape, flags, grid, alpha=0.25, cgMaxIterFac=1.0, cgAccuracy=1e-4):
        """grid: float[n] (Real) or float[n, 3] (Vec3/MAC), updated in place"""
        st = self.Stats()
        ncomp = 1 if grid.ndim == 1 else 3
        self.lib.og_cg_solve_diffusion(self.dims(shape), self._arr(flags, np.int32), self._arr(grid), ncomp,
                                       alpha, cgMaxIterFac, cgAccuracy, C.byref(st))
        return self._stats(st)

    def make_rhs_we(self, shape, flags, ut, utm1, s, crankNic=False):
        rhs = np.zeros_like(ut)
        self.lib.og_make_rhs_we(self.dims(shape), self._arr(flags, np.int32), self._arr(rhs), self._arr(ut),
                                self._arr(utm1), s, int(bool(crankNic)))
        return rhs

    def cg_solve_we(self, shape, flags, ut, utm1, out, dt=1.0, crankNic=False, cSqr=0.25, cgMaxIterFac=1.5,
                    cgAccuracy=1e-5):
        st = self.Stats()
        self.lib.og_cg_solve_we(self.dims(shape), self._arr(flags, np.int32), self._arr(ut), self._arr(utm1),
                                self._arr(out), int(bool(crankNic)), cSqr, cgMaxIterFac, cgAccuracy, dt,
                                C.byref(st))
        return self._stats(st)

    # -- multigrid -----------------------------------------------------------
    def mg_create(self, shape):
        return self.lib.og_mg_create(self.dims(shape))

    def mg_levels(self, mg):
        """[(size, types int8[n], stencil storage [n, S])] for every level."""
        out = []
        for l in range(self.lib.og_mg_num_levels(mg)):
            a, b, c = C.c_int(), C.c_int(), C.c_int()
            self.lib.og_mg_level_size(mg, l, C.byref(a), C.byref(b), C.byref(c))
            n = a.value * b.value * c.value
            t = np.zeros(n, np.int8)
            self.lib.og_mg_level_types(mg, l, t.ctypes.data_as(C.c_void_p))
            m = np.zeros(n * 14, self.dtype)
            ss = self.lib.og_mg_level_matrix(mg, l, m.ctypes.data_as(C.c_void_p))
            out.append(((a.value, b.value, c.value), t, m[:n * ss].reshape(n, ss)))
        return out

    def mg_set_a(self, mg, A):
        self.lib.og_mg_set_a(mg, *[self._arr(a) for a in A])

    def mg_vcycle(self, mg, rhs, acc=None):
        if acc is not None:
            self.lib.og_mg_set_coarsest_accuracy(mg, acc)
        z = np.zeros_like(rhs)
        self.lib.og_mg_vcycle(mg, self._arr(rhs), self._arr(z))
        return z

    def mg_destroy(self, mg):
        self.lib.og_mg_destroy(mg)

    # -- shapes ---------------------------------------------------------------
    def apply_shape(self, shape, kind, a, b, r, grid, gridType, value, respectFlags=None):
        av = np.asarray(a, self.dtype)
        bv = np.asarray(b, self.dtype)
        vv = np.asarray(value if np.ndim(value) else [value, 0, 0], self.dtype)
        gdt = np.int32 if gridType == TYPE_INT else self.dtype
        self.lib.og_apply_shape(self.dims(shape), kind, self._arr(av), self._arr(bv), r,
                                self._arr(grid, gdt), gridType, self._arr(vv),
                                self._arr(respectFlags, np.int32))
