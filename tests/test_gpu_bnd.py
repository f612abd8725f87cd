"""Second-order obstacle and free-surface boundaries on the GPU against the
oracle: MakeRhs with fractions, obstacle velocity, surface tension and
per-cell correction, correctVelocity's ghost-fluid terms and setWallBcs with
phiObs bit for bit; the pressure solve on the fractions / ghost-fluid matrix
(plain CG against the oracle's CG on the same matrix: iterations within one,
pressure within the CG tolerance; multigrid-preconditioned CG against the
stopping rule measured with the oracle's matrix); the whole solvePressure
through the Python face.  2D and 3D, liquid with empty cells above, obstacle
cells, outflow cells."""
import ctypes as C

import numpy as np
import pytest

from mantaflow_amd import _abi
from mantaflow_amd._abi import lib
from oracle import EMPTY, FLUID, OBSTACLE, OUTFLOW, PC_NONE
from tests.helpers import GpuScene, assert_bitexact

pytestmark = pytest.mark.gpu

SHAPES = [(26, 22, 1), (17, 15, 13)]


def _case(shape, seed, outflow=False):
    """Walls, a few obstacle cells, liquid below a wavy surface (empty
    above), the level set of that surface, fractions that close obstacle
    faces and open the rest partly, curvature, obstacle velocity."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    n = nx * ny * nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    surf = 0.6 * ny + 1.5 * np.sin(0.7 * i) + (np.cos(0.5 * k) if nz > 1 else 0)
    phi = (j + 0.5 - surf).astype(np.float32)          # < 0 in the liquid
    f = np.where(phi < 0, FLUID, EMPTY).astype(np.int32)
    wall = (i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1)
    if nz > 1:
        wall |= (k == 0) | (k == nz - 1)
    f[wall] = OBSTACLE
    for _ in range(3):
        x, y = rng.integers(2, nx - 3), rng.integers(2, ny // 2)
        z = rng.integers(2, nz - 2) if nz > 1 else 0
        f[z, y, x] = OBSTACLE
    if outflow:
        f[..., ny - 2, 1:-1][f[..., ny - 2, 1:-1] == EMPTY] = EMPTY | OUTFLOW
    flags = f.reshape(-1)
    ob = (flags & OBSTACLE) != 0
    fr = rng.uniform(0.3, 1.0, (n, 3)).astype(np.float32)
    F = fr.reshape(nz, ny, nx, 3)
    O = ob.reshape(nz, ny, nx)
    F[..., 0][O | np.roll(O, 1, axis=2)] = 0
    F[..., 1][O | np.roll(O, 1, axis=1)] = 0
    if nz > 1:
        F[..., 2][O | np.roll(O, 1, axis=0)] = 0
    else:
        F[..., 2] = 0
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    obvel = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    if nz == 1:
        vel[:, 2] = 0
        obvel[:, 2] = 0
    curv = rng.uniform(-0.2, 0.2, n).astype(np.float32)
    pcc = rng.uniform(-0.05, 0.05, n).astype(np.float32)
    return flags, phi.reshape(-1), np.ascontiguousarray(fr), vel, obvel, curv, pcc


KINDS = {
    "fractions": dict(fractions=True),
    "fractions_obvel": dict(fractions=True, obvel=True),
    "ghost": dict(phi=True),
    "ghost_surftens": dict(phi=True, curv=True),
    "pcc": dict(pcc=True),
    "all": dict(fractions=True, obvel=True, phi=True, curv=True, pcc=True),
}


def _inputs(sc, o, kind, phi, fr, obvel, curv, pcc):
    """(oracle bnd struct, C ABI struct, grids kept alive)"""
    k = KINDS[kind]
    g = {}
    if k.get("phi"):
        g["phi"] = sc.real(phi)
    if k.get("fractions"):
        g["fractions"] = sc.mac(fr)
    if k.get("obvel"):
        g["obvel"] = sc.mac(obvel)
    if k.get("curv"):
        g["curv"] = sc.real(curv)
    if k.get("pcc"):
        g["perCellCorr"] = sc.real(pcc)
    b = _abi.PressureBnd()
    lib().mfx_pressure_bnd_default(C.byref(b))
    for name, grid in g.items():
        setattr(b, name, grid._h)
    b.surfTens = 0.3
    ob = o.bnd(phi=phi if "phi" in g else None, fractions=fr if "fractions" in g else None,
               obvel=obvel if "obvel" in g else None, curv=curv if "curv" in g else None,
               perCellCorr=pcc if "perCellCorr" in g else None, surfTens=0.3)
    return ob, b, g


def _opts(**kw):
    o = _abi.PressureOpts()
    lib().mfx_pressure_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("compat", [0, 1])
def test_rhs_bnd_bitexact(manta, oracle, shape, kind, compat):
    flags, phi, fr, vel, obvel, curv, pcc = _case(shape, 11)
    sc = GpuScene(manta, shape)
    f, v, r = sc.flags(flags), sc.mac(vel), sc.real()
    ob, b, keep = _inputs(sc, oracle, kind, phi, fr, obvel, curv, pcc)
    assert lib().mfx_compute_pressure_rhs_bnd(r._h, v._h, f._h, C.byref(_opts(enforceCompatibility=compat)),
                                              C.byref(b)) == 0, lib().mfx_last_error()
    ref, cnt, s = oracle.make_rhs_bnd(shape, flags, vel, ob)
    if not compat:
        assert_bitexact(r.toNumpy(), ref, "rhs %s" % kind)
        return
    # enforceCompatibility adds (Real)(-sum / cnt) to every cell; sum is a
    # double sum of the cells' values whose order differs between the GPU's
    # block partials and the oracle's loop (the reference's own threaded
    # reduction has no fixed order either), so the shift may differ by one
    # float ulp, and each cell by one ulp of its value
    corr = np.float32(-s / np.float32(cnt))
    ref = (ref + corr).astype(np.float32)
    got = r.toNumpy()
    tol = 2 * np.spacing(np.abs(ref).astype(np.float32) + np.abs(corr))
    assert np.all(np.abs(got - ref) <= tol), np.abs(got - ref).max()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["ghost", "ghost_surftens", "all"])
@pytest.mark.parametrize("outflow", [False, True])
def test_correct_velocity_bnd_bitexact(manta, oracle, shape, kind, outflow):
    flags, phi, fr, vel, obvel, curv, pcc = _case(shape, 12, outflow)
    n = int(np.prod(shape))
    p = np.random.default_rng(3).uniform(-1, 1, n).astype(np.float32)
    # a few fluid cells next to the surface with a level set that clamps
    phi = phi.copy()
    phi[np.nonzero(flags & FLUID)[0][::17]] = 0.3
    sc = GpuScene(manta, shape)
    f, v, pr = sc.flags(flags), sc.mac(vel), sc.real(p)
    ob, b, keep = _inputs(sc, oracle, kind, phi, fr, obvel, curv, pcc)
    assert lib().mfx_correct_velocity_bnd(v._h, pr._h, f._h, C.byref(b)) == 0, lib().mfx_last_error()
    ref = vel.copy()
    oracle.correct_velocity_bnd(shape, flags, ref, p, ob)
    assert_bitexact(v.toNumpy(), ref, "correctVelocity %s" % kind)


@pytest.mark.parametrize("shape", [(26, 22, 1), (17, 15, 13), (70, 9, 5)])
@pytest.mark.parametrize("sdf", ["plane", "sphere"])
def test_set_wall_bcs_frac_bitexact(manta, oracle, shape, sdf):
    nx, ny, nz = shape
    n = nx * ny * nz
    k, j, i = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    if sdf == "plane":
        phio = 0.6 * (i - nx / 3) + 0.8 * (j - ny / 2) + (0.3 * (k - nz / 2) if nz > 1 else 0)
    else:
        phio = np.sqrt((i - nx / 2) ** 2 + (j - ny / 2) ** 2 + ((k - nz / 2) ** 2 if nz > 1 else 0)) - min(nx, ny) / 4
    phio = phio.reshape(-1).astype(np.float32)
    flags = np.where(phio < 0, OBSTACLE, FLUID).astype(np.int32)
    flags[::13] = EMPTY
    vel = np.random.default_rng(4).uniform(-1, 1, (n, 3)).astype(np.float32)
    fr = np.ones((n, 3), np.float32)
    sc = GpuScene(manta, shape)
    f, v, frg, po = sc.flags(flags), sc.mac(vel), sc.mac(fr), sc.real(phio)
    manta.setWallBcs(flags=f, vel=v, fractions=frg, phiObs=po)
    ref = vel.copy()
    oracle.set_wall_bcs_frac(shape, flags, ref, phio)
    assert_bitexact(v.toNumpy(), ref, "setWallBcs frac %s" % sdf)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["fractions", "ghost", "all"])
def test_solve_bnd_plain_cg(manta, oracle, shape, kind):
    """PcNone on the float matrix: the oracle's CG runs the same operations,
    so iterations agree within one and pressures within the tolerance."""
    flags, phi, fr, vel, obvel, curv, pcc = _case(shape, 13)
    n = int(np.prod(shape))
    sc = GpuScene(manta, shape)
    f, v, r, pr = sc.flags(flags), sc.mac(vel), sc.real(), sc.real()
    ob, b, keep = _inputs(sc, oracle, kind, phi, fr, obvel, curv, pcc)
    o = _opts(preconditioner=0, cgAccuracy=1e-5, cgMaxIterFac=20.0)
    assert lib().mfx_compute_pressure_rhs_bnd(r._h, v._h, f._h, C.byref(o), C.byref(b)) == 0
    st = _abi.SolveStats()
    assert lib().mfx_solve_pressure_system_bnd(r._h, pr._h, f._h, C.byref(o), C.byref(b), C.byref(st)) == 0, \
        lib().mfx_last_error()
    rhs, _, _ = oracle.make_rhs_bnd(shape, flags, vel, ob)
    ref_p = np.zeros(n, np.float32)
    ref = oracle.solve_pressure_bnd(shape, vel.copy(), ref_p, flags, ob, cgAccuracy=1e-5, preconditioner=PC_NONE,
                                    cgMaxIterFac=20.0)
    assert st.converged and abs(st.iterations - ref["iterations"]) <= 1, (st.iterations, ref["iterations"])
    gp = pr.toNumpy()
    scale = max(float(np.abs(ref_p).max()), 1e-6)
    assert np.abs(gp - ref_p).max() <= 1e-3 * scale
    # the GPU pressure satisfies the stopping rule with the oracle's matrix
    A = oracle.make_laplace_bnd(shape, flags, ob)
    res = rhs - oracle.apply_matrix(shape, flags, gp, A)
    assert np.abs(res).max() < 2e-5


@pytest.mark.parametrize("shape", [(40, 36, 34), (48, 44, 1)])
@pytest.mark.parametrize("pc", ["mgstatic", "mic"])
def test_solve_bnd_multigrid(manta, oracle, shape, pc):
    """Multigrid preconditioner built from the flags' plain Laplacian, CG on
    the fractions / ghost-fluid matrix: converges, and the pressure meets
    the stopping rule measured with the oracle's matrix; the corrected
    velocity matches the oracle's correction of that pressure."""
    flags, phi, fr, vel, obvel, curv, pcc = _case(shape, 14)
    sc = GpuScene(manta, shape)
    f, v, pr = sc.flags(flags), sc.mac(vel), sc.real()
    ob, b, keep = _inputs(sc, oracle, "all", phi, fr, obvel, curv, pcc)
    grids = {k: g for k, g in keep.items()}
    pcv = manta.PcMGStatic if pc == "mgstatic" else manta.PcMIC
    manta.solvePressure(flags=f, vel=v, pressure=pr, cgAccuracy=1e-4, preconditioner=pcv, cgMaxIterFac=20,
                        phi=grids["phi"], fractions=grids["fractions"], obvel=grids["obvel"], curv=grids["curv"],
                        perCellCorr=grids["perCellCorr"], surfTens=0.3)
    gst = sc.s._solve_stats
    print("%s %s: %d iterations" % (pc, shape, gst["iterations"]))
    assert gst["converged"], gst
    gp = pr.toNumpy()
    rhs, _, _ = oracle.make_rhs_bnd(shape, flags, vel, ob)
    A = oracle.make_laplace_bnd(shape, flags, ob)
    # the stopping rule reads the recursive residual (|r|_inf < 1e-4); the
    # true residual in float may sit slightly above it
    assert np.abs(rhs - oracle.apply_matrix(shape, flags, gp, A)).max() < 2e-4
    ref = vel.copy()
    oracle.correct_velocity_bnd(shape, flags, ref, gp, ob)
    assert_bitexact(v.toNumpy(), ref, "velocity after the solve")


def test_batch_rejects_boundary_inputs(manta):
    sc = GpuScene(manta, (16, 16, 1))
    f = sc.s.create(manta.FlagGrid)
    f.initDomain()
    f.fillGrid()
    v, pr, phi = sc.mac(), sc.real(), sc.real()
    with pytest.raises(RuntimeError, match="solvePressure only"):
        manta.solvePressureBatch([(v, pr, f)], phi=phi)
