"""GPU parity for solver options and grid operations beyond the default
path, through libmfx's C ABI (via the manta module), against the oracle or
numpy restatements of the reference's kernels:

- zero-pressure fixing, explicit and engaged by cgAccuracy < 1e-7, with the
  preferred top-centre cells blocked (pressure.cpp:226-245,349-388);
- enforceCompatibility (pressure.cpp:297-298) and retRhs (:516-518);
- useL2Norm, the sum-of-squares stopping rule (conjugategrad.cpp:267-271);
- PcMGDynamic (pressure.cpp:418-433) and the PcMIC stand-in after the flags
  change, against a fresh solver; PcMGStatic's kept hierarchy and releaseMG;
- grid reductions, setBound / setBoundNeumann / setBoundMAC, countCells,
  gridMaxDiff*, getComponent / setComponent / copyMacToVec3
  (grid.cpp:315-430,585-648,680-685)."""
import numpy as np
import pytest

from oracle import EMPTY, FLUID, OBSTACLE, PC_MGSTATIC, PC_NONE
from tests.helpers import GpuScene, assert_bitexact, component_demean, make_flags, smooth_vel

pytestmark = pytest.mark.gpu


def _closed_case(oracle, rng, shape, obstacles=3):
    """initDomain walls (no empty cells), fluid interior with obstacles, a
    velocity with wall BCs applied."""
    flags = make_flags(oracle, shape, bw=0, obstacles=obstacles, rng=rng)
    vel = smooth_vel(rng, shape, 2, 1.0)
    oracle.set_wall_bcs(shape, flags, vel)
    return flags, vel


def _pressure_scale(p):
    return max(float(np.abs(p).max()), 1e-6)


def _same_pressure(gp, rp, flags, shape, tol):
    """pressures equal modulo one constant per fluid region (a region without
    a pinned cell is singular, so CG may leave a different constant there)"""
    a, b = component_demean(gp, flags, shape), component_demean(rp, flags, shape)
    return float(np.abs(a - b).max()) <= tol * _pressure_scale(rp)


@pytest.mark.parametrize("shape", [(40, 36, 1), (24, 22, 20)])
@pytest.mark.parametrize("block", [0, 1, 3])
def test_zero_pressure_fixing(manta, oracle, rng, shape, block):
    """The pinned cell equals the reference's choice (top centre, one or two
    cells below, else the first fluid cell in ijk order inside the boundary),
    its pressure is 0, and the solve matches the oracle's PcNone solve of the
    same pinned system; retRhs returns the pinned right-hand side."""
    nx, ny, nz = shape
    flags, vel = _closed_case(oracle, rng, shape)
    g = flags.reshape(nz, ny, nx)
    tcx, tcz = nx // 2, nz // 2 if nz > 1 else 0
    for o in range(block):  # obstacles on the preferred positions
        g[tcz, ny - 1 - o, tcx] = OBSTACLE
    fix = oracle.find_fix_cell(shape, flags)
    assert fix >= 0
    n = nx * ny * nz
    sc = GpuScene(manta, shape)
    f, v, pr, rr = sc.flags(flags), sc.mac(vel), sc.real(), sc.real()
    manta.solvePressure(flags=f, vel=v, pressure=pr, cgAccuracy=1e-5, preconditioner=manta.PcNone,
                        cgMaxIterFac=30, zeroPressureFixing=True, retRhs=rr)
    st = sc.s._solve_stats
    assert st["fixPidx"] == fix and st["converged"]
    gp = pr.toNumpy()
    assert gp[fix] == 0.0
    ref_v, ref_p = vel.copy(), np.zeros(n, np.float32)
    ost = oracle.solve_pressure(shape, ref_v, ref_p, flags, cgAccuracy=1e-5, preconditioner=PC_NONE,
                                cgMaxIterFac=30, zeroPressureFixing=True)
    assert ost["fixPidx"] == fix
    assert abs(st["iterations"] - ost["iterations"]) <= 1
    assert _same_pressure(gp, ref_p, flags, shape, 1e-4)
    assert np.abs(v.toNumpy() - ref_v).max() <= 1e-4
    assert_bitexact(rr.toNumpy(), ost["rhs"], "retRhs (pinned)")


@pytest.mark.parametrize("shape", [(40, 36, 1), (24, 22, 20)])
def test_zero_pressure_fixing_engaged_by_accuracy(manta, oracle, rng, shape):
    """cgAccuracy < 1e-7 pins a cell without zeroPressureFixing; with empty
    cells in the domain nothing is pinned."""
    flags, vel = _closed_case(oracle, rng, shape)
    fix = oracle.find_fix_cell(shape, flags)
    sc = GpuScene(manta, shape)
    f, v, pr = sc.flags(flags), sc.mac(vel), sc.real()
    manta.solvePressure(flags=f, vel=v, pressure=pr, cgAccuracy=5e-8, preconditioner=manta.PcNone,
                        cgMaxIterFac=2)
    st = sc.s._solve_stats
    assert st["fixPidx"] == fix >= 0
    assert pr.toNumpy()[fix] == 0.0
    nx, ny, nz = shape
    flags2 = flags.copy()
    flags2.reshape(nz, ny, nx)[nz // 2, ny // 2, 2] = EMPTY
    assert oracle.find_fix_cell(shape, flags2) == -1
    f.fromNumpy(flags2)
    manta.solvePressure(flags=f, vel=v, pressure=pr, cgAccuracy=5e-8, preconditioner=manta.PcNone,
                        cgMaxIterFac=2, zeroPressureFixing=True)
    assert sc.s._solve_stats["fixPidx"] == -1


@pytest.mark.parametrize("shape", [(40, 36, 1), (24, 22, 20)])
def test_enforce_compatibility(manta, oracle, rng, shape):
    """computePressureRhs with enforceCompatibility subtracts the mean of the
    fluid divergence from every cell: rhs += (Real)(-sum / (Real)cnt), the
    sum in FP64 (pressure.cpp:297-298); solvePressure with it matches the
    oracle."""
    nx, ny, nz = shape
    n = nx * ny * nz
    flags = make_flags(oracle, shape, bw=0, obstacles=3, rng=rng)
    vel = smooth_vel(rng, shape, 2, 1.0)  # no wall BCs: the RHS has a non-zero mean
    ref_rhs, cnt, s = oracle.make_rhs(shape, flags, vel)
    corr = np.float32(-s / np.float64(np.float32(cnt)))
    exp = (ref_rhs + corr).astype(np.float32)
    assert abs(float(corr)) > 1e-4
    sc = GpuScene(manta, shape)
    f, v, pr, rhs = sc.flags(flags), sc.mac(vel), sc.real(), sc.real()
    manta.computePressureRhs(rhs, v, pr, f, enforceCompatibility=True)
    got = rhs.toNumpy()
    # the FP64 sum is taken in another order: allow one float rounding step
    assert np.abs(got - exp).max() <= 2 * np.spacing(np.float32(np.abs(exp).max()))
    rr = sc.real()
    manta.solvePressure(flags=f, vel=v, pressure=pr, cgAccuracy=1e-5, preconditioner=manta.PcNone,
                        cgMaxIterFac=30, enforceCompatibility=True, retRhs=rr)
    st = sc.s._solve_stats
    ref_v, ref_p = vel.copy(), np.zeros(n, np.float32)
    ost = oracle.solve_pressure(shape, ref_v, ref_p, flags, cgAccuracy=1e-5, preconditioner=PC_NONE,
                                cgMaxIterFac=30, enforceCompatibility=True)
    assert abs(st["iterations"] - ost["iterations"]) <= 1
    assert np.abs(rr.toNumpy() - ost["rhs"]).max() <= 2 * np.spacing(np.float32(np.abs(exp).max()))
    assert _same_pressure(pr.toNumpy(), ref_p, flags, shape, 1e-4)
    assert np.abs(v.toNumpy() - ref_v).max() <= 1e-4


@pytest.mark.parametrize("shape", [(40, 36, 1), (24, 22, 20)])
def test_use_l2_norm(manta, oracle, rng, shape):
    """useL2Norm: the residual is the sum of squares (no square root) and the
    solve stops when it falls below cgAccuracy; PcNone matches the oracle's
    iteration count within one and its pressure within the tolerance."""
    n = int(np.prod(shape))
    flags = make_flags(oracle, shape, bw=1, open_bound="Y", obstacles=2, rng=rng)
    oracle.reset_outflow(shape, flags)
    vel = smooth_vel(rng, shape, 2, 1.0)
    oracle.set_wall_bcs(shape, flags, vel)
    acc = 1e-6
    sc = GpuScene(manta, shape)
    f, v, pr = sc.flags(flags), sc.mac(vel), sc.real()
    manta.solvePressure(flags=f, vel=v, pressure=pr, cgAccuracy=acc, preconditioner=manta.PcNone,
                        cgMaxIterFac=30, useL2Norm=True)
    st = sc.s._solve_stats
    ref_v, ref_p = vel.copy(), np.zeros(n, np.float32)
    ost = oracle.solve_pressure(shape, ref_v, ref_p, flags, cgAccuracy=acc, preconditioner=PC_NONE,
                                cgMaxIterFac=30, useL2Norm=True)
    assert st["converged"] and st["resNorm"] < acc
    assert abs(st["iterations"] - ost["iterations"]) <= 1
    rhs, _, _ = oracle.make_rhs(shape, flags, vel)
    r = rhs.astype(np.float64) - oracle.apply_matrix(shape, flags, pr.toNumpy(), oracle.make_laplace(shape, flags))
    assert float((r * r).sum()) < 2 * acc
    assert _same_pressure(pr.toNumpy(), ref_p, flags, shape, 1e-3)
    # the max-norm rule needs a different iteration count at the same value
    pr2 = sc.real()
    v.fromNumpy(vel)
    manta.solvePressure(flags=f, vel=v, pressure=pr2, cgAccuracy=acc, preconditioner=manta.PcNone,
                        cgMaxIterFac=30, useL2Norm=False)
    assert sc.s._solve_stats["iterations"] != st["iterations"]


def _moved(oracle, flags, shape):
    """The same domain with one more obstacle box (a moving obstacle)."""
    nx, ny, nz = shape
    f2 = flags.copy()
    g = f2.reshape(nz, ny, nx)
    if nz > 1:
        g[nz // 2:nz // 2 + 3, ny // 3:ny // 3 + 3, nx // 4:nx // 4 + 4] = OBSTACLE
    else:
        g[0, ny // 3:ny // 3 + 3, nx // 4:nx // 4 + 4] = OBSTACLE
    return f2


def _residual_ok(oracle, shape, flags, vel, p, acc):
    rhs, _, _ = oracle.make_rhs(shape, flags, vel)
    r = rhs - oracle.apply_matrix(shape, flags, p, oracle.make_laplace(shape, flags))
    return float(np.abs(r).max()) < 2 * acc


@pytest.mark.parametrize("shape", [(48, 44, 1), (28, 26, 24)])
def test_mg_dynamic_and_mic_after_flag_change(manta, oracle, rng, shape):
    """PcMGDynamic rebuilds its hierarchy every solve, and the multigrid
    hierarchy standing in for PcMIC is rebuilt when the matrix changes: after
    an obstacle moves, both give bit for bit what a fresh solver gives, and
    they converge against the oracle's matrix.  PcMGStatic keeps its first
    hierarchy, as the reference's static mode does, until releaseMG."""
    flags_a = make_flags(oracle, shape, bw=1, open_bound="Y", obstacles=2, rng=rng)
    oracle.reset_outflow(shape, flags_a)
    flags_b = _moved(oracle, flags_a, shape)
    vel = smooth_vel(rng, shape, 2, 1.0)
    vel_b = vel.copy()
    oracle.set_wall_bcs(shape, flags_b, vel_b)
    oracle.set_wall_bcs(shape, flags_a, vel)
    acc = 1e-4
    kw = dict(cgAccuracy=acc, cgMaxIterFac=20)

    def fresh(pc):
        sc2 = GpuScene(manta, shape)
        f2, v2, p2 = sc2.flags(flags_b), sc2.mac(vel_b), sc2.real()
        manta.solvePressure(flags=f2, vel=v2, pressure=p2, preconditioner=pc, **kw)
        return p2.toNumpy(), v2.toNumpy(), sc2.s._solve_stats["iterations"]

    for pc in (manta.PcMGDynamic, manta.PcMIC):
        sc = GpuScene(manta, shape)
        f, v, p = sc.flags(flags_a), sc.mac(vel), sc.real()
        manta.solvePressure(flags=f, vel=v, pressure=p, preconditioner=pc, **kw)
        assert sc.s._solve_stats["converged"]
        assert _residual_ok(oracle, shape, flags_a, vel, p.toNumpy(), acc)
        f.fromNumpy(flags_b)
        v.fromNumpy(vel_b)
        manta.solvePressure(flags=f, vel=v, pressure=p, preconditioner=pc, **kw)
        st = sc.s._solve_stats
        if shape[2] == 1 and pc == manta.PcMIC:
            assert st["preconditioner"] == 0  # 2D: mIC is off in the reference too
        assert st["converged"] and _residual_ok(oracle, shape, flags_b, vel_b, p.toNumpy(), acc)
        rp, rv, rit = fresh(pc)
        assert st["iterations"] == rit
        assert_bitexact(p.toNumpy(), rp, "pressure after flag change, pc %d" % pc)
        assert_bitexact(v.toNumpy(), rv, "velocity after flag change, pc %d" % pc)

    # PcMGStatic: kept hierarchy (reference semantics), fresh after releaseMG
    sc = GpuScene(manta, shape)
    f, v, p = sc.flags(flags_a), sc.mac(vel), sc.real()
    manta.solvePressure(flags=f, vel=v, pressure=p, preconditioner=manta.PcMGStatic, **kw)
    f.fromNumpy(flags_b)
    manta.releaseMG(sc.s)
    v.fromNumpy(vel_b)
    manta.solvePressure(flags=f, vel=v, pressure=p, preconditioner=manta.PcMGStatic, **kw)
    rp, rv, rit = fresh(manta.PcMGStatic)
    assert sc.s._solve_stats["iterations"] == rit
    assert_bitexact(p.toNumpy(), rp, "PcMGStatic after releaseMG")
    manta.releaseMG()  # all solvers


# ---------------------------------------------------------------------------
# grid operations (grid.cpp) against numpy restatements
# ---------------------------------------------------------------------------
def _ijk(shape):
    nx, ny, nz = shape
    return np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")


def _inner(shape, bnd):
    nx, ny, nz = shape
    k, j, i = _ijk(shape)
    m = (i >= bnd) & (i < nx - bnd) & (j >= bnd) & (j < ny - bnd)
    if nz > 1:
        m &= (k >= bnd) & (k < nz - bnd)
    return m.reshape(-1)


def _np_norm(v):
    """vectorbase.h:380-384 in float"""
    l = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
    eps2 = np.float32(1e-6) * np.float32(1e-6)
    r = np.sqrt(l)
    r = np.where(np.abs(l.astype(np.float64) - 1.0) < eps2, np.float32(1.0), r)
    return np.where(l <= eps2, np.float32(0.0), r).astype(np.float32)


@pytest.mark.parametrize("shape", [(19, 17, 1), (13, 12, 11)])
def test_grid_reductions(manta, rng, shape):
    """getMax/getMin/getMaxAbs/getL1/getL2 (grid.cpp:315-389) for Real, Int
    and Vec3 grids; FP64 sums returned as Real."""
    n = int(np.prod(shape))
    sc = GpuScene(manta, shape)
    a = rng.uniform(-3, 2, n).astype(np.float32)
    g = sc.real(a)
    assert g.getMax() == float(a.max()) and g.getMin() == float(a.min())
    assert g.getMaxAbs() == float(max(abs(a.min()), abs(a.max())))
    for bnd in (0, 1, 2):
        m = _inner(shape, bnd)
        l1 = np.float32(np.abs(a[m].astype(np.float64)).sum())
        assert abs(g.getL1(bnd) - float(l1)) <= float(np.spacing(l1))
        l2 = np.float32(np.sqrt((a[m].astype(np.float64) ** 2).sum()))
        assert abs(g.getL2(bnd) - float(l2)) <= float(np.spacing(l2))
    ia = rng.integers(-50, 40, n).astype(np.int32)
    gi = sc.s.create(manta.IntGrid)
    gi.fromNumpy(ia)
    assert gi.getMax() == float(ia.max()) and gi.getMin() == float(ia.min())
    assert gi.getMaxAbs() == float(max(abs(ia.min()), abs(ia.max())))
    assert gi.getL1(1) == float(np.abs(ia[_inner(shape, 1)]).sum())
    va = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    va[0] = (1.0, 0.0, 0.0)   # |v|^2 == 1 exactly: norm() returns 1
    va[1] = (0.0, 0.0, 0.0)
    gv = sc.vec(va)
    ns = (va[:, 0] * va[:, 0] + va[:, 1] * va[:, 1]) + va[:, 2] * va[:, 2]
    assert gv.getMax() == float(np.sqrt(ns.max())) and gv.getMaxAbs() == float(np.sqrt(ns.max()))
    assert gv.getMin() == float(np.sqrt(ns.min()))
    for bnd in (0, 1):
        m = _inner(shape, bnd)
        l1 = np.float32(_np_norm(va)[m].astype(np.float64).sum())
        assert abs(gv.getL1(bnd) - float(l1)) <= float(np.spacing(l1))
        l2 = np.float32(np.sqrt(ns[m].astype(np.float64).sum()))
        assert abs(gv.getL2(bnd) - float(l2)) <= float(np.spacing(l2))


@pytest.mark.parametrize("shape", [(19, 17, 1), (13, 12, 11)])
def test_set_bound_ops(manta, rng, shape):
    """setBound, setBoundNeumann (grid.cpp:585-627) and setBoundMAC with and
    without normalOnly (:630-648), bit for bit against numpy."""
    nx, ny, nz = shape
    n = nx * ny * nz
    is3 = nz > 1
    k, j, i = [x.reshape(-1) for x in _ijk(shape)]
    sc = GpuScene(manta, shape)
    a = rng.uniform(-1, 1, n).astype(np.float32)
    for w in (0, 1, 2):
        g = sc.real(a)
        g.setBound(2.5, w)
        m = (i <= w) | (i >= nx - 1 - w) | (j <= w) | (j >= ny - 1 - w)
        if is3:
            m |= (k <= w) | (k >= nz - 1 - w)
        assert_bitexact(g.toNumpy(), np.where(m, np.float32(2.5), a), "setBound w=%d" % w)
        g = sc.real(a)
        g.setBoundNeumann(w)
        si, sj, sk = i.copy(), j.copy(), k.copy()
        si[i <= w] = w + 1
        si[i >= nx - 1 - w] = nx - 2 - w
        sj[j <= w] = w + 1
        sj[j >= ny - 1 - w] = ny - 2 - w
        if is3:
            sk[k <= w] = w + 1
            sk[k >= nz - 1 - w] = nz - 2 - w
        src = si + nx * sj + (nx * ny * sk if is3 else 0)
        assert_bitexact(g.toNumpy(), a[src], "setBoundNeumann w=%d" % w)
    va = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    val = np.array([0.5, -0.25, 0.75], np.float32)
    for w in (0, 1, 2):
        for normal in (False, True):
            g = sc.mac(va)
            g.setBoundMAC(manta.vec3(*val), w, normal)
            e = va.copy()
            zlo = (k <= w - 1) if is3 else np.zeros(n, bool)
            zhi = (k >= nz - 1 - w) if is3 else np.zeros(n, bool)
            if normal:
                mx = (i <= w) | (i >= nx - w)
                my = (j <= w) | (j >= ny - w)
                mz = ((k <= w) | (k >= nz - w)) if is3 else np.zeros(n, bool)
            else:
                mx = (i <= w) | (i >= nx - w) | (j <= w - 1) | (j >= ny - 1 - w) | zlo | zhi
                my = (i <= w - 1) | (i >= nx - 1 - w) | (j <= w) | (j >= ny - w) | zlo | zhi
                mz = (i <= w - 1) | (i >= nx - 1 - w) | (j <= w - 1) | (j >= ny - 1 - w)
                if is3:
                    mz |= (k <= w) | (k >= nz - w)
            e[mx, 0] = val[0]
            e[my, 1] = val[1]
            e[mz, 2] = val[2]
            assert_bitexact(g.toNumpy(), e, "setBoundMAC w=%d normalOnly=%s" % (w, normal))


@pytest.mark.parametrize("shape", [(19, 17, 1), (13, 12, 11)])
def test_count_cells_maxdiff_components(manta, oracle, rng, shape):
    """countCells with bnd and mask (grid.cpp:384-396), gridMaxDiff /
    gridMaxDiffInt / gridMaxDiffVec3 (:400-430), getComponent /
    setComponent (:680-685) and copyMacToVec3 (:433-440)."""
    n = int(np.prod(shape))
    flags = make_flags(oracle, shape, bw=1, open_bound="Y", obstacles=2, rng=rng)
    sc = GpuScene(manta, shape)
    f = sc.flags(flags)
    mask = sc.real(rng.uniform(0, 1, n).astype(np.float32))
    for flag in (FLUID, OBSTACLE, EMPTY | 16, 2 | 1):
        for bnd in (0, 1, 2):
            hit = ((flags & flag) != 0) & (_inner(shape, bnd) if bnd > 0 else True)
            assert f.countCells(flag, bnd, mask) == int(hit.sum())
            assert_bitexact(mask.toNumpy(), hit.astype(np.float32), "countCells mask")
    a = rng.uniform(-1, 1, n).astype(np.float32)
    b = a.copy()
    b[7] += np.float32(0.375)
    b[n - 3] -= np.float32(0.5)
    # fabs of the float difference, maximum in double, returned as Real
    assert manta.gridMaxDiff(sc.real(a), sc.real(b)) == float(np.abs(a - b).max())
    ia = rng.integers(-9, 9, n).astype(np.int32)
    ib = ia.copy()
    ib[5] += 7
    g1, g2 = sc.s.create(manta.IntGrid), sc.s.create(manta.IntGrid)
    g1.fromNumpy(ia)
    g2.fromNumpy(ib)
    assert manta.gridMaxDiffInt(g1, g2) == 7.0
    va = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    vb = va + rng.uniform(-0.1, 0.1, (n, 3)).astype(np.float32)
    d = np.abs(va.astype(np.float64) - vb.astype(np.float64))
    exp = np.float32(((d[:, 0] + d[:, 1]) + d[:, 2]).max())
    assert manta.gridMaxDiffVec3(sc.vec(va), sc.vec(vb)) == float(exp)
    gv, gr = sc.vec(va), sc.real()
    for c in range(3):
        manta.getComponent(gv, gr, c)
        assert_bitexact(gr.toNumpy(), va[:, c], "getComponent %d" % c)
    gr.fromNumpy(a)
    manta.setComponent(gr, gv, 1)
    e = va.copy()
    e[:, 1] = a
    assert_bitexact(gv.toNumpy(), e, "setComponent")
    gm = sc.mac(vb)
    manta.copyMacToVec3(gm, gv)
    assert_bitexact(gv.toNumpy(), vb, "copyMacToVec3")
