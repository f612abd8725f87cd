"""Pins of the oracle's second-order obstacle and free-surface boundary
functions (pressure.cpp:32-84,111-219,455-476, conjugategrad.h:154-187,
extforces.cpp:220-315), CPU only.  Each check is a closed form or an
identity that a wrong sign, index, face or clamp would break:

- the ghost-fluid factor (thetaHelper / ghostFluidHelper) on hand-derived
  level-set pairs, clamps included;
- MakeRhs with fractions = 1 equals the plain RHS bit for bit; the obstacle
  velocity term with vel = 0 equals the fraction term with fractions 1 - f
  and vel = obvel bit for bit; perCellCorr adds exactly; the RHS of an
  all-fluid interior telescopes to the boundary flux; surface tension on a
  hand-built cell;
- MakeLaplaceMatrix with the fractions updateFractions would give a voxel
  domain (1 open, 0 at obstacles) equals the plain matrix; random fractions
  give a symmetric positive definite system with an empty cell; the
  ghost-fluid diagonal on a hand-built cell;
- correctVelocity's ghost-fluid terms and the clamped-face replacement on
  hand-built cells;
- the solve: the CG result against a dense FP64 solve of the same system;
- setWallBcs with phiObs: for a planar obstacle SDF the face velocity loses
  exactly its component along the plane normal (central differences of a
  linear function are exact), other faces keep their values bit for bit."""
import numpy as np
import pytest

from oracle import EMPTY, FLUID, OBSTACLE, PC_NONE, Oracle


@pytest.fixture(scope="module")
def o():
    return Oracle()


@pytest.mark.parametrize("pin,pout,expect", [
    (-0.25, 0.75, -3.0),       # theta 0.25
    (-1.0, 1.0, -1.0),         # theta 0.5
    (0.5, 0.5, -1.0),          # denom > -1e-4: theta 0.5
    (-2.0, 0.5, -0.25),        # theta 0.8
    (-3.0, -1.0, 0.0),         # theta 1.5 -> clamped to 1
    (0.1, 0.6, 1e-4),          # theta < 0 -> 0 -> below gfClamp: returns gfClamp itself
    (-1e-5, 0.99999, 1e-4),    # theta 1e-5 < gfClamp
])
def test_ghost_fluid_factor(o, pin, pout, expect):
    got = o.ghost_fluid_factor(pin, pout, 1e-4)
    assert abs(got - expect) <= 1e-6 * max(1.0, abs(expect)), (pin, pout, got)


def _domain(shape, rng=None, obstacles=0, empty_top=False):
    nx, ny, nz = shape
    f = np.full((nz, ny, nx), FLUID, np.int32)
    f[:, :, 0] = f[:, :, -1] = OBSTACLE
    f[:, 0, :] = f[:, -1, :] = OBSTACLE
    if nz > 1:
        f[0] = f[-1] = OBSTACLE
    if empty_top:
        if nz > 1:
            f[1:-1, -2, 1:-1] = EMPTY
        else:
            f[:, -2, 1:-1] = EMPTY
    for _ in range(obstacles):
        x, y = rng.integers(2, max(3, nx - 3)), rng.integers(2, max(3, ny - 3))
        z = rng.integers(2, max(3, nz - 2)) if nz > 1 else 0
        f[z, y, x] = OBSTACLE
    return f.reshape(-1)


def _voxel_fractions(flags, shape):
    """1 on faces between two non-obstacle cells, 0 on faces touching an
    obstacle (lower neighbour outside the grid counts as obstacle)."""
    nx, ny, nz = shape
    ob = ((flags & OBSTACLE) != 0).reshape(nz, ny, nx)
    fr = np.zeros((nz, ny, nx, 3), np.float32)
    lo = np.ones_like(ob)
    lo[:, :, 1:] = ob[:, :, :-1]
    fr[..., 0] = (~ob & ~lo).astype(np.float32)
    lo = np.ones_like(ob)
    lo[:, 1:, :] = ob[:, :-1, :]
    fr[..., 1] = (~ob & ~lo).astype(np.float32)
    if nz > 1:
        lo = np.ones_like(ob)
        lo[1:] = ob[:-1]
        fr[..., 2] = (~ob & ~lo).astype(np.float32)
    return fr.reshape(-1, 3)


@pytest.mark.parametrize("shape", [(9, 8, 1), (7, 6, 5)])
def test_rhs_fractions_one_equals_plain(o, shape):
    rng = np.random.default_rng(1)
    n = int(np.prod(shape))
    flags = _domain(shape, rng, obstacles=2)
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    plain, cnt0, s0 = o.make_rhs(shape, flags, vel)
    b = o.bnd(fractions=np.ones((n, 3), np.float32))
    got, cnt1, s1 = o.make_rhs_bnd(shape, flags, vel, b)
    assert np.array_equal(got, plain) and cnt0 == cnt1 and s0 == s1


@pytest.mark.parametrize("shape", [(9, 8, 1), (7, 6, 5)])
def test_rhs_obvel_term_is_the_complementary_fraction_term(o, shape):
    rng = np.random.default_rng(2)
    n = int(np.prod(shape))
    flags = _domain(shape, rng, obstacles=2)
    fr = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    w = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    a, _, _ = o.make_rhs_bnd(shape, flags, np.zeros((n, 3), np.float32), o.bnd(fractions=fr, obvel=w))
    b, _, _ = o.make_rhs_bnd(shape, flags, w, o.bnd(fractions=(np.float32(1) - fr).astype(np.float32)))
    assert np.array_equal(a, b)


def test_rhs_per_cell_correction_adds(o):
    shape = (8, 7, 6)
    rng = np.random.default_rng(3)
    n = int(np.prod(shape))
    flags = _domain(shape, rng, obstacles=3)
    vel = rng.integers(-4, 5, (n, 3)).astype(np.float32)
    pcc = rng.integers(-3, 4, n).astype(np.float32)
    a, _, _ = o.make_rhs_bnd(shape, flags, vel, o.bnd())
    b, _, _ = o.make_rhs_bnd(shape, flags, vel, o.bnd(perCellCorr=pcc))
    fl = (flags & FLUID) != 0
    nx, ny, nz = shape
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    m = fl & inner.reshape(-1)
    assert np.array_equal(b[m], a[m] + pcc[m]) and np.array_equal(b[~m], a[~m])


def test_rhs_fractions_telescope_to_boundary_flux(o):
    """All interior cells fluid: sum over cells of sum_faces (+-) f u equals
    the flux through the faces of the interior box."""
    shape = (9, 8, 7)
    nx, ny, nz = shape
    n = nx * ny * nz
    rng = np.random.default_rng(4)
    flags = _domain(shape)
    fr = rng.uniform(0.2, 1, (n, 3)).astype(np.float32)
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    rhs, cnt, s = o.make_rhs_bnd(shape, flags, vel, o.bnd(fractions=fr))
    F = (fr.astype(np.float64) * vel).reshape(nz, ny, nx, 3)
    flux = (F[1:-1, 1:-1, 1, 0].sum() - F[1:-1, 1:-1, nx - 1, 0].sum() + F[1:-1, 1, 1:-1, 1].sum() -
            F[1:-1, ny - 1, 1:-1, 1].sum() + F[1, 1:-1, 1:-1, 2].sum() - F[nz - 1, 1:-1, 1:-1, 2].sum())
    assert cnt == (nx - 2) * (ny - 2) * (nz - 2)
    assert abs(float(rhs.astype(np.float64).sum()) - flux) < 1e-4 and abs(s - flux) < 1e-4


def _one_cell_case():
    """5x5x1: centre cell (2,2) fluid, its -x and +y neighbours empty, the
    rest fluid inside a wall ring"""
    shape = (5, 5, 1)
    f = np.full((1, 5, 5), FLUID, np.int32)
    f[0, 0, :] = f[0, -1, :] = f[0, :, 0] = f[0, :, -1] = OBSTACLE
    f[0, 2, 1] = EMPTY    # (i=1, j=2): -x neighbour of (2,2)
    f[0, 3, 2] = EMPTY    # (i=2, j=3): +y neighbour
    return shape, f.reshape(-1)


def test_surface_tension_rhs_closed_form(o):
    shape, flags = _one_cell_case()
    n = 25
    phi = np.full(n, -1.0, np.float32)
    c = 2 + 5 * 2
    phi[c] = -0.25
    phi[c - 1] = 0.75      # theta 0.25 -> factor -3
    phi[c + 5] = 0.25      # theta 0.5 -> factor -1
    curv = np.zeros(n, np.float32)
    curv[c], curv[c - 1], curv[c + 5] = 0.5, 2.0, 4.0
    vel = np.zeros((n, 3), np.float32)
    st = 0.1
    rhs, _, _ = o.make_rhs_bnd(shape, flags, vel, o.bnd(phi=phi, curv=curv, surfTens=st))
    expect = st * (2.0 - (-3.0) * 0.5) + st * (4.0 - (-1.0) * 0.5)
    assert abs(rhs[c] - expect) < 1e-6, (rhs[c], expect)


@pytest.mark.parametrize("shape", [(9, 8, 1), (7, 6, 5)])
def test_laplace_voxel_fractions_equal_plain(o, shape):
    rng = np.random.default_rng(5)
    flags = _domain(shape, rng, obstacles=3, empty_top=True)
    A = o.make_laplace(shape, flags)
    B = o.make_laplace_bnd(shape, flags, o.bnd(fractions=_voxel_fractions(flags, shape)))
    for a, b in zip(A, B):
        assert np.array_equal(a, b)


def _dense(o, shape, flags, A):
    """dense FP64 matrix of ApplyMatrix over the fluid cells"""
    n = int(np.prod(shape))
    fl = np.nonzero(flags & FLUID)[0]
    M = np.zeros((fl.size, fl.size))
    for a, ix in enumerate(fl):
        e = np.zeros(n, np.float32)
        e[ix] = 1
        col = o.apply_matrix(shape, flags, e, A).astype(np.float64)
        M[:, a] = col[fl]
    return M, fl


def test_laplace_random_fractions_spd(o):
    shape = (6, 6, 5)
    rng = np.random.default_rng(6)
    n = int(np.prod(shape))
    flags = _domain(shape, rng, obstacles=2, empty_top=True)
    fr = rng.uniform(0.1, 1, (n, 3)).astype(np.float32)
    A = o.make_laplace_bnd(shape, flags, o.bnd(fractions=fr))
    M, _ = _dense(o, shape, flags, A)
    assert np.allclose(M, M.T)
    assert np.linalg.eigvalsh(M).min() > 1e-6


def test_ghost_fluid_diagonal_closed_form(o):
    shape, flags = _one_cell_case()
    n = 25
    c = 12
    phi = np.full(n, -1.0, np.float32)
    phi[c] = -0.25
    phi[c - 1] = 0.75      # factor -3
    phi[c + 5] = 0.1       # theta 0.25/0.35 -> 1 - 1.4 = -0.4
    A = o.make_laplace_bnd(shape, flags, o.bnd(phi=phi))
    P = o.make_laplace(shape, flags)
    f2 = 1.0 - 1.0 / (0.25 / 0.35)
    assert abs(A[0][c] - (P[0][c] + 3.0 - f2)) < 1e-6, (A[0][c], P[0][c])
    # every other fluid row: minus the factor of each empty face neighbour
    # (the factor itself is pinned above); off-diagonals untouched
    for ix in range(n):
        if ix == c:
            continue
        if not (flags[ix] & FLUID):
            assert A[0][ix] == P[0][ix]
            continue
        want = float(P[0][ix])
        for off in (-1, 1, -5, 5):
            if flags[ix + off] & EMPTY:
                want -= o.ghost_fluid_factor(phi[ix], phi[ix + off], 1e-4)
        assert abs(A[0][ix] - want) < 1e-6, ix
    assert all(np.array_equal(a, b) for a, b in zip(A[1:], P[1:]))


def test_correct_velocity_ghost_fluid_closed_form(o):
    shape, flags = _one_cell_case()
    n = 25
    c = 12
    phi = np.full(n, -1.0, np.float32)
    phi[c] = -0.25
    phi[c - 1] = 0.75                      # factor(c, -X) = -3
    p = np.zeros(n, np.float32)
    p[c], p[c - 5], p[c + 4], p[c + 5] = 2.0, 0.5, 0.75, 0.3
    vel = np.zeros((n, 3), np.float32)
    vel[c] = (1.0, 1.0, 0.0)
    vel[c + 5] = (0.25, 0.5, 0.0)         # the empty cell above c
    ref = vel.copy()
    o.correct_velocity(shape, flags, ref, p)   # the plain correction
    got = vel.copy()
    o.correct_velocity_bnd(shape, flags, got, p, o.bnd(phi=phi))
    # fluid cell c, empty -x neighbour: += p[c] * factor = 2 * -3
    assert abs(got[c, 0] - (ref[c, 0] + 2.0 * -3.0)) < 1e-6
    # empty cell above c (fluid lower neighbour c): y face -= p[c] * factor(c, +Y)
    fy = o.ghost_fluid_factor(phi[c], phi[c + 5], 1e-4)
    assert abs(got[c + 5, 1] - (ref[c + 5, 1] - 2.0 * fy)) < 1e-6
    # its x face: the lower x neighbour (i=1, j=3) is fluid, factor from phi (-1 vs -1): theta 0.5
    assert abs(got[c + 5, 0] - (ref[c + 5, 0] - p[c + 4] * -1.0)) < 1e-6


def test_replace_clamped_ghost_fluid_velocity(o):
    shape, flags = _one_cell_case()
    n = 25
    c = 12
    phi = np.full(n, -1.0, np.float32)
    phi[c] = 0.2          # the fluid cell's level set is positive: theta clamps to 0 < gfClamp
    phi[c - 1] = 0.7
    p = np.zeros(n, np.float32)
    vel = np.random.default_rng(7).uniform(-1, 1, (n, 3)).astype(np.float32)
    got = vel.copy()
    o.correct_velocity_bnd(shape, flags, got, p, o.bnd(phi=phi))
    # empty cell c-1 has the fluid +x neighbour c whose face towards it was
    # clamped: its x face takes c's x face
    assert got[c - 1, 0] == got[c, 0]


@pytest.mark.parametrize("shape,pc", [((10, 9, 1), PC_NONE), ((8, 7, 6), PC_NONE)])
def test_solve_bnd_against_dense(shape, pc):
    o64 = Oracle(f64=True)
    rng = np.random.default_rng(8)
    n = int(np.prod(shape))
    flags = _domain(shape, rng, obstacles=2, empty_top=True)
    fr = rng.uniform(0.3, 1, (n, 3))
    phi = rng.uniform(-1, 0.5, n)
    vel = rng.uniform(-1, 1, (n, 3))
    b = o64.bnd(phi=phi, fractions=fr)
    A = o64.make_laplace_bnd(shape, flags, b)
    rhs, _, _ = o64.make_rhs_bnd(shape, flags, vel.copy(), b)
    # dense solve in FP64 over the fluid cells
    fl = np.nonzero(flags & FLUID)[0]
    Md = np.zeros((fl.size, fl.size))
    for a_, ix in enumerate(fl):
        e = np.zeros(n)
        e[ix] = 1
        Md[:, a_] = o64.apply_matrix(shape, flags, e, A)[fl]
    xd = np.linalg.solve(Md, rhs[fl])
    p = np.zeros(n)
    v = vel.copy()
    st = o64.solve_pressure_bnd(shape, v, p, flags, b, cgAccuracy=1e-12, preconditioner=pc, cgMaxIterFac=20.0)
    assert np.abs(p[fl] - xd).max() < 1e-8 * max(1.0, np.abs(xd).max()), st


@pytest.mark.parametrize("normal", [(1.0, 0.0, 0.0), (0.6, 0.8, 0.0), (0.48, 0.64, 0.6)])
def test_wall_bcs_frac_planar_sdf(o, normal):
    shape = (9, 8, 7) if normal[2] else (9, 8, 1)
    nx, ny, nz = shape
    n = nx * ny * nz
    a, b_, c = normal
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    # cell-centred SDF of the plane n . (x - x0) = 0 through the domain centre
    phi = (a * (i + 0.5 - nx / 2) + b_ * (j + 0.5 - ny / 2) + c * ((k + 0.5 - nz / 2) if nz > 1 else 0)).reshape(-1)
    phi = phi.astype(np.float32)
    flags = np.where(phi < 0, OBSTACLE, FLUID).astype(np.int32)
    vel = np.random.default_rng(9).uniform(-1, 1, (n, 3)).astype(np.float32)
    if nz == 1:
        vel[:, 2] = 0
    got = vel.copy()
    o.set_wall_bcs_frac(shape, flags, got, phi)
    V = vel.astype(np.float64).reshape(nz, ny, nx, 3)
    F = flags.reshape(nz, ny, nx)
    G = got.reshape(nz, ny, nx, 3)
    nv = np.array(normal) / np.linalg.norm(normal)
    changed = 0
    for kk in range(nz):
        for jj in range(ny):
            for ii in range(nx):
                inner = 1 <= ii < nx - 1 and 1 <= jj < ny - 1 and (nz == 1 or 1 <= kk < nz - 1)
                cur = F[kk, jj, ii]
                for comp, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    if comp == 2 and nz == 1:
                        continue
                    touch = inner and (cur & (FLUID | OBSTACLE)) and \
                        ((cur & OBSTACLE) or (F[kk - dk, jj - dj, ii - di] & OBSTACLE))
                    if not touch:
                        assert G[kk, jj, ii, comp] == vel.reshape(nz, ny, nx, 3)[kk, jj, ii, comp]
                        continue
                    # getAtMAC of the face: its own component plus the 4-face
                    # averages of the others
                    u = np.zeros(3)
                    u[comp] = V[kk, jj, ii, comp]
                    for oc in range(3):
                        if oc == comp or (oc == 2 and nz == 1):
                            continue
                        e = (0, 0, 0)
                        e = [0, 0, 0]
                        e[oc] = 1
                        s = 0.0
                        for (xi, yj, zk) in ((0, 0, 0), (-di, -dj, -dk)):
                            for t in (0, 1):
                                s += V[kk + zk + t * e[2], jj + yj + t * e[1], ii + xi + t * e[0], oc]
                        u[oc] = 0.25 * s
                    want = u[comp] - np.dot(nv, u) * nv[comp]
                    assert abs(G[kk, jj, ii, comp] - want) < 2e-6, (ii, jj, kk, comp)
                    changed += 1
    assert changed > 0


def bnd_pins():
    """The checks above as (name, function of a float oracle) pairs."""
    out = []
    for args in [(-0.25, 0.75, -3.0), (-2.0, 0.5, -0.25), (-3.0, -1.0, 0.0), (0.1, 0.6, 1e-4)]:
        out.append(("ghost_factor%s" % (args,), lambda o, a=args: test_ghost_fluid_factor(o, *a)))
    for shape in [(9, 8, 1), (7, 6, 5)]:
        out.append(("rhs_fr1%s" % (shape,), lambda o, s=shape: test_rhs_fractions_one_equals_plain(o, s)))
        out.append(("rhs_obvel%s" % (shape,), lambda o, s=shape: test_rhs_obvel_term_is_the_complementary_fraction_term(o, s)))
        out.append(("lap_voxel%s" % (shape,), lambda o, s=shape: test_laplace_voxel_fractions_equal_plain(o, s)))
    out += [("rhs_pcc", test_rhs_per_cell_correction_adds), ("rhs_telescope", test_rhs_fractions_telescope_to_boundary_flux),
            ("rhs_surftens", test_surface_tension_rhs_closed_form), ("lap_spd", test_laplace_random_fractions_spd),
            ("ghost_diag", test_ghost_fluid_diagonal_closed_form),
            ("correct_gf", test_correct_velocity_ghost_fluid_closed_form),
            ("replace_clamped", test_replace_clamped_ghost_fluid_velocity)]
    for nv in [(1.0, 0.0, 0.0), (0.6, 0.8, 0.0), (0.48, 0.64, 0.6)]:
        out.append(("wall_frac%s" % (nv,), lambda o, v=nv: test_wall_bcs_frac_planar_sdf(o, v)))
    return out


def test_planted_mistakes_in_the_boundary_functions_are_rejected(tmp_path):
    """Oracle builds with mistakes 21-24 (the list at the top of
    oracle/mfx_oracle.c) each fail at least one of the checks above."""
    import os
    import subprocess
    odir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    for k in (21, 22, 23, 24):
        subprocess.run(["make", "-s", "-C", odir, "mutant", "MUT=%d" % k, "OUT=%s" % tmp_path], check=True,
                       capture_output=True)
        o = Oracle(path=str(tmp_path / ("liboracle_m%d.so" % k)))
        failed = []
        for name, pin in bnd_pins():
            try:
                pin(o)
            except AssertionError:
                failed.append(name)
        assert failed, "planted mistake %d passes every boundary check" % k
