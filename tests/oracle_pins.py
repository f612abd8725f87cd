"""Checks that pin the oracle's MAC and centred-Vec3 advection, the outflow
extrapolation and the 3D wall boundary conditions against facts that do not
depend on the oracle: mirror equivariance of the whole advection step,
closed forms for linear fields and integer shifts, a hand-derived outflow
column, and a face-centred statement of the wall rules.

Each ``pin_*`` takes an oracle object and raises AssertionError on a
mismatch.  ``tests/test_oracle_mac.py`` runs them against the oracle;
``tests/test_oracle_mutants.py`` runs them against builds of the oracle with
planted mistakes (the list at the top of oracle/mfx_oracle.c) and requires
every planted mistake to be rejected by at least one of them.

Reference sources: advection.cpp:44-115,191-235,270-287,326-436 (SL-MAC,
MacCormack correct/clamp, outflow), grid.h:424-470 (MAC accessors),
interpol.h:50-90 (trilinear lookup), extforces.cpp:186-217 (KnSetWallBcs)."""
import numpy as np

from oracle import EMPTY, FLUID, OBSTACLE, OUTFLOW, STICK, TYPE_MAC, TYPE_REAL, TYPE_VEC3
from tests.helpers import make_flags

MARGIN = 4  # cells next to the mirrored sides, where face 0 has no mirror image


def _mirror_cell(a, ax):
    return np.ascontiguousarray(np.flip(a, ax))


def _mirror_mac(v, axis):
    """Mirror a MAC field (nz, ny, nx, 3) across grid axis `axis` (0 = x).
    Tangential components flip with the cells; the normal component at face
    j of the mirrored grid is minus the original's at face N - j (face 0 has
    no image and is set to 0)."""
    ax = 2 - axis
    w = np.flip(v, ax).copy()
    comp = w[..., axis]
    moved = np.zeros_like(comp)
    dst = [slice(None)] * 3
    src = [slice(None)] * 3
    dst[ax] = slice(1, None)
    src[ax] = slice(None, -1)
    moved[tuple(dst)] = -comp[tuple(src)]
    w[..., axis] = moved
    return np.ascontiguousarray(w)


def pin_mac_mirror(o, is3d, order, clampMode, axis, seed=11):
    """MAC self-advection commutes with mirroring the domain (flags and
    velocity) across an axis: the invariant of the reference's
    tools/tests/test_2005_symmAdv.py, stated for a whole random problem."""
    rng = np.random.default_rng(seed)
    shape = (18, 16, 14) if is3d else (20, 18, 1)
    nx, ny, nz = shape
    flags = make_flags(o, shape, bw=1, obstacles=4, rng=rng).reshape(nz, ny, nx)
    vel = rng.uniform(-1.0, 1.0, (nz, ny, nx, 3)).astype(np.float32)
    if not is3d:
        vel[..., 2] = 0
    dt = np.float32(0.8)
    a = vel.copy().reshape(-1, 3)
    o.advect(shape, dt, flags.reshape(-1).copy(), a.copy(), a, TYPE_MAC, order, 1.0, clampMode)
    fm = _mirror_cell(flags, 2 - axis).reshape(-1)
    vm = _mirror_mac(vel, axis).reshape(-1, 3)
    b = vm.copy()
    o.advect(shape, dt, fm, vm, b, TYPE_MAC, order, 1.0, clampMode)
    ma = _mirror_mac(a.reshape(nz, ny, nx, 3), axis)
    bb = b.reshape(nz, ny, nx, 3)
    sl = [slice(None)] * 3
    sl[2 - axis] = slice(MARGIN, shape[axis] - MARGIN + 1)
    diff = np.abs(ma[tuple(sl)] - bb[tuple(sl)]).max()
    assert diff < 1e-5, "MAC advection not mirror-equivariant (axis %d): %g" % (axis, diff)
    # the problem must exercise the path: the step changed the field
    assert np.abs(a.reshape(nz, ny, nx, 3) - vel)[tuple(sl)].max() > 0.1


def _face_positions(shape):
    """Physical positions of the x, y and z faces of every cell:
    x-face (i, j+1/2, k+1/2), y-face (i+1/2, j, k+1/2), z-face (i+1/2, j+1/2, k)."""
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = np.stack([i + 0.5, j + 0.5, k + 0.5], -1).astype(np.float64)
    out = []
    for a in range(3):
        p = c.copy()
        p[..., a] -= 0.5
        if nz == 1:
            p[..., 2] = 0.5
        out.append(p)
    return out


def _linear_mac(shape, a0, G):
    """MAC field of the linear velocity u(x) = a0 + G x sampled at the faces."""
    pos = _face_positions(shape)
    v = np.zeros(pos[0].shape, np.float64)
    for c in range(3):
        v[..., c] = a0[c] + pos[c] @ G[c]
    return v


def pin_mac_linear_field(o, is3d, order, clampMode, which):
    """SL-MAC / MacCormack-MAC self-advection of a linear velocity field:
    every face value becomes u_c(x_f - u(x_f) dt) exactly (trilinear
    interpolation and the 4-face averages of getAtMACX/Y/Z are exact on linear
    fields).  For order 2 the gradient G is nilpotent (G^2 = 0), so the
    backward trace returns u exactly, the correction vanishes and the clamp
    keeps the forward value (it lies inside the 8 corners of its cell)."""
    shape = (16, 15, 14) if is3d else (18, 17, 1)
    nx, ny, nz = shape
    a0 = np.array([0.12, -0.08, 0.1 if is3d else 0.0])
    if order == 1:
        G = np.array([[0.02, -0.03, 0.015], [0.025, 0.01, -0.02], [-0.015, 0.02, 0.01]])
    else:
        # rank one, G = s v w^T with w.v = 0
        vv, ww = ((1, 1, 0), (1, -1, 0.5)) if which == 0 else ((1, 0, 1), (1, 0.5, -1))
        G = 0.03 * np.outer(vv, ww)
    if not is3d:
        G[2, :] = 0
        G[:, 2] = 0
    u = _linear_mac(shape, a0, G)
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    dt = 1.0
    g = u.astype(np.float32).reshape(-1, 3).copy()
    o.advect(shape, np.float32(dt), flags, g.copy(), g, TYPE_MAC, order, 1.0, clampMode)
    pos = _face_positions(shape)
    exp = np.zeros_like(u)
    for c in range(3):
        uf = a0 + pos[c] @ G.T             # u at the c-faces
        back = pos[c] - uf * dt
        if not is3d:
            back[..., 2] = 0.5
        exp[..., c] = a0[c] + back @ G[c]
    got = g.reshape(nz, ny, nx, 3)
    m = 3
    sl = (slice(m, nz - m) if is3d else slice(None), slice(m, ny - m), slice(m, nx - m))
    comps = 3 if is3d else 2
    diff = np.abs(got[sl][..., :comps] - exp[sl][..., :comps]).max()
    assert diff < 2e-5, "linear MAC field not reproduced: %g" % diff


def pin_integer_shift(o, gridType, order, clampMode, is3d, direction):
    """A uniform velocity of one cell per step along an axis translates any
    MAC or centred Vec3 field by exactly one cell (trilinear weights 0/1, zero
    MacCormack correction, forward value kept by the clamp)."""
    rng = np.random.default_rng(21 + direction)
    shape = (14, 12, 10) if is3d else (16, 14, 1)
    nx, ny, nz = shape
    n = nx * ny * nz
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    axis, sign = direction // 2, (1 if direction % 2 == 0 else -1)
    if axis == 2 and not is3d:
        axis = 1
    vel = np.zeros((n, 3), np.float32)
    vel[:, axis] = sign
    src = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    if not is3d:
        src[:, 2] = 0
    g = src.copy()
    o.advect(shape, np.float32(1.0), flags, vel, g, gridType, order, 1.0, clampMode)
    s4, g4 = src.reshape(nz, ny, nx, 3), g.reshape(nz, ny, nx, 3)
    ax = 2 - axis
    m = 2
    sl = [slice(m, nz - m) if is3d else slice(None), slice(m, ny - m), slice(m, nx - m)]
    shifted = np.roll(s4, sign, ax)
    assert np.array_equal(g4[tuple(sl)], shifted[tuple(sl)]), \
        "type %d not translated exactly by one cell along %d" % (gridType, axis)


def pin_centered_linear(o, gridType, order, clampMode):
    """Centred grids (Real, Vec3): a linear field under a uniform non-integer
    velocity is translated exactly (order 2: the backward trace returns the
    original, the correction vanishes, the clamp keeps the forward value);
    under a non-uniform velocity (order 1) the traced point uses the cell's
    centred velocity 0.5*(u_i + u_i+1) per axis (getCentered)."""
    rng = np.random.default_rng(5)
    shape = (16, 15, 14)
    nx, ny, nz = shape
    n = nx * ny * nz
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.stack([i + 0.5, j + 0.5, k + 0.5], -1)
    coef = np.array([[0.2, 0.03, -0.02, 0.01], [-0.1, 0.01, 0.02, -0.03], [0.3, -0.02, 0.01, 0.02]])
    f = lambda P, c: coef[c, 0] + P @ coef[c, 1:]
    nc = 1 if gridType == TYPE_REAL else 3
    src = np.stack([f(X, c) for c in range(nc)], -1).astype(np.float32).reshape(n, nc)
    if order == 2:
        vel = np.tile(np.array([0.37, -0.61, 0.29], np.float32), (n, 1))
        cen = vel.reshape(nz, ny, nx, 3).astype(np.float64)
    else:
        vel = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
        v4 = vel.reshape(nz, ny, nx, 3).astype(np.float64)
        cen = np.stack([0.5 * (v4[..., 0] + np.roll(v4[..., 0], -1, 2)),
                        0.5 * (v4[..., 1] + np.roll(v4[..., 1], -1, 1)),
                        0.5 * (v4[..., 2] + np.roll(v4[..., 2], -1, 0))], -1)
    g = src.reshape(-1).copy() if nc == 1 else src.copy()
    o.advect(shape, np.float32(1.0), flags, vel, g, gridType, order, 1.0, clampMode)
    back = X - cen
    exp = np.stack([f(back, c) for c in range(nc)], -1)
    got = g.reshape(nz, ny, nx, nc)
    sl = (slice(3, -3),) * 3
    diff = np.abs(got[sl] - exp[sl]).max()
    assert diff < 2e-5, "linear centred field not reproduced (type %d): %g" % (gridType, diff)


def pin_outflow_column(o, is3d):
    """extrapolateVelConvectiveBC + copyChangedVels on a hand-built open top:
    fluid rows j = 1..5, outflow rows j = 6, 7, walls at the sides and bottom,
    velocity constant within each row (V5, V6, V7 in fluid/outflow cells, a
    large value in walls, which the bulk average must not see).  Derivation
    with dt = 0.5 (timeStep = max(1, 4 dt) = 2), d = vel - velPrev:
      row 6: only the y axis finds a fluid cell, at distance 1 (j = 5):
             v = d / (2 * max(1, bulk_y)) + V5, bulk_y = mean(V5, V6, V7).y
      row 7: the y axis finds fluid at distance 2 (j = 5), the value added is
             that of the adjacent cell (j = 6):
             v = d / (2 * max(1, bulk_y)) + V6, bulk_y = mean(V6, V7).y
    x and z find no fluid within two cells, so each outflow cell has one
    contribution.  All other cells keep their values."""
    nx, ny = 5, 8
    nz = 5 if is3d else 1
    shape = (nx, ny, nz)
    n = nx * ny * nz
    fl = np.full((nz, ny, nx), OBSTACLE, np.int32)
    zs = slice(1, nz - 1) if is3d else slice(None)
    fl[zs, 1:6, 1:nx - 1] = FLUID
    fl[zs, 6:8, 1:nx - 1] = OUTFLOW | EMPTY
    V = {5: (0.25, 2.5, 0.5), 6: (0.5, 3.0, -0.25), 7: (0.75, 3.5, 0.125)}
    vel = np.full((nz, ny, nx, 3), 100.0, np.float32)
    for j in range(1, 8):
        vel[zs, j, 1:nx - 1] = V.get(j, (0.1 * j, 0.2 * j, -0.1 * j))
    if not is3d:
        vel[..., 2] = np.where(vel[..., 2] == 100.0, 100.0, 0.0)
    delta = np.array([0.75, 1.5, 0.375 if is3d else 0.0], np.float32)
    prev = vel - delta
    v = vel.reshape(-1, 3).copy()
    o.apply_outflow_bc(shape, fl.reshape(-1).copy(), v, prev.reshape(-1, 3).copy(), np.float32(0.5))
    got = v.reshape(nz, ny, nx, 3)
    f32 = np.float32
    b6 = f32(2.0) * f32(3.0)       # bulk_y(row 6) = (2.5 + 3 + 3.5) / 3
    b7 = f32(2.0) * f32(3.25)      # bulk_y(row 7) = (3 + 3.5) / 2
    exp6 = (delta / b6).astype(np.float32) + np.array(V[5], np.float32)
    exp7 = (delta / b7).astype(np.float32) + np.array(V[6], np.float32)
    assert np.array_equal(exp6[:2], np.array([0.375, 2.75], np.float32))  # the derivation, by hand
    g6 = got[zs, 6, 1:nx - 1].reshape(-1, 3)
    g7 = got[zs, 7, 1:nx - 1].reshape(-1, 3)
    comps = 3 if is3d else 2
    assert np.array_equal(g6[:, :comps], np.tile(exp6[:comps], (len(g6), 1))), (g6[0], exp6)
    assert np.array_equal(g7[:, :comps], np.tile(exp7[:comps], (len(g7), 1))), (g7[0], exp7)
    keep = np.ones((nz, ny, nx), bool)
    keep[zs, 6:8, 1:nx - 1] = False
    assert np.array_equal(got[keep], vel[keep])


def pin_wall_bcs_3d(o, seed=3):
    """KnSetWallBcs on a hand-built 3D case with an obstacle, a stick obstacle,
    an empty cell and per-cell obstacle velocities.  Stated face by face:
    the a-face between cells L = c - e_a and c takes obvel_a(c) iff both cells
    are fluid or obstacle and at least one is an obstacle; then a fluid cell
    with a stick neighbour along axis b (either side) has its components
    other than b set to 0; every other face is unchanged."""
    rng = np.random.default_rng(seed)
    shape = (7, 6, 6)
    nx, ny, nz = shape
    fl = o.init_domain(shape).reshape(nz, ny, nx)
    o.fill_grid(shape, fl.reshape(-1))
    fl[2, 2, 2] = OBSTACLE
    fl[3, 3, 4] = OBSTACLE | STICK
    fl[1, 4, 1] = EMPTY
    fl[4, 1, 3] = OBSTACLE
    vel = rng.uniform(0.5, 1.5, (nz, ny, nx, 3)).astype(np.float32)
    obv = rng.uniform(-3, -2, (nz, ny, nx, 3)).astype(np.float32)
    v = vel.reshape(-1, 3).copy()
    o.set_wall_bcs(shape, fl.reshape(-1).copy(), v, obv.reshape(-1, 3).copy())
    got = v.reshape(nz, ny, nx, 3)
    # explicit faces around the obstacle at (i,j,k) = (2,2,2)
    assert got[2, 2, 2, 0] == obv[2, 2, 2, 0]   # x-face between fluid (1,2,2) and the obstacle
    assert got[2, 2, 3, 0] == obv[2, 2, 3, 0]   # x-face between the obstacle and fluid (3,2,2)
    assert got[2, 3, 2, 1] == obv[2, 3, 2, 1]   # y-face above the obstacle, value of the fluid cell
    assert got[3, 2, 2, 2] == obv[3, 2, 2, 2]   # z-face above the obstacle
    assert got[1, 2, 2, 0] == vel[1, 2, 2, 0]   # fluid-fluid x-face of cell (2,2,1)
    # fluid (3,3,3) has the stick cell (4,3,3) as x neighbour: y and z zero,
    # and its x component stays (its lower x neighbour is fluid)
    assert got[3, 3, 3, 1] == 0 and got[3, 3, 3, 2] == 0 and got[3, 3, 3, 0] == vel[3, 3, 3, 0]
    # the empty cell (1,4,1) is untouched although its neighbour is a wall
    assert np.array_equal(got[1, 4, 1], vel[1, 4, 1])
    # every face, by the face-centred statement
    fo = (fl & (FLUID | OBSTACLE)) != 0
    ob = (fl & OBSTACLE) != 0
    fluid = (fl & FLUID) != 0
    stick = (fl & STICK) != 0
    exp = vel.copy()
    for a in range(3):
        ax = 2 - a
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(None, -1)
        hi[ax] = slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        hit = fo[hi] & fo[lo] & (ob[hi] | ob[lo])
        e = exp[hi + (a,)]
        e[hit] = obv[hi + (a,)][hit]
    for b in range(3):
        ax = 2 - b
        nb = np.zeros_like(stick)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(None, -1)
        hi[ax] = slice(1, None)
        nb[tuple(hi)] |= stick[tuple(lo)]
        nb[tuple(lo)] |= stick[tuple(hi)]
        m = nb & fluid
        for c in range(3):
            if c != b:
                exp[..., c][m] = 0
    assert np.array_equal(got, exp), np.argwhere(got != exp)[:5]


def all_pins():
    """(name, callable(o)) for every pin, with the parameter sets the tests use."""
    pins = []
    for is3d in (False, True):
        for order in (1, 2):
            for cm in ((1, 2) if order == 2 else (2,)):
                for axis in ((0, 1, 2) if is3d else (0, 1)):
                    pins.append(("mac_mirror-%dd-o%d-c%d-ax%d" % (3 if is3d else 2, order, cm, axis),
                                 lambda o, a=(is3d, order, cm, axis): pin_mac_mirror(o, *a)))
    for is3d in (False, True):
        for order, cm in ((1, 2), (2, 1), (2, 2)):
            for which in ((0, 1) if (order == 2 and is3d) else (0,)):
                pins.append(("mac_linear-%dd-o%d-c%d-g%d" % (3 if is3d else 2, order, cm, which),
                             lambda o, a=(is3d, order, cm, which): pin_mac_linear_field(o, *a)))
    for gt in (TYPE_MAC, TYPE_VEC3):
        for is3d in (False, True):
            for order, cm in ((1, 2), (2, 1), (2, 2)):
                for direction in range(6 if is3d else 4):
                    pins.append(("shift-t%d-%dd-o%d-c%d-d%d" % (gt, 3 if is3d else 2, order, cm, direction),
                                 lambda o, a=(gt, order, cm, is3d, direction): pin_integer_shift(o, *a)))
    for gt in (TYPE_REAL, TYPE_VEC3):
        for order, cm in ((1, 2), (2, 1), (2, 2)):
            pins.append(("centered_linear-t%d-o%d-c%d" % (gt, order, cm),
                         lambda o, a=(gt, order, cm): pin_centered_linear(o, *a)))
    for is3d in (False, True):
        pins.append(("outflow-%dd" % (3 if is3d else 2), lambda o, a=is3d: pin_outflow_column(o, a)))
    pins.append(("wall_bcs_3d", pin_wall_bcs_3d))
    return pins + high_order_pins() + gridcg_client_pins()


# ---------------------------------------------------------------------------
# higher-order lookups and traces: orderSpace 2 (Catmull-Rom cubic,
# util/interpolHigh.h) and orderTrace 2 (explicit midpoint, advection.cpp:33-37,58-73)
# ---------------------------------------------------------------------------
def _cr_weights(t):
    """Catmull-Rom weights of the points at offsets -1, 0, 1, 2 for the
    fraction t: the cardinal-spline basis with tension 0, written as weights
    (the reference writes the same segment as polynomial coefficients)."""
    t2, t3 = t * t, t * t * t
    return np.array([(-t + 2 * t2 - t3) / 2, (2 - 5 * t2 + 3 * t3) / 2, (t + 4 * t2 - 3 * t3) / 2, (-t2 + t3) / 2])


def pin_cubic_weights(o, is3d, vec, seed=31):
    """The cubic lookup equals the tensor-product Catmull-Rom weight sum over
    the 4x4(x4) cells around the point (cell centres at i + 1/2); where that
    support leaves the grid it is the trilinear lookup, exactly."""
    rng = np.random.default_rng(seed)
    shape = (11, 10, 9) if is3d else (13, 11, 1)
    nx, ny, nz = shape
    n = nx * ny * nz
    data = rng.uniform(-1, 1, (n, 3) if vec else n).astype(np.float32)
    arr = (data[:, 1] if vec else data).reshape(nz, ny, nx).astype(np.float64)
    look = (lambda p: o.interpol_cubic_component(shape, data, 1, p)) if vec else \
        (lambda p: o.interpol_cubic(shape, data, p))
    lin = (lambda p: o.lib.og_interpol_component(o.dims(shape), o._arr(data), 1, *[o.creal(x) for x in p])) if vec \
        else (lambda p: o.interpol(shape, data, p))
    for _ in range(150):
        p = rng.uniform([2.0, 2.0, 2.0], [nx - 1.5, ny - 1.5, max(nz - 1.5, 2.01)])
        if not is3d:
            p[2] = 0.5
        q = p - 0.5
        b = np.floor(q).astype(int)
        wx, wy = _cr_weights(q[0] - b[0]), _cr_weights(q[1] - b[1])
        if is3d:
            wz = _cr_weights(q[2] - b[2])
            blk = arr[b[2] - 1:b[2] + 3, b[1] - 1:b[1] + 3, b[0] - 1:b[0] + 3]
            ref = np.einsum("k,j,i,kji->", wz, wy, wx, blk)
        else:
            blk = arr[0, b[1] - 1:b[1] + 3, b[0] - 1:b[0] + 3]
            ref = np.einsum("j,i,ji->", wy, wx, blk)
        assert abs(look(p) - ref) < 2e-5, ("cubic", p, look(p), ref)
    for _ in range(60):  # supports that leave the grid: trilinear fallback
        p = rng.uniform([0.2, 0.2, 0.2], [1.4, ny - 0.2, max(nz - 0.2, 1.0)])
        if not is3d:
            p[2] = 0.5
        assert look(p) == lin(p), ("fallback", p)


def pin_cubic_quadratic(o, is3d, vec):
    """Catmull-Rom segments reproduce quadratics, so the tensor-product cubic
    lookup reproduces q(x) = prod over axes of a quadratic, exactly up to
    rounding, wherever its support is inside the grid."""
    rng = np.random.default_rng(41)
    shape = (12, 11, 10) if is3d else (14, 12, 1)
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = rng.uniform(-0.05, 0.05, (3, 3))
    qa = lambda x, a: 0.5 + c[a, 0] + c[a, 1] * x + c[a, 2] * x * x / 4
    q = lambda x, y, z: qa(x, 0) * qa(y, 1) * (qa(z, 2) if is3d else 1.0)
    vals = q(i + 0.5, j + 0.5, k + 0.5).reshape(-1)
    data = np.repeat(vals[:, None], 3, 1).astype(np.float32) if vec else vals.astype(np.float32)
    for _ in range(120):
        p = rng.uniform([2.0, 2.0, 2.0], [nx - 1.5, ny - 1.5, max(nz - 1.5, 2.01)])
        if not is3d:
            p[2] = 0.5
        got = o.interpol_cubic_component(shape, data, 2, p) if vec else o.interpol_cubic(shape, data, p)
        assert abs(got - q(*p)) < 3e-6, (p, got, q(*p))


def _linear_vel_mac(shape, a0, G):
    return _linear_mac(shape, a0, G).astype(np.float32).reshape(-1, 3)


def pin_midpoint_centered(o, gridType, orderSpace, is3d):
    """orderTrace 2 on Real / centred Vec3 grids with a linear velocity u and a
    linear source f: the cell value becomes f(p0 - u(p0 - u(p0) dt/2) dt),
    p0 the cell centre (getCentered and interpolMAC are exact on linear
    fields, and so are the trilinear and cubic lookups of f)."""
    shape = (16, 15, 14) if is3d else (18, 17, 1)
    nx, ny, nz = shape
    n = nx * ny * nz
    a0 = np.array([0.15, -0.1, 0.12 if is3d else 0.0])
    G = np.array([[0.02, -0.015, 0.01], [0.01, 0.02, -0.02], [-0.01, 0.015, 0.02]])
    if not is3d:
        G[2, :] = 0
        G[:, 2] = 0
    vel = _linear_vel_mac(shape, a0, G)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P0 = np.stack([i + 0.5, j + 0.5, (k + 0.5) if is3d else np.full(i.shape, 0.5)], -1)
    cf = np.array([[0.3, 0.04, -0.03, 0.02], [-0.2, 0.01, 0.03, -0.01], [0.1, -0.02, 0.01, 0.03]])
    nc = 1 if gridType == TYPE_REAL else 3
    f = lambda P, cc: cf[cc, 0] + P[..., 0] * cf[cc, 1] + P[..., 1] * cf[cc, 2] + (P[..., 2] * cf[cc, 3] if is3d else 0)
    src = np.stack([f(P0, cc) for cc in range(nc)], -1).astype(np.float32).reshape(n, nc)
    g = src.reshape(-1).copy() if nc == 1 else src.copy()
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    dt = 0.9
    o.advect_hi(shape, np.float32(dt), flags, vel, g, gridType, 1, 1.0, 2, orderSpace, 2)
    u = lambda P: a0 + P @ G.T
    P2 = P0 - u(P0 - u(P0) * dt / 2) * dt
    exp = np.stack([f(P2, cc) for cc in range(nc)], -1)
    m = 3
    sl = (slice(m, nz - m) if is3d else slice(None), slice(m, ny - m), slice(m, nx - m))
    diff = np.abs(g.reshape(nz, ny, nx, nc)[sl] - exp[sl]).max()
    assert diff < 2e-5, ("midpoint centred", gridType, orderSpace, diff)


def pin_midpoint_mac(o, orderSpace, is3d, self_adv):
    """orderTrace 2 on MAC grids, velocity u and advected MAC field s both
    linear (s = u for self-advection).  The midpoint trace of face c at f0
    uses the advected field (advection.cpp:59-73): q = p0 - s(f0 - s(f0) dt/2) dt
    with p0 the cell centre; the value is s_c(q - e_c/2) with the trilinear
    lookup and s_c(q) with the cubic one (interpolCubicMAC looks up at
    q + e_c/2 in cell-centred coordinates, interpolHigh.h:172-179); in 2D the
    cubic z component is 0."""
    shape = (16, 15, 14) if is3d else (18, 17, 1)
    nx, ny, nz = shape
    aU = np.array([0.1, -0.12, 0.08 if is3d else 0.0])
    GU = np.array([[0.015, -0.02, 0.01], [0.02, 0.01, -0.015], [-0.01, 0.02, 0.015]])
    aS = np.array([-0.15, 0.1, -0.05 if is3d else 0.0])
    GS = np.array([[-0.02, 0.01, 0.015], [0.015, -0.02, 0.01], [0.01, -0.015, -0.02]])
    if self_adv:
        aS, GS = aU, GU
    for Gm in (GU, GS):
        if not is3d:
            Gm[2, :] = 0
            Gm[:, 2] = 0
    vel = _linear_vel_mac(shape, aU, GU)
    g = _linear_vel_mac(shape, aS, GS).copy()
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    dt = 0.9
    o.advect_hi(shape, np.float32(dt), flags, g.copy() if self_adv else vel, g, TYPE_MAC, 1, 1.0, 2, orderSpace, 2)
    sfun = lambda P: aS + P @ GS.T
    pos = _face_positions(shape)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P0 = np.stack([i + 0.5, j + 0.5, (k + 0.5) if is3d else np.full(i.shape, 0.5)], -1)
    exp = np.zeros(pos[0].shape)
    for c in range(3):
        f0 = pos[c]
        q = P0 - sfun(f0 - sfun(f0) * dt / 2) * dt
        if orderSpace == 1:
            q = q.copy()
            q[..., c] -= 0.5
        exp[..., c] = aS[c] + q @ GS[c]
    if not is3d and orderSpace == 2:
        exp[..., 2] = 0
    got = g.reshape(nz, ny, nx, 3)
    m = 3
    sl = (slice(m, nz - m) if is3d else slice(None), slice(m, ny - m), slice(m, nx - m))
    comps = 3 if is3d else 2
    diff = np.abs(got[sl][..., :comps] - exp[sl][..., :comps]).max()
    assert diff < 2e-5, ("midpoint MAC", orderSpace, self_adv, diff)


def pin_high_order_maccormack(o, gridType, orderSpace, orderTrace):
    """MacCormack with the higher-order lookups and traces keeps its defining
    property: a linear field under a uniform velocity is translated exactly
    (backward trace returns the original, zero correction, forward value kept
    by the clamp) for both clamp modes.  MAC grids with the cubic lookup are
    translated from the half-cell-shifted point of interpolCubicMAC."""
    shape = (16, 15, 14)
    nx, ny, nz = shape
    n = nx * ny * nz
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    nc = 1 if gridType == TYPE_REAL else 3
    u0 = np.array([0.37, -0.61, 0.29])
    if gridType == TYPE_MAC:
        G = np.zeros((3, 3))
        src = _linear_vel_mac(shape, np.array([0.2, -0.1, 0.3]), np.array([[0, 0.02, -0.01], [0.01, 0, 0.02], [-0.02, 0.01, 0]]))
        pos = _face_positions(shape)
        Gs = np.array([[0, 0.02, -0.01], [0.01, 0, 0.02], [-0.02, 0.01, 0]])
        a0 = np.array([0.2, -0.1, 0.3])
        exp = np.stack([a0[c] + (pos[c] - u0) @ Gs[c] for c in range(3)], -1)
        if orderSpace == 2:
            exp = np.stack([a0[c] + (pos[c] + 0.5 * np.eye(3)[c] - u0) @ Gs[c] for c in range(3)], -1)
        vel = np.tile(u0.astype(np.float32), (n, 1))
        g = src.copy()
    else:
        X = np.stack([i + 0.5, j + 0.5, k + 0.5], -1)
        cf = np.array([[0.2, 0.03, -0.02, 0.01], [-0.1, 0.01, 0.02, -0.03], [0.3, -0.02, 0.01, 0.02]])
        f = lambda P, cc: cf[cc, 0] + P @ cf[cc, 1:]
        src = np.stack([f(X, cc) for cc in range(nc)], -1).astype(np.float32).reshape(n, nc)
        exp = np.stack([f(X - u0, cc) for cc in range(nc)], -1)
        vel = np.tile(u0.astype(np.float32), (n, 1))
        g = src.reshape(-1).copy() if nc == 1 else src.copy()
    for cm in (1, 2):
        gg = g.copy()
        o.advect_hi(shape, np.float32(1.0), flags, vel, gg, gridType, 2, 1.0, cm, orderSpace, orderTrace)
        got = gg.reshape(nz, ny, nx, -1)
        sl = (slice(4, -4),) * 3
        diff = np.abs(got[sl] - exp.reshape(nz, ny, nx, -1)[sl]).max()
        assert diff < 3e-5, ("high-order MacCormack", gridType, orderSpace, orderTrace, cm, diff)


def high_order_pins():
    pins = []
    for is3d in (False, True):
        for vec in (False, True):
            pins.append(("cubic_weights-%dd-v%d" % (3 if is3d else 2, vec), lambda o, a=(is3d, vec): pin_cubic_weights(o, *a)))
            pins.append(("cubic_quadratic-%dd-v%d" % (3 if is3d else 2, vec), lambda o, a=(is3d, vec): pin_cubic_quadratic(o, *a)))
        for os_ in (1, 2):
            for gt in (TYPE_REAL, TYPE_VEC3):
                pins.append(("midpoint_centered-%dd-t%d-s%d" % (3 if is3d else 2, gt, os_),
                             lambda o, a=(gt, os_, is3d): pin_midpoint_centered(o, *a)))
            for self_adv in (False, True):
                pins.append(("midpoint_mac-%dd-s%d-self%d" % (3 if is3d else 2, os_, self_adv),
                             lambda o, a=(os_, is3d, self_adv): pin_midpoint_mac(o, *a)))
    for gt in (TYPE_REAL, TYPE_VEC3, TYPE_MAC):
        for os_, ot in ((2, 1), (1, 2), (2, 2)):
            if gt == TYPE_MAC and ot == 2:
                continue  # the MAC midpoint trace follows the advected field, not vel (pin_midpoint_mac)
            pins.append(("hi_maccormack-t%d-s%d-tr%d" % (gt, os_, ot),
                         lambda o, a=(gt, os_, ot): pin_high_order_maccormack(o, *a)))
    return pins


# ---------------------------------------------------------------------------
# other GridCg clients: cgSolveDiffusion (conjugategrad.cpp:347-423) and
# cgSolveWE (plugin/waves.cpp:69-141), pinned by eigenvector closed forms of
# the operators they solve and a closed form of the Crank-Nicolson RHS
# ---------------------------------------------------------------------------
def _interior_modes(shape, modes, kind):
    """Product of 1-D eigenvectors on the cells 1..m of each axis (0 on the
    outer layer) and the matching eigenvalue of the 5/7-point Laplacian:
    kind 'dirichlet' (neighbours outside the interior held at 0): sin(pi a i/(m+1)),
    2 - 2 cos(pi a/(m+1)); kind 'neumann' (those neighbours dropped):
    cos(pi a (i - 1/2)/m), 2 - 2 cos(pi a/m)."""
    nx, ny, nz = shape
    dims = [nx, ny] + ([nz] if nz > 1 else [])
    field = np.ones((nz, ny, nx), np.float64)
    lam = 0.0
    for ax, (n, a) in enumerate(zip(dims, modes)):
        m = n - 2
        i = np.arange(n, dtype=np.float64)
        if kind == "dirichlet":
            e = np.sin(np.pi * a * i / (m + 1))
            lam += 2 - 2 * np.cos(np.pi * a / (m + 1))
        else:
            e = np.cos(np.pi * a * (i - 0.5) / m)
            lam += 2 - 2 * np.cos(np.pi * a / m)
        e[0] = e[-1] = 0.0
        shp = [1, 1, 1]
        shp[2 - ax] = n
        field = field * e.reshape(shp)
    return field, lam


def pin_diffusion_eigen(o, is3d, vec):
    """Walls on every side, fluid inside: cgSolveDiffusion solves
    (I + alpha L) x = b on the interior with L the Dirichlet Laplacian (the
    wall cells are identity rows holding 0), so a Dirichlet eigenvector b of
    eigenvalue lam gives x = b / (1 + alpha lam).  Vec3/MAC grids: each
    solved component (two in 2D) with its own mode; the z component of a 2D
    grid is left alone."""
    shape = (14, 12, 10) if is3d else (18, 14, 1)
    nx, ny, nz = shape
    n = nx * ny * nz
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    alpha = 0.3
    modes = [(1, 2, 1), (2, 1, 3), (3, 2, 2)]
    if not vec:
        v, lam = _interior_modes(shape, modes[0], "dirichlet")
        g = v.reshape(-1).astype(np.float32)
        o.cg_solve_diffusion(shape, flags, g, alpha=alpha, cgMaxIterFac=4.0, cgAccuracy=1e-7)
        exp = v.reshape(-1) / (1 + alpha * lam)
        assert np.abs(g - exp).max() < 2e-5, ("diffusion real", is3d, np.abs(g - exp).max())
        return
    g = np.zeros((n, 3), np.float32)
    exps = []
    for c in range(3):
        v, lam = _interior_modes(shape, modes[c], "dirichlet")
        g[:, c] = (c + 1) * v.reshape(-1)
        exps.append((c + 1) * v.reshape(-1) / (1 + alpha * lam))
    g0 = g.copy()
    o.cg_solve_diffusion(shape, flags, g, alpha=alpha, cgMaxIterFac=4.0, cgAccuracy=1e-7)
    for c in range(3 if is3d else 2):
        d = np.abs(g[:, c] - exps[c]).max()
        assert d < 5e-5, ("diffusion vec component", is3d, c, d)
    if not is3d:
        assert np.array_equal(g[:, 2], g0[:, 2]), "2D diffusion touched the z component"


def pin_diffusion_obstacle_rows(o):
    """Obstacle rows of the diffusion matrix are identity (A0 = 1, no
    off-diagonals) whatever alpha is; every other interior row has diagonal
    1 + alpha * (2 dim) and off-diagonals -alpha toward every upper
    neighbour (the matrix is built on an all-fluid grid); the outer layer's
    non-obstacle rows are 1 on the diagonal."""
    shape = (9, 8, 7)
    rng = np.random.default_rng(5)
    flags = make_flags(o, shape, bw=1, obstacles=3, rng=rng)
    A0, Ai, Aj, Ak = o.diffusion_matrix(shape, flags, 0.4)
    nx, ny, nz = shape
    f = flags.reshape(nz, ny, nx)
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    obs = (f & OBSTACLE) != 0
    A0, Ai, Aj, Ak = [a.reshape(nz, ny, nx) for a in (A0, Ai, Aj, Ak)]
    assert np.all(A0[obs] == 1) and np.all(Ai[obs] == 0) and np.all(Aj[obs] == 0) and np.all(Ak[obs] == 0)
    m = inner & ~obs
    assert np.allclose(A0[m], 1 + 0.4 * 6) and np.all(Ai[m] == np.float32(-0.4)) and np.all(Ak[m] == np.float32(-0.4))
    outer = ~inner & ~obs
    assert np.all(A0[outer] == 1) and np.all(Ai[outer] == 0)


def pin_we_eigen(o, is3d):
    """Walls on every side, fluid inside: cgSolveWE solves (I + s L) x = b
    with s = dt^2 cSqr / 2 and L the Neumann Laplacian of the interior (the
    pressure matrix: wall neighbours are not counted), b = 2 ut - utm1.  A
    Neumann eigenvector gives out = b / (1 + s lam); afterwards ut = out and
    utm1 = the old ut."""
    shape = (12, 11, 10) if is3d else (16, 13, 1)
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    v, lam = _interior_modes(shape, (2, 1, 1), "neumann")
    ut = v.reshape(-1).astype(np.float32)
    utm1 = (0.3 * v.reshape(-1)).astype(np.float32)
    ut0 = ut.copy()
    out = np.zeros_like(ut)
    dt, cSqr = 1.5, 0.25
    s = np.float32(np.float32(dt * dt) * np.float32(cSqr) * 0.5)
    o.cg_solve_we(shape, flags, ut, utm1, out, dt=dt, cSqr=cSqr, cgMaxIterFac=4.0, cgAccuracy=1e-7)
    exp = (2 * ut0.astype(np.float64) - 0.3 * v.reshape(-1)) / (1 + float(s) * lam)
    assert np.abs(out - exp).max() < 2e-5, ("wave solve", is3d, np.abs(out - exp).max())
    assert np.array_equal(ut, out) and np.array_equal(utm1, ut0)


def pin_we_rhs_crank_nicolson(o, is3d):
    """MakeRhsWE with crankNic: for a quadratic ut = a x^2 + b y^2 + c x y (+ e z^2)
    the 5-point x-y stencil -4 ut + ut(i+-1) + ut(j+-1) is 2a + 2b at every
    interior cell, so rhs = 2 ut - utm1 + s (2a + 2b); the stencil has no z
    terms, so e drops out in 3D.  The outer layer stays 0."""
    shape = (11, 9, 7) if is3d else (13, 10, 1)
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a, b, c, e = 0.011, 0.007, 0.003, 0.02
    ut = (a * x * x + b * y * y + c * x * y + e * z * z).astype(np.float32).reshape(-1)
    utm1 = (0.5 * ut).astype(np.float32)
    flags = o.init_domain(shape)
    o.fill_grid(shape, flags)
    s = 0.37
    rhs = o.make_rhs_we(shape, flags, ut, utm1, np.float32(s), True).reshape(nz, ny, nx)
    exp = (2.0 * ut - utm1).reshape(nz, ny, nx) + np.float32(s) * (2 * a + 2 * b)
    inner = (slice(1, -1) if is3d else slice(None), slice(1, -1), slice(1, -1))
    assert np.abs(rhs[inner] - exp[inner]).max() < 1e-5, ("crankNic rhs", np.abs(rhs[inner] - exp[inner]).max())
    mask = np.ones_like(rhs, bool)
    mask[inner] = False
    assert np.all(rhs[mask] == 0)
    plain = o.make_rhs_we(shape, flags, ut, utm1, np.float32(s), False).reshape(nz, ny, nx)
    assert np.abs(plain[inner] - (2.0 * ut - utm1).reshape(nz, ny, nx)[inner]).max() < 1e-5


def gridcg_client_pins():
    pins = []
    for is3d in (False, True):
        for vec in (False, True):
            pins.append(("diffusion_eigen-%dd-v%d" % (3 if is3d else 2, vec),
                         lambda o, a=(is3d, vec): pin_diffusion_eigen(o, *a)))
        pins.append(("we_eigen-%dd" % (3 if is3d else 2), lambda o, a=is3d: pin_we_eigen(o, a)))
        pins.append(("we_rhs_cn-%dd" % (3 if is3d else 2), lambda o, a=is3d: pin_we_rhs_crank_nicolson(o, a)))
    pins.append(("diffusion_rows", pin_diffusion_obstacle_rows))
    return pins
