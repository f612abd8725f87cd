"""Parity of advectSemiLagrange with orderSpace 2 (Catmull-Rom cubic lookups,
util/interpolHigh.h) and orderTrace 2 (explicit midpoint trace,
advection.cpp:33-37,58-73) against the oracle, bit for bit: Real, Vec3 and
MAC grids, first order and MacCormack with both clamp modes, 2D and 3D, with
obstacles and an open top (MAC outflow extrapolation)."""
import numpy as np
import pytest

from oracle import TYPE_MAC, TYPE_REAL, TYPE_VEC3
from tests.helpers import GpuScene, assert_bitexact, make_flags, rand_vel

pytestmark = pytest.mark.gpu

SHAPES = [(26, 22, 1), (17, 13, 11), (20, 24, 18)]
ORDERS = [(1, 2), (2, 1), (2, 2)]
HI = [(2, 1), (1, 2), (2, 2)]


def _inputs(oracle, rng, shape, open_bound):
    n = int(np.prod(shape))
    flags = make_flags(oracle, shape, bw=1, open_bound=open_bound, obstacles=3, rng=rng)
    if open_bound:
        oracle.reset_outflow(shape, flags)
    vel = rand_vel(rng, n, 1.6)
    if shape[2] == 1:
        vel[:, 2] = 0
    return flags, vel


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("order,clamp", ORDERS)
@pytest.mark.parametrize("space,trace", HI)
def test_advect_real_hi(manta, oracle, rng, shape, order, clamp, space, trace):
    n = int(np.prod(shape))
    flags, vel = _inputs(oracle, rng, shape, "Y")
    dens = rng.uniform(0, 1, n).astype(np.float32)
    dt = 1.2
    sc = GpuScene(manta, shape, dt)
    f, v, d = sc.flags(flags), sc.mac(vel), sc.real(dens)
    ref = dens.copy()
    oracle.advect_hi(shape, np.float32(dt), flags, vel, ref, TYPE_REAL, order, 1.0, clamp, space, trace)
    manta.advectSemiLagrange(flags=f, vel=v, grid=d, order=order, clampMode=clamp, orderSpace=space,
                             orderTrace=trace)
    assert_bitexact(d.toNumpy(), ref, "advect real orderSpace %d orderTrace %d" % (space, trace))


@pytest.mark.parametrize("shape", SHAPES[:2])
@pytest.mark.parametrize("order,clamp", ORDERS)
@pytest.mark.parametrize("space,trace", HI)
def test_advect_vec3_hi(manta, oracle, rng, shape, order, clamp, space, trace):
    n = int(np.prod(shape))
    flags, vel = _inputs(oracle, rng, shape, "")
    g = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    dt = 0.9
    sc = GpuScene(manta, shape, dt)
    f, v, gg = sc.flags(flags), sc.mac(vel), sc.vec(g)
    ref = g.copy()
    oracle.advect_hi(shape, np.float32(dt), flags, vel, ref, TYPE_VEC3, order, 0.7, clamp, space, trace)
    manta.advectSemiLagrange(flags=f, vel=v, grid=gg, order=order, strength=0.7, clampMode=clamp,
                             orderSpace=space, orderTrace=trace)
    assert_bitexact(gg.toNumpy(), ref, "advect vec3 orderSpace %d orderTrace %d" % (space, trace))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("order,clamp", ORDERS)
@pytest.mark.parametrize("space,trace", HI)
@pytest.mark.parametrize("open_bound", ["", "Y"])
def test_advect_mac_self_hi(manta, oracle, rng, shape, order, clamp, space, trace, open_bound):
    flags, vel = _inputs(oracle, rng, shape, open_bound)
    dt = 1.2
    sc = GpuScene(manta, shape, dt)
    f, v = sc.flags(flags), sc.mac(vel)
    ref = vel.copy()
    oracle.advect_hi(shape, np.float32(dt), flags, ref, ref, TYPE_MAC, order, 1.0, clamp, space, trace)
    manta.setDebugLevel(0)
    try:
        manta.advectSemiLagrange(flags=f, vel=v, grid=v, order=order, strength=1.0, clampMode=clamp,
                                 orderSpace=space, orderTrace=trace)
    finally:
        manta.setDebugLevel(1)
    assert_bitexact(v.toNumpy(), ref, "advect mac orderSpace %d orderTrace %d" % (space, trace))


def test_advect_mac_other_grid_midpoint(manta, oracle, rng):
    """A MAC grid advected by a different velocity with orderTrace 2: the
    midpoint trace takes its velocities from the advected grid itself, not
    from vel (advection.cpp:59-73)."""
    shape = (19, 16, 14)
    n = int(np.prod(shape))
    flags, vel = _inputs(oracle, rng, shape, "")
    other = rand_vel(rng, n, 1.1)
    sc = GpuScene(manta, shape, 1.0)
    f, v, o = sc.flags(flags), sc.mac(vel), sc.mac(other)
    ref = other.copy()
    oracle.advect_hi(shape, np.float32(1.0), flags, vel, ref, TYPE_MAC, 2, 1.0, 2, 1, 2)
    manta.advectSemiLagrange(flags=f, vel=v, grid=o, order=2, orderTrace=2)
    assert_bitexact(o.toNumpy(), ref, "advect other mac grid, midpoint trace")


def test_advect_unknown_orders_raise(manta, rng):
    shape = (12, 10, 8)
    sc = GpuScene(manta, shape, 1.0)
    n = int(np.prod(shape))
    f = sc.s.create(manta.FlagGrid)
    f.initDomain(boundaryWidth=1)
    f.fillGrid()
    v, d = sc.mac(rand_vel(rng, n, 0.5)), sc.real(np.ones(n, np.float32))
    with pytest.raises(RuntimeError, match="Unknown interpolation order 3"):
        manta.advectSemiLagrange(flags=f, vel=v, grid=d, orderSpace=3)
    with pytest.raises(RuntimeError, match="Unknown backtracing order 0"):
        manta.advectSemiLagrange(flags=f, vel=v, grid=d, orderTrace=0)
