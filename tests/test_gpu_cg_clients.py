"""The other GridCg clients on the GPU against the oracle: cgSolveDiffusion
(conjugategrad.cpp:347-423) on Real, Vec3 and MAC grids and cgSolveWE
(plugin/waves.cpp:87-141) with and without the Crank-Nicolson RHS, 2D and 3D,
with obstacles and empty cells.  The CG runs the same iteration as the
oracle; reductions are summed in another order, so iteration counts agree
within one and solutions within the CG tolerance (stated per test)."""
import numpy as np
import pytest

from oracle import EMPTY, FLUID, OBSTACLE
from tests.helpers import GpuScene, make_flags

pytestmark = pytest.mark.gpu

SHAPES = [(40, 34, 1), (22, 20, 18)]


def _flags(oracle, rng, shape, empty):
    flags = make_flags(oracle, shape, bw=1, obstacles=3, rng=rng)
    if empty:
        nx, ny, nz = shape
        g = flags.reshape(nz, ny, nx)
        band = g[..., ny - 5:ny - 1, :] if nz == 1 else g[1:-1, ny - 5:ny - 1, 1:-1]
        band[(band & FLUID) != 0] = EMPTY
    return flags


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["real", "vec", "mac"])
@pytest.mark.parametrize("empty", [False, True])
def test_cg_solve_diffusion(manta, oracle, rng, shape, kind, empty):
    n = int(np.prod(shape))
    flags = _flags(oracle, rng, shape, empty)
    acc = 1e-5
    sc = GpuScene(manta, shape)
    f = sc.flags(flags)
    if kind == "real":
        val = rng.uniform(-1, 1, n).astype(np.float32)
        g = sc.real(val)
    else:
        val = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        g = sc.vec(val) if kind == "vec" else sc.mac(val)
    ref = val.copy()
    ost = oracle.cg_solve_diffusion(shape, flags, ref, alpha=0.35, cgMaxIterFac=2.0, cgAccuracy=acc)
    manta.cgSolveDiffusion(flags=f, grid=g, alpha=0.35, cgMaxIterFac=2.0, cgAccuracy=acc)
    st = sc.s._solve_stats
    got = g.toNumpy()
    assert st["converged"] and not ost["diverged"]
    assert abs(st["iterations"] - ost["iterations"]) <= 1, (st["iterations"], ost["iterations"])
    # ||r||_inf < acc and A >= I on the fluid cells: |x - x_ref| <= 2 acc
    assert np.abs(got - ref).max() <= 4 * acc, np.abs(got - ref).max()
    if kind != "real" and shape[2] == 1:
        assert np.array_equal(got[:, 2], val[:, 2]), "2D diffusion touched the z component"


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("crank", [False, True])
def test_cg_solve_we(manta, oracle, rng, shape, crank):
    n = int(np.prod(shape))
    flags = _flags(oracle, rng, shape, False)
    ut0 = rng.uniform(-1, 1, n).astype(np.float32)
    utm0 = rng.uniform(-1, 1, n).astype(np.float32)
    dt, acc = 1.3, 1e-6
    sc = GpuScene(manta, shape, dt)
    f, ut, utm1, out = sc.flags(flags), sc.real(ut0), sc.real(utm0), sc.real()
    r_ut, r_utm1, r_out = ut0.copy(), utm0.copy(), np.zeros(n, np.float32)
    ost = oracle.cg_solve_we(shape, flags, r_ut, r_utm1, r_out, dt=dt, crankNic=crank, cSqr=0.3,
                             cgMaxIterFac=2.0, cgAccuracy=acc)
    manta.setDebugLevel(0)
    try:
        manta.cgSolveWE(flags=f, ut=ut, utm1=utm1, out=out, crankNic=crank, cSqr=0.3, cgMaxIterFac=2.0,
                        cgAccuracy=acc)
    finally:
        manta.setDebugLevel(1)
    st = sc.s._solve_stats
    assert st["converged"] and not ost["diverged"]
    assert abs(st["iterations"] - ost["iterations"]) <= 1, (st["iterations"], ost["iterations"])
    assert np.abs(out.toNumpy() - r_out).max() <= 4 * acc
    assert np.array_equal(ut.toNumpy(), out.toNumpy())
    assert np.array_equal(utm1.toNumpy(), ut0)
