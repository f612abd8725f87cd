"""Batched pressure solves (uneven 2D and 3D problems with 0-20 % obstacles
and seeded divergent velocities, the batched-solve workload at small sizes):
every problem's result from one batched solve is bit-identical to solving it
alone, because the batch runs the same kernel bodies and reduces each
problem's partial sums in the same order."""
import numpy as np
import pytest

from helpers import GpuScene, obstacle_flags, smooth_vel

pytestmark = pytest.mark.gpu

PROBLEMS = [((20, 20, 20), 0.0), ((33, 27, 40), 0.1), ((48, 48, 48), 0.2),
            ((40, 40, 1), 0.05), ((96, 70, 1), 0.15), ((130, 130, 1), 0.2), ((9, 9, 9), 0.0)]


def make_problem(manta, oracle, shape, frac, seed):
    rng = np.random.default_rng(seed)
    flags = obstacle_flags(oracle, shape, frac, rng)
    vel = smooth_vel(rng, shape, passes=3)
    oracle.set_wall_bcs(shape, flags, vel)
    g = GpuScene(manta, shape)
    return g, g.flags(flags), g.mac(vel), g.real()


@pytest.mark.parametrize("pc", ["PcMGStatic", "PcMIC"])
def test_batch_equals_single(manta, oracle, pc):
    kw = dict(preconditioner=getattr(manta, pc), cgAccuracy=1e-4)
    single, batch = [], []
    for k, (shape, frac) in enumerate(PROBLEMS):
        single.append(make_problem(manta, oracle, shape, frac, 100 + k))
        batch.append(make_problem(manta, oracle, shape, frac, 100 + k))
    for g, f, v, p in single:
        manta.solvePressure(flags=f, vel=v, pressure=p, **kw)
    manta.solvePressureBatch([(v, p, f) for g, f, v, p in batch], **kw)
    for (g1, f1, v1, p1), (g2, f2, v2, p2), (shape, _) in zip(single, batch, PROBLEMS):
        s1, s2 = manta.lastSolveStats(g1.s), manta.lastSolveStats(g2.s)
        print(shape, pc, "iterations", s1["iterations"], s2["iterations"], "rel.res", s1["relResidual"])
        assert s1["converged"] and s2["converged"]
        assert s1["iterations"] == s2["iterations"]
        np.testing.assert_array_equal(p1.toNumpy(), p2.toNumpy())
        np.testing.assert_array_equal(v1.toNumpy(), v2.toNumpy())


def test_batch_repeated_direct_capture_replay(manta, oracle):
    """The same batch solved twelve times in a row: the chunk autotuner's
    eight trial solves (four plane-chunk policies, queued launches), then the
    chosen policy's queued launches, the capture of its two-iteration graph
    and the replays.  Every pass starts from the same velocity and must give
    results bit-identical to single solves."""
    kw = dict(preconditioner=manta.PcMGStatic, cgAccuracy=1e-4)
    single, batch, vel0 = [], [], []
    for k, (shape, frac) in enumerate(PROBLEMS):
        single.append(make_problem(manta, oracle, shape, frac, 200 + k))
        g, f, v, p = make_problem(manta, oracle, shape, frac, 200 + k)
        batch.append((g, f, v, p))
        vel0.append(v.toNumpy().copy())
    for g, f, v, p in single:
        manta.solvePressure(flags=f, vel=v, pressure=p, **kw)
    for rep in range(12):
        for (g, f, v, p), v0 in zip(batch, vel0):
            v.fromNumpy(v0)
        manta.solvePressureBatch([(v, p, f) for g, f, v, p in batch], **kw)
        for (g1, f1, v1, p1), (g2, f2, v2, p2) in zip(single, batch):
            s1, s2 = manta.lastSolveStats(g1.s), manta.lastSolveStats(g2.s)
            assert s2["converged"] and s1["iterations"] == s2["iterations"], rep
            np.testing.assert_array_equal(p1.toNumpy(), p2.toNumpy())
            np.testing.assert_array_equal(v1.toNumpy(), v2.toNumpy())


def test_batch_rejects_shared_solver(manta, oracle):
    g, f, v, p = make_problem(manta, oracle, (16, 16, 16), 0.0, 1)
    with pytest.raises(RuntimeError, match="own solver"):
        manta.solvePressureBatch([(v, p, f), (v, p, f)])
