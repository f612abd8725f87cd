"""The batched-solve and serving workloads at their full sizes, on the GPU,
with per-problem / per-request checks that do not depend on the GPU path:

- batched solves: the 64 problems of workloads.batched_problems() (32 2D
  grids with sides 64-512, 32 3D grids with sides 32-160, 0-20 % obstacles,
  54 M cells).  Every problem converges, its FP64 relative residual
  ||b - A x||_2 / ||b||_2, recomputed on the CPU with the oracle's
  double-precision build from the downloaded pressure, is <= 1e-4, the
  corrected velocity is divergence-free to the same tolerance, the batched
  result equals the single solve bit for bit, and the smallest 2D and 3D
  problems agree with the oracle's own CG solve of the same system;
- serving: the 24 sessions of serve.make_sessions(24) (64^2 to 192^3 cells,
  log-uniform, 2D and 3D) under a stream of step and solve requests: the
  served (continuously batched, four workers) replay leaves every session
  bit-identical to the serial replay, every request's projection meets
  relative residual 1e-4 (FP64, on the device), and each session's last
  projection is checked on the CPU from the downloaded state."""
import numpy as np
import pytest

from oracle import OUTFLOW, PC_NONE

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _fp64_relres(o64, shape, flags, vel, p):
    """||b - A p||_2 / ||b||_2 in FP64 from the oracle's double build, with
    b = rhs(vel before the correction)."""
    v = vel.astype(np.float64)
    b, _, _ = o64.make_rhs(shape, flags, v)
    A = o64.make_laplace(shape, flags)
    r = b - o64.apply_matrix(shape, flags, p.astype(np.float64), A)
    nb = float(np.sqrt((b * b).sum()))
    return float(np.sqrt((r * r).sum())) / (nb if nb > 0 else 1.0), nb


def _outflow_adjacent(flags, shape):
    nx, ny, nz = shape
    out = (flags.reshape(nz, ny, nx) & OUTFLOW) != 0
    m = np.zeros_like(out)
    m[:, :, :-1] |= out[:, :, 1:]
    m[:, :-1, :] |= out[:, 1:, :]
    if nz > 1:
        m[:-1, :, :] |= out[1:, :, :]
    return m.reshape(-1)


def test_batched_problems_full_size(manta, oracle, oracle64):
    from mantaflow_amd import workloads
    kw = dict(cgAccuracy=TOL, preconditioner=manta.PcMGStatic)
    batch = workloads.batched_problems()
    single = workloads.batched_problems()
    assert len(batch) == 64 and sum(p.cells for p in batch) > 50e6
    manta.solvePressureBatch([p.problem() for p in batch], **kw)
    for p in single:
        manta.solvePressure(flags=p.flags, vel=p.vel, pressure=p.pressure, **kw)
    small = sorted(range(64), key=lambda i: batch[i].cells)
    check_oracle = {small[0]} | set([i for i in small if batch[i].shape[2] > 1][:1])
    worst = 0.0
    for i, (pb, ps) in enumerate(zip(batch, single)):
        sb, ss = manta.lastSolveStats(pb.s), manta.lastSolveStats(ps.s)
        assert sb["converged"] and sb["relResidual"] <= TOL, (i, pb.shape, sb)
        assert sb["iterations"] == ss["iterations"], i
        gp, gv = pb.pressure.toNumpy(), pb.vel.toNumpy()
        np.testing.assert_array_equal(gp, ps.pressure.toNumpy(), err_msg="pressure %d" % i)
        np.testing.assert_array_equal(gv, ps.vel.toNumpy(), err_msg="velocity %d" % i)
        flags = pb.flags.toNumpy()
        v0 = pb.vel0.copy()
        oracle.set_wall_bcs(pb.shape, flags, v0)   # the velocity the solve started from
        rel, nb = _fp64_relres(oracle64, pb.shape, flags, v0, gp)
        assert rel <= 1.05 * TOL, (i, pb.shape, rel)
        worst = max(worst, rel)
        # divergence after the correction equals -(b - A p): the same bound
        d, _, _ = oracle64.make_rhs(pb.shape, flags, gv.astype(np.float64))
        assert float(np.sqrt((d * d).sum())) <= 1.05 * TOL * nb, i
        if i in check_oracle:
            ref_v, ref_p = v0.copy(), np.zeros(pb.cells, np.float32)
            # plain CG: robust on obstacle-enclosed pockets (singular, compatible)
            st = oracle.solve_pressure(pb.shape, ref_v, ref_p, flags, cgAccuracy=TOL, preconditioner=PC_NONE,
                                       relTol=TOL, maxIterOverride=4000)
            assert st["relResidual"] <= TOL
            scale = max(float(np.abs(ref_v).max()), 1e-6)
            assert float(np.abs(gv - ref_v).max()) <= 2e-2 * scale, (i, pb.shape)
    print("64 batched problems: worst FP64 relative residual (CPU recomputation) %.3g" % worst)


def test_served_sessions_full_size(manta, oracle64):
    from mantaflow_amd import serve
    shapes = serve.make_sessions(24)
    assert max(a * b * c for a, b, c in shapes) > 100 ** 3
    a = [serve.PlumeScene(sh, seed=20 + i) for i, sh in enumerate(shapes)]
    b = [serve.PlumeScene(sh, seed=20 + i) for i, sh in enumerate(shapes)]
    trace = serve.make_trace(72, len(shapes), rate=1e9, seed=8)
    ser = serve.Server(a, max_batch=1, workers=1)
    ser.run(trace, realtime=False)
    srv = serve.Server(b, workers=4)
    lat, _ = srv.run(trace, realtime=False)
    assert sorted(lat) == sorted(r.rid for r in trace) and max(srv.batch_sizes) > 1
    n_req = 0
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.steps == y.steps
        for g in ("density", "vel", "pressure"):
            np.testing.assert_array_equal(getattr(x, g).toNumpy(), getattr(y, g).toNumpy(),
                                          err_msg="session %d %s" % (i, g))
        for res in (ser.results, srv.results):
            for r in res:
                if r.session == i:
                    assert r.status == 0 and r.converged and r.relResidual <= TOL, (i, r.relResidual)
                    n_req += 1
        if any(r.session == i for r in srv.results):
            # the last projection, checked on the CPU: undo the correction to
            # recover its right-hand side, b = rhs(u) + A p (correctVelocity's
            # identity, away from cells whose upper face borders an outflow cell)
            sh = y.shape
            flags, u, p = y.flags.toNumpy(), y.vel.toNumpy().astype(np.float64), y.pressure.toNumpy()
            r, _, _ = oracle64.make_rhs(sh, flags, u)
            A = oracle64.make_laplace(sh, flags)
            bvec = r + oracle64.apply_matrix(sh, flags, p.astype(np.float64), A)
            keep = ~_outflow_adjacent(flags, sh)
            nr = float(np.sqrt((r[keep] ** 2).sum()))
            nb = float(np.sqrt((bvec[keep] ** 2).sum()))
            # the downloaded velocity is float: its divergence carries a
            # rounding error of a few ulps of |u| per cell, which bounds what
            # this reconstruction can show (a projection of an already
            # divergence-free field has ||b|| at that level)
            floor = 8 * 2.0 ** -24 * float(np.sqrt((u ** 2).sum()))
            assert nr <= 1.5 * TOL * nb + floor, (i, sh, nr, nb, floor)
    assert n_req == 2 * len(trace)
