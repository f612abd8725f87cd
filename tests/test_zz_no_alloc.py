"""The allocation-free solve (mfx_solve_pressure_noalloc) through the C ABI
against mfx_solve_pressure with the same options (unpreconditioned CG): the
pressure, the corrected velocity and the statistics must be identical, on two
calls in a row (the second on a smaller grid, so the library array holds the
first call's values), and sizes or preconditioners it does not take are
rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from mantaflow_amd import _abi
from mantaflow_amd._abi import lib
from tests.helpers import GpuScene, assert_bitexact, make_flags, smooth_vel

pytestmark = pytest.mark.gpu


def _problem(manta, oracle, rng, shape, open_bound):
    flags = make_flags(oracle, shape, bw=1, open_bound=open_bound, obstacles=3, rng=rng)
    if open_bound:
        oracle.reset_outflow(shape, flags)
    vel = smooth_vel(rng, shape, 2, 1.0)
    return flags, vel


def _opts(**kw):
    o = _abi.PressureOpts()
    lib().mfx_pressure_opts_default(C.byref(o))
    o.preconditioner = 0
    o.cgAccuracy = 1e-4
    # plain CG needs more than the default 1.5 * max(n) iterations here
    o.cgMaxIterFac = 8.0
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _solve(manta, shape, flags, vel, noalloc, **kw):
    sc = GpuScene(manta, shape)
    f, v, p = sc.flags(flags), sc.mac(vel), sc.real()
    o = _opts(**kw)
    st = _abi.SolveStats()
    if noalloc:
        rc = lib().mfx_solve_pressure_noalloc(v._h, p._h, f._h, C.byref(o), C.byref(st))
    else:
        rc = lib().mfx_solve_pressure(v._h, p._h, f._h, C.byref(o), C.byref(st), None)
    assert rc == 0, lib().mfx_last_error()
    return p.toNumpy(), v.toNumpy(), st


def _same(a, b, what):
    pa, va, sa = a
    pb, vb, sb = b
    assert_bitexact(pa, pb, what + " pressure")
    assert_bitexact(va, vb, what + " velocity")
    assert (sa.iterations, sa.resNorm, sa.relResidual, sa.fixPidx, sa.converged) == \
        (sb.iterations, sb.resNorm, sb.relResidual, sb.fixPidx, sb.converged), what + " stats"


@pytest.mark.parametrize("shape1,shape2,open_bound", [
    ((40, 36, 28), (21, 17, 13), "Y"),
    ((120, 96, 1), (50, 44, 1), "xX"),
])
def test_noalloc_equals_pooled_solve(manta, oracle, rng, shape1, shape2, open_bound):
    for shape in (shape1, shape2):
        flags, vel = _problem(manta, oracle, rng, shape, open_bound)
        ref = _solve(manta, shape, flags, vel, False)
        got = _solve(manta, shape, flags, vel, True)
        assert ref[2].converged
        _same(got, ref, "noalloc %s" % (shape,))


def test_noalloc_options(manta, oracle, rng):
    """Zero-pressure fixing (closed domain, no empty cells), compatibility
    and the L2 stopping rule go through the same device code in both entry
    points."""
    shape = (24, 22, 20)
    flags = make_flags(oracle, shape, bw=1, obstacles=2, rng=rng)
    vel = smooth_vel(rng, shape, 2, 1.0)
    for kw in (dict(zeroPressureFixing=1), dict(enforceCompatibility=1, cgAccuracy=1e-5),
               dict(useL2Norm=1, cgAccuracy=1e-6)):
        ref = _solve(manta, shape, flags, vel, False, **kw)
        got = _solve(manta, shape, flags, vel, True, **kw)
        _same(got, ref, "noalloc %s" % kw)
        if "zeroPressureFixing" in kw:
            assert ref[2].fixPidx >= 0


def test_noalloc_rejects(manta):
    # more cells than the library array holds
    sc = GpuScene(manta, (128, 128, 65))
    f, v, p = sc.s.create(manta.FlagGrid), sc.mac(), sc.real()
    st = _abi.SolveStats()
    o = _opts()
    assert lib().mfx_solve_pressure_noalloc(v._h, p._h, f._h, C.byref(o), C.byref(st)) == -1
    assert b"maximum" in lib().mfx_last_error()
    # a multigrid preconditioner needs its hierarchy
    sc = GpuScene(manta, (16, 16, 16))
    f, v, p = sc.s.create(manta.FlagGrid), sc.mac(), sc.real()
    o = _opts(preconditioner=3)
    assert lib().mfx_solve_pressure_noalloc(v._h, p._h, f._h, C.byref(o), C.byref(st)) == -1
    assert b"unpreconditioned" in lib().mfx_last_error()
