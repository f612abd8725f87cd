"""Serving: a mixed stream of plume-step and pressure-solve requests over
sessions of different sizes (2D and 3D).  Each session's state after the
served (continuously batched) replay is bit-identical to replaying the same
requests one at a time on the plain path: a result does not depend on what
it was batched with."""
import numpy as np
import pytest

from mantaflow_amd import serve

pytestmark = pytest.mark.gpu

SHAPES = [(24, 24, 24), (48, 48, 1), (40, 40, 40), (64, 64, 1), (16, 16, 16), (100, 100, 1), (33, 33, 33)]


def state(sess):
    return [sess.density.toNumpy().copy(), sess.vel.toNumpy().copy(), sess.pressure.toNumpy().copy()]


def test_served_equals_serial():
    a = [serve.PlumeScene(sh, seed=10 + i) for i, sh in enumerate(SHAPES)]
    b = [serve.PlumeScene(sh, seed=10 + i) for i, sh in enumerate(SHAPES)]
    trace = serve.make_trace(40, len(SHAPES), rate=1e9, seed=3)
    lat_s, _ = serve.serve_serial(a, trace, realtime=False)
    srv = serve.Server(b, max_batch=8)
    lat_b, _ = srv.run(trace, realtime=False)
    assert len(lat_s) == len(lat_b) == len(trace)
    assert srv.ticks < len(trace) and max(srv.batch_sizes) > 1
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.steps == y.steps
        for u, v, name in zip(state(x), state(y), ("density", "vel", "pressure")):
            np.testing.assert_array_equal(u, v, err_msg="session %d %s" % (i, name))


def test_admission_respects_memory():
    """One request per session per tick, in arrival order; nothing admitted
    beyond the memory budget; a request larger than the whole budget fails
    with the engine's message."""
    need = 6 * 4 * 16 ** 3
    trace = [serve.Request(0, 0.0, 0, serve.REQ_SOLVE), serve.Request(1, 0.0, 1, serve.REQ_SOLVE),
             serve.Request(2, 0.0, 0, serve.REQ_STEP)]
    sessions = [serve.PlumeScene((16, 16, 16), seed=1), serve.PlumeScene((16, 16, 16), seed=2)]
    srv = serve.Server(sessions, max_batch=8, workers=1, memory_limit=1 << 40)
    srv.run(trace, realtime=False)
    ticks = {}
    for x in srv.results:
        ticks.setdefault(x.tick, []).append(x.tag)
    assert sorted(sorted(v) for v in ticks.values()) == [[0, 1], [2]]
    assert sessions[0].steps == 1 and sessions[1].steps == 0
    # budget for one session's solve at a time: every tick holds one request
    srv = serve.Server(sessions, max_batch=8, workers=4, memory_limit=need)
    srv.run(trace, realtime=False)
    assert srv.batch_sizes == [1, 1, 1]
    starts = sorted((x.done, x.tag) for x in srv.results)
    assert [t for _, t in starts][-1] == 2  # the session's step after its solve
    with pytest.raises(RuntimeError, match="no request fits"):
        serve.Server(sessions, max_batch=8, workers=2, memory_limit=need - 1).run(trace[:1], realtime=False)


def test_engine_plain_path_equals_plugin_calls():
    """The engine's plain path (one request per tick) leaves a session in
    the same state, bit for bit, as the same requests run from Python one
    plugin call at a time (PlumeScene.run_single, the scene script's loop)."""
    shapes = [(20, 20, 20), (40, 40, 1)]
    a = [serve.PlumeScene(sh, seed=50 + i) for i, sh in enumerate(shapes)]
    b = [serve.PlumeScene(sh, seed=50 + i) for i, sh in enumerate(shapes)]
    trace = serve.make_trace(12, len(shapes), rate=1e9, seed=11)
    for r in sorted(trace, key=lambda r: r.arrival):
        a[r.session].run_single(r.kind)
    srv = serve.Server(b, max_batch=1, workers=1)
    lat, _ = srv.run(trace, realtime=False)
    assert sorted(lat) == sorted(r.rid for r in trace) and srv.batch_sizes == [1] * len(trace)
    assert all(x.iterations > 0 and x.relResidual < 1e-4 for x in srv.results)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.steps == y.steps
        for u, v, name in zip(state(x), state(y), ("density", "vel", "pressure")):
            np.testing.assert_array_equal(u, v, err_msg="session %d %s" % (i, name))


def test_threaded_served_equals_serial():
    """Four worker threads, ticks in flight at once on their sessions'
    streams (real-time arrivals and all at once): the same per-session
    states as the serial replay, every request answered once."""
    for realtime, rate in ((False, 1e9), (True, 400.0)):
        a = [serve.PlumeScene(sh, seed=30 + i) for i, sh in enumerate(SHAPES)]
        b = [serve.PlumeScene(sh, seed=30 + i) for i, sh in enumerate(SHAPES)]
        trace = serve.make_trace(40, len(SHAPES), rate=rate, seed=5)
        serve.serve_serial(a, trace, realtime=False)
        srv = serve.Server(b, max_batch=8, workers=4)
        lat, _ = srv.run(trace, realtime=realtime)
        assert sorted(lat) == sorted(r.rid for r in trace)
        assert sum(srv.batch_sizes) == len(trace)
        for i, (x, y) in enumerate(zip(a, b)):
            assert x.steps == y.steps
            for u, v, name in zip(state(x), state(y), ("density", "vel", "pressure")):
                np.testing.assert_array_equal(u, v, err_msg="realtime=%s session %d %s" % (realtime, i, name))


def test_session_checkpoint_resume(tmp_path):
    """A session saved after some requests and restored into a fresh session
    of the same shape continues bit-identically."""
    a = serve.PlumeScene((20, 20, 20), seed=4)
    for kind in (serve.REQ_STEP, serve.REQ_SOLVE, serve.REQ_STEP):
        a.run_single(kind)
    a.save(str(tmp_path / "ck"))
    b = serve.PlumeScene((20, 20, 20), seed=99)  # different source values until restored
    b.load(str(tmp_path / "ck"))
    assert b.steps == a.steps
    for kind in (serve.REQ_STEP, serve.REQ_STEP, serve.REQ_SOLVE):
        a.run_single(kind)
        b.run_single(kind)
    for u, v, name in zip(state(a), state(b), ("density", "vel", "pressure")):
        np.testing.assert_array_equal(u, v, err_msg=name)
    with pytest.raises(RuntimeError, match="shape"):
        serve.PlumeScene((16, 16, 16), seed=1).load(str(tmp_path / "ck"))
