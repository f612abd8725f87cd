"""Serving loop for smoke-step and pressure-solve requests on scenes that stay
resident in HBM: the Python face of the library's serving engine
(mfx_serve_* in include/mfx.h).

A request names a session (one plume scene: flags, MAC velocity, density,
pressure, source) and a kind: "step" = one plume step (density source,
MacCormack advection of density and velocity, resetOutflow, setWallBcs,
addBuoyancy, solvePressure, setWallBcs, as scenes/plume_2d.py and
scenes/simpleplume.py do), "solve" = the pressure projection of the
session's current velocity.

Server (continuous batching, native threads): the engine's dispatcher admits
the requests that have arrived, whose session is idle, while batch slots and
free HBM allow, and hands them to worker threads (one tick in flight per
worker): one request per worker while no more are ready than there are free
workers (the plain path, so a small request runs beside a large one instead
of queueing behind it), else all of them as one batched tick - their
pre-solve phases on their own HIP streams, one mfx_solve_pressure_batch
(one launch per CG/multigrid phase for every admitted problem), then the
post phases.  Requests arriving during a tick are admitted as workers free
up.  A session's results do not depend on what it was batched with (the
batched solve is bit-identical to a single solve).

serve_serial replays the same trace on the plain path: the same engine with
one worker and one request per tick, each with its own solvePressure.
PlumeScene.run_single is the same plain path driven from Python, one plugin
call at a time, as a scene script would.
"""
import collections
import ctypes as C
import itertools
import math
import time

import numpy as np

from . import _abi
from . import manta as m
from ._abi import check
from .manta import lib  # runs any held-back plugin call first

REQ_STEP, REQ_SOLVE = "step", "solve"
Request = collections.namedtuple("Request", "rid arrival session kind")


class PlumeScene:
    """One session: an open-top smoke plume resident on the GPU."""

    def __init__(self, shape, seed=1, rel_tol=1e-4, preconditioner=None):
        nx, ny, nz = shape
        self.shape = tuple(shape)
        self.dim = 3 if nz > 1 else 2
        s = m.Solver(name="session", gridSize=m.vec3(nx, ny, nz), dim=self.dim)
        s.timestep = 1.0
        s.mfxRelTol = rel_tol
        self.s = s
        self.pc = m.PcMGStatic if preconditioner is None else preconditioner
        self.flags = s.create(m.FlagGrid)
        self.vel = s.create(m.MACGrid)
        self.density = s.create(m.RealGrid)
        self.pressure = s.create(m.RealGrid)
        self.flags.initDomain(boundaryWidth=1)
        self.flags.fillGrid()
        m.setOpenBound(self.flags, 1, "Y", m.FlagOutflow | m.FlagEmpty)
        mask = s.create(m.RealGrid)
        if self.dim == 3:
            src = m.Cylinder(s, center=m.vec3(0.5 * nx, 0.1 * ny, 0.5 * nz), radius=0.14 * nx,
                             z=m.vec3(0, 0.02 * ny, 0))
        else:
            src = m.Box(s, p0=m.vec3(0.4 * nx, 0.05 * ny, 0), p1=m.vec3(0.6 * nx, 0.1 * ny, 1))
        src.applyToGrid(grid=mask, value=1.0)
        val = np.random.default_rng(seed).uniform(0.8, 1.0, nx * ny * nz).astype(np.float32)
        self.srcval = s.create(m.RealGrid)
        self.srcval.fromNumpy(val)
        self.srcval.mult(mask)
        self.keep = s.create(m.RealGrid)
        self.keep.setConst(1.0)
        self.keep.sub(mask)
        self.gravity = m.vec3(0, -1e-3, 0)
        self.steps = 0

    @property
    def cells(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    # phases of one step
    def pre(self):
        d, v, f = self.density, self.vel, self.flags
        d.mult(self.keep)
        d.add(self.srcval)
        m.advectSemiLagrange(flags=f, vel=v, grid=d, order=2)
        m.advectSemiLagrange(flags=f, vel=v, grid=v, order=2, strength=1.0)
        m.resetOutflow(flags=f, real=d)
        m.setWallBcs(flags=f, vel=v)
        m.addBuoyancy(density=d, vel=v, gravity=self.gravity, flags=f)

    def post(self, kind):
        if kind == REQ_STEP:
            m.setWallBcs(flags=self.flags, vel=self.vel)
            self.steps += 1

    def problem(self):
        return (self.vel, self.pressure, self.flags)

    def session_struct(self):
        """The session as the serving engine takes it (mfx_serve_session)."""
        g = _abi.ServeSession()
        g.flags, g.vel, g.pressure = self.flags._h.value, self.vel._h.value, self.pressure._h.value
        g.density, g.keep, g.srcval = self.density._h.value, self.keep._h.value, self.srcval._h.value
        g.gravity[:] = [self.gravity.x, self.gravity.y, self.gravity.z]
        return g

    def run_single(self, kind):
        """The plain path for one request."""
        if kind == REQ_STEP:
            self.pre()
        m.solvePressure(flags=self.flags, vel=self.vel, pressure=self.pressure, cgAccuracy=1e-4,
                        preconditioner=self.pc)
        self.post(kind)
        self.s.sync()

    # checkpoint / resume: every grid of the session in the reference's .uni
    # format (one file each) plus the step counter; a restored session
    # continues exactly as the saved one would have
    _GRIDS = ("flags", "vel", "density", "pressure", "srcval", "keep")

    def save(self, directory):
        import json
        import os
        os.makedirs(directory, exist_ok=True)
        for name in self._GRIDS:
            getattr(self, name).save(os.path.join(directory, name + ".uni"))
        with open(os.path.join(directory, "session.json"), "w") as f:
            json.dump({"shape": list(self.shape), "steps": self.steps,
                       "gravity": [self.gravity.x, self.gravity.y, self.gravity.z]}, f)

    def load(self, directory):
        import json
        import os
        with open(os.path.join(directory, "session.json")) as f:
            meta = json.load(f)
        if tuple(meta["shape"]) != self.shape:
            raise RuntimeError("session checkpoint of shape %s, session %s" % (tuple(meta["shape"]), self.shape))
        for name in self._GRIDS:
            getattr(self, name).load(os.path.join(directory, name + ".uni"))
        self.steps = int(meta["steps"])
        self.gravity = m.vec3(*meta["gravity"])

    def memory(self):
        b = C.c_size_t()
        check(lib().mfx_scene_memory(self.s._h, C.byref(b)))
        return b.value

    def sync(self):
        self.s.sync()


def device_memory():
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().mfx_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value


def make_sessions(n, seed=9, min_cells=64 * 64, max_cells=192 ** 3):
    """Session sizes log-uniform in cell count, 2D or 3D by the draw; rounded
    to a square or a cube."""
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(n):
        cells = math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))
        if rng.random() < 0.5:
            a = max(8, int(round(cells ** 0.5)))
            shapes.append((a, a, 1))
        else:
            a = max(8, int(round(cells ** (1.0 / 3.0))))
            shapes.append((a, a, a))
    return shapes


def make_trace(n_req, n_sessions, rate=200.0, seed=8):
    """Poisson arrivals (mean `rate` per second), sessions and kinds drawn
    uniformly."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    sess = rng.integers(0, n_sessions, n_req)
    kinds = np.where(rng.random(n_req) < 0.5, REQ_STEP, REQ_SOLVE)
    return [Request(i, float(t[i]), int(sess[i]), str(kinds[i])) for i in range(n_req)]


def _percentiles(lat):
    a = np.sort(np.asarray(lat, np.float64))
    if len(a) == 0:
        return 0.0, 0.0
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


class Server:
    """Continuous-batching engine over resident sessions (the library's
    mfx_serve_*): `workers` ticks in flight at once, batched ticks of at most
    `max_batch` requests, admission by free device memory less
    `hbm_reserve` (or by `memory_limit` bytes when that is set)."""

    def __init__(self, sessions, max_batch=64, hbm_reserve=2 << 30, workers=1, memory_limit=0):
        self.sessions = sessions
        self.max_batch = max_batch
        self.hbm_reserve = hbm_reserve
        self.workers = max(1, int(workers))
        self.memory_limit = int(memory_limit)
        self.ticks = 0
        self.batch_sizes = []
        self.results = []

    def _options(self):
        o = _abi.ServeOpts()
        lib().mfx_serve_opts_default(C.byref(o))
        o.maxBatch = int(self.max_batch)
        o.workers = self.workers
        o.hbmReserve = int(self.hbm_reserve)
        o.memoryLimit = self.memory_limit
        first = self.sessions[0]
        o.solve.cgAccuracy = 1e-4
        o.solve.preconditioner = int(first.pc)
        o.solve.relTol = float(getattr(first.s, "mfxRelTol", 0.0) or 0.0)
        return o

    def run(self, trace, realtime=True):
        """Replay `trace` (arrival times in seconds from the start; all at
        once when realtime is False); returns per-request latency (s, from
        the request's arrival to its completion) and the makespan."""
        L = lib()
        h = C.c_void_p()
        check(L.mfx_serve_create(C.byref(self._options()), C.byref(h)))
        res = []
        try:
            for sess in self.sessions:
                check(L.mfx_serve_add_session(h, C.byref(sess.session_struct()), None))
            t = C.c_double()
            check(L.mfx_serve_time(h, C.byref(t)))
            start = t.value
            w0 = time.perf_counter()
            order = sorted(trace, key=lambda r: r.arrival)
            # requests of one arrival instant (all of them when not in real
            # time) are queued together
            groups = [order] if not realtime else [list(g) for _, g in itertools.groupby(order, lambda r: r.arrival)]
            for g in groups:
                if realtime:
                    # sleep until shortly before the arrival, then spin: a
                    # request is submitted within microseconds of its
                    # arrival time (sleep alone overshoots by ~0.1 ms)
                    wait = g[0].arrival - (time.perf_counter() - w0)
                    if wait > 1e-3:
                        time.sleep(wait - 1e-3)
                    while time.perf_counter() - w0 < g[0].arrival:
                        pass
                n = len(g)
                sess = (C.c_int * n)(*[r.session for r in g])
                kinds = (C.c_int * n)(*[_abi.REQ_STEP if r.kind == REQ_STEP else _abi.REQ_SOLVE for r in g])
                tags = (C.c_longlong * n)(*[r.rid for r in g])
                check(L.mfx_serve_submit_many(h, n, sess, kinds, tags))
            check(L.mfx_serve_drain(h, 3600.0))
            buf = (_abi.ServeResult * 256)()
            n = C.c_int()
            while True:
                check(L.mfx_serve_poll(h, buf, len(buf), C.byref(n), 0.0))
                if n.value == 0:
                    break
                res.extend(_abi.ServeResult.from_buffer_copy(buf[i]) for i in range(n.value))
        finally:
            check(L.mfx_serve_destroy(h))
        self.results = res
        bad = [x for x in res if x.status != 0]
        if bad:
            raise RuntimeError(bad[0].error.decode(errors="replace"))
        ticks = {}
        for x in res:
            ticks[x.tick] = x.batch
            if x.kind == _abi.REQ_STEP:
                self.sessions[x.session].steps += 1
        self.ticks += len(ticks)
        self.batch_sizes.extend(ticks[k] for k in sorted(ticks))
        arrival = {r.rid: r.arrival for r in trace}
        lat = {x.tag: x.done - start - (arrival[x.tag] if realtime else 0.0) for x in res}
        makespan = max((x.done for x in res), default=start) - start
        return lat, makespan


def serve_serial(sessions, trace, realtime=True):
    """Plain path: requests one at a time in arrival order, each with its
    own solvePressure (the engine with one worker and one request per tick)."""
    return Server(sessions, max_batch=1, workers=1).run(trace, realtime)


def summarize(lat, makespan):
    p50, p99 = _percentiles(list(lat.values()))
    return {"requests": len(lat), "solves_per_s": len(lat) / makespan if makespan > 0 else 0.0,
            "p50_ms": p50 * 1e3, "p99_ms": p99 * 1e3, "makespan_s": makespan}
