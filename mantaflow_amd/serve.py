"""Serving loop for smoke-step and pressure-solve requests on scenes that stay
resident in HBM.

A request names a session (one plume scene: flags, MAC velocity, density,
pressure, source) and a kind: "step" = one plume step (density source,
MacCormack advection of density and velocity, resetOutflow, setWallBcs,
addBuoyancy, solvePressure, setWallBcs, as scenes/plume_2d.py and
scenes/simpleplume.py do), "solve" = the pressure projection of the
session's current velocity.

Server (continuous batching): every engine tick admits the requests that
have arrived, whose session is idle, while batch slots and free HBM allow;
their pre-solve phases run on their own HIP streams (so they overlap on the
device), then all their pressure systems are solved by one batched solve
(mfx_solve_pressure_batch: one launch per CG/multigrid phase for every
admitted problem), then the post phases.  Requests arriving during a tick
are admitted by the next one.  A session's results do not depend on what it
was batched with (the batched solve is bit-identical to a single solve).

With workers > 1 the ticks run on worker threads (one per hardware queue of
the GPU; the library call releases the interpreter lock), so a small request
no longer waits for a large one to finish: the dispatcher hands the ready
requests of idle sessions to free workers, one request per worker while
there are no more than free workers, else all of them as one batched tick.

serve_serial replays the same trace on the plain path: one request at a
time, each with its own solvePressure.
"""
import collections
import ctypes as C
import math
import queue
import threading
import time

import numpy as np

from . import manta as m
from ._abi import check, lib

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
    """Continuous-batching engine over resident sessions (workers: ticks in
    flight at once, each on its own thread)."""

    def __init__(self, sessions, max_batch=64, hbm_reserve=2 << 30, workers=1):
        self.sessions = sessions
        self.max_batch = max_batch
        self.hbm_reserve = hbm_reserve
        self.workers = max(1, int(workers))
        self.ticks = 0
        self.batch_sizes = []
        self._lock = threading.Lock()

    def _admit(self, ready, free_bytes):
        chosen, used, budget = [], set(), free_bytes - self.hbm_reserve
        for r in ready:
            if len(chosen) >= self.max_batch:
                break
            if r.session in used:
                continue  # one request per session per tick, in arrival order
            need = 6 * 4 * self.sessions[r.session].cells  # rhs and CG vectors of the solve
            if need > budget:
                continue
            budget -= need
            used.add(r.session)
            chosen.append(r)
        return chosen

    def _count(self, n):
        with self._lock:
            self.ticks += 1
            self.batch_sizes.append(n)

    def tick(self, batch):
        if len(batch) == 1:
            # nothing to share: the plain path has the lowest latency
            self.sessions[batch[0].session].run_single(batch[0].kind)
            self._count(1)
            return
        for r in batch:
            if r.kind == REQ_STEP:
                self.sessions[r.session].pre()
        m.solvePressureBatch([self.sessions[r.session].problem() for r in batch], cgAccuracy=1e-4,
                             preconditioner=self.sessions[batch[0].session].pc)
        for r in batch:
            self.sessions[r.session].post(r.kind)
        for r in batch:
            self.sessions[r.session].sync()
        self._count(len(batch))

    def run(self, trace, realtime=True):
        """Replay `trace`; returns per-request latency (s) and the makespan."""
        if self.workers > 1:
            return self._run_workers(trace, realtime)
        pending = collections.deque(sorted(trace, key=lambda r: r.arrival))
        ready = []
        lat = {}
        t0 = time.perf_counter()
        while pending or ready:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                ready.append(pending.popleft())
            if not ready:
                time.sleep(max(0.0, pending[0].arrival - now))
                continue
            free, _ = device_memory()
            batch = self._admit(ready, free)
            if not batch:
                raise RuntimeError("serve: no request fits in free device memory")
            self.tick(batch)
            done = time.perf_counter() - t0
            ids = {r.rid for r in batch}
            for r in batch:
                lat[r.rid] = done - (r.arrival if realtime else 0.0)
            ready = [r for r in ready if r.rid not in ids]
        return lat, time.perf_counter() - t0

    def _split(self, batch, free_workers):
        """Few ready requests (no more than the free workers): one tick each,
        in flight at once (latency); more: one batched tick (a batch shares
        every launch, which is what raises throughput)."""
        if len(batch) <= free_workers:
            return [[r] for r in batch]
        return [batch]

    def _run_workers(self, trace, realtime):
        pending = collections.deque(sorted(trace, key=lambda r: r.arrival))
        ready = []
        busy = set()
        lat = {}
        work_q, done_q = queue.Queue(), queue.Queue()
        t0 = time.perf_counter()

        def worker():
            while True:
                batch = work_q.get()
                if batch is None:
                    return
                err = None
                try:
                    self.tick(batch)
                except BaseException as e:  # reported by the dispatcher
                    err = e
                done_q.put((batch, time.perf_counter() - t0, err))

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.workers)]
        for th in threads:
            th.start()
        in_flight = 0
        try:
            while pending or ready or in_flight:
                now = time.perf_counter() - t0
                while pending and (not realtime or pending[0].arrival <= now):
                    ready.append(pending.popleft())
                free_workers = self.workers - in_flight
                if free_workers > 0:
                    cand = [r for r in ready if r.session not in busy]
                    batch = []
                    if cand:
                        free, _ = device_memory()
                        batch = self._admit(cand, free)
                        if not batch and in_flight == 0:
                            raise RuntimeError("serve: no request fits in free device memory")
                    if batch:
                        for group in self._split(batch, free_workers):
                            busy.update(r.session for r in group)
                            ids = {r.rid for r in group}
                            ready = [r for r in ready if r.rid not in ids]
                            work_q.put(group)
                            in_flight += 1
                        continue
                # wait for a completion, or for the next arrival
                wait = None
                if realtime and pending:
                    wait = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                if in_flight == 0:
                    time.sleep(wait or 0.0)
                    continue
                try:
                    group, done, err = done_q.get(timeout=wait)
                except queue.Empty:
                    continue
                in_flight -= 1
                if err is not None:
                    raise err
                for r in group:
                    lat[r.rid] = done - (r.arrival if realtime else 0.0)
                    busy.discard(r.session)
        finally:
            for _ in threads:
                work_q.put(None)
            for th in threads:
                th.join()
        return lat, time.perf_counter() - t0


def serve_serial(sessions, trace, realtime=True):
    """Plain path: requests one at a time in arrival order."""
    lat = {}
    t0 = time.perf_counter()
    for r in sorted(trace, key=lambda r: r.arrival):
        if realtime:
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
        sessions[r.session].run_single(r.kind)
        lat[r.rid] = time.perf_counter() - t0 - (r.arrival if realtime else 0.0)
    return lat, time.perf_counter() - t0


def summarize(lat, makespan):
    p50, p99 = _percentiles(list(lat.values()))
    return {"requests": len(lat), "solves_per_s": len(lat) / makespan if makespan > 0 else 0.0,
            "p50_ms": p50 * 1e3, "p99_ms": p99 * 1e3, "makespan_s": makespan}
