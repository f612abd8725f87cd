// Serving engine behind mfx_serve_*: continuous batching of plume-step and
// pressure-solve requests over scenes resident in HBM.
//
// Threads: one dispatcher and `workers` tick runners, all native (no
// interpreter lock anywhere on the path).  The dispatcher owns admission:
// it takes ready requests of idle sessions in arrival order, one per
// session, within maxBatch and the memory budget, and hands them to free
// workers - one request per worker while no more are ready than there are
// free workers (a small request then runs beside a large one instead of
// waiting for it), else all of them as one batched tick.  A worker runs a
// single request on the plain path (mfx_solve_pressure) and a batched tick
// as the requests' pre-solve phases on their own streams, one
// mfx_solve_pressure_batch and the post phases.  Every library call a tick
// makes is on its sessions' scenes, and a session is never in two ticks at
// once, so ticks in flight do not share a stream or a grid.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "mfx_internal.h"

struct mfx_scene : public mfx::Scene {};
struct mfx_grid : public mfx::Grid {};

namespace mfx {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
	long long tag;
	int session, kind;
	double submitted;
	// the session's grids, copied at dispatch (under the lock) so that a
	// running tick never reads the session table, which add_session may grow
	mfx_serve_session g;
	Scene* scene;
};

struct Session {
	mfx_serve_session g;
	Scene* scene;
	size_t need;  // bytes counted against the budget while admitted
	bool busy = false;
};

}  // namespace
}  // namespace mfx

struct mfx_server {
	mfx_serve_opts opts;
	int device = 0;
	mfx::Clock::time_point t0;
	std::mutex mu;
	std::condition_variable cvDispatch, cvWork, cvDone;
	std::vector<mfx::Session> sessions;
	std::deque<mfx::Request> ready;
	std::deque<std::vector<mfx::Request>> work;
	std::deque<mfx_serve_result> done;
	long long outstanding = 0;  // submitted, not finished
	int inFlight = 0;           // ticks handed to workers, not finished
	size_t heldBytes = 0;       // budget held by admitted requests
	long long ticks = 0;
	bool stop = false;
	std::thread dispatcher;
	std::vector<std::thread> workers;

	double now() const { return std::chrono::duration<double>(mfx::Clock::now() - t0).count(); }
};

namespace mfx {
namespace {

Grid* G(const mfx_grid* g) { return const_cast<Grid*>(static_cast<const Grid*>(g)); }

void copy_msg(char* dst, size_t cap, const char* src) {
	std::strncpy(dst, src ? src : "", cap - 1);
	dst[cap - 1] = 0;
}

mfx_serve_result make_result(const Request& r) {
	mfx_serve_result o;
	std::memset(&o, 0, sizeof(o));
	o.tag = r.tag;
	o.session = r.session;
	o.kind = r.kind;
	o.submitted = r.submitted;
	o.relResidual = 0.0;
	return o;
}

// -- the phases of one request (the scene scripts' step loop) ---------------
int step_pre(const mfx_serve_session& s) {
	// density source, advection of density then velocity (MacCormack), the
	// open top, wall BCs and buoyancy - scenes/simpleplume.py's loop body
	// up to solvePressure
	if (mfx_grid_mult(s.density, s.keep) || mfx_grid_add(s.density, s.srcval)) return -1;
	if (mfx_advect_semi_lagrange(s.flags, s.vel, s.density, 2, 1.0f, 1, 2, 1)) return -1;
	if (mfx_advect_semi_lagrange(s.flags, s.vel, s.vel, 2, 1.0f, 1, 2, 1)) return -1;
	if (mfx_reset_outflow(s.flags, nullptr, s.density)) return -1;
	// setWallBcs then addBuoyancy, one pass
	return mfx_set_wall_bcs_buoyancy(s.flags, s.vel, s.density, s.gravity, 1.0f, 1);
}

int step_post(const mfx_serve_session& s, int kind) {
	return kind == MFX_REQ_STEP ? mfx_set_wall_bcs(s.flags, s.vel, nullptr) : 0;
}

// solvePressure; for a step its velocity correction and the setWallBcs that
// follows run as one pass
int solve_single(const mfx_serve_session& s, int kind, const mfx_pressure_opts& opts, mfx_solve_stats* st) {
	mfx_pressure_opts o = opts;
	o.deferCorrection = kind == MFX_REQ_STEP ? 1 : 0;
	if (mfx_solve_pressure(s.vel, s.pressure, s.flags, &o, st, nullptr)) return -1;
	return kind == MFX_REQ_STEP ? mfx_correct_velocity_wall_bcs(s.vel, s.pressure, s.flags) : 0;
}

// Runs one tick; fills res[i].status / error / stats.
void run_tick(mfx_server* sv, const std::vector<Request>& batch, std::vector<mfx_serve_result>& res) {
	const size_t n = batch.size();
	std::vector<mfx_solve_stats> st(n);
	auto fail = [&](size_t from, size_t to) {
		const char* m = mfx_last_error();
		for (size_t i = from; i < to; ++i) {
			res[i].status = -1;
			copy_msg(res[i].error, sizeof(res[i].error), m);
		}
	};
	auto sess = [&](size_t i) -> const mfx_serve_session& { return batch[i].g; };
	if (n == 1) {
		const mfx_serve_session& s = sess(0);
		if ((batch[0].kind == MFX_REQ_STEP && step_pre(s)) || solve_single(s, batch[0].kind, sv->opts.solve, &st[0]) ||
		    mfx_scene_sync(static_cast<mfx_scene*>(batch[0].scene))) {
			fail(0, 1);
			return;
		}
	} else {
		for (size_t i = 0; i < n; ++i)
			if (batch[i].kind == MFX_REQ_STEP && step_pre(sess(i))) {
				fail(0, n);
				return;
			}
		std::vector<mfx_grid*> V(n), P(n);
		std::vector<const mfx_grid*> F(n);
		for (size_t i = 0; i < n; ++i) {
			V[i] = sess(i).vel;
			P[i] = sess(i).pressure;
			F[i] = sess(i).flags;
		}
		mfx_pressure_opts o = sv->opts.solve;
		o.deferCorrection = 0;
		if (mfx_solve_pressure_batch((int)n, V.data(), P.data(), F.data(), &o, st.data())) {
			fail(0, n);
			return;
		}
		for (size_t i = 0; i < n; ++i)
			if (step_post(sess(i), batch[i].kind)) {
				fail(0, n);
				return;
			}
		for (size_t i = 0; i < n; ++i)
			if (mfx_scene_sync(static_cast<mfx_scene*>(batch[i].scene))) {
				fail(0, n);
				return;
			}
	}
	for (size_t i = 0; i < n; ++i) {
		res[i].iterations = st[i].iterations;
		res[i].converged = st[i].converged;
		res[i].relResidual = st[i].relResidual;
	}
}

void worker_loop(mfx_server* sv) {
	(void)hipSetDevice(sv->device);
	std::unique_lock<std::mutex> lk(sv->mu);
	for (;;) {
		sv->cvWork.wait(lk, [&] { return sv->stop || !sv->work.empty(); });
		if (sv->work.empty()) return;  // stop, nothing left to run
		std::vector<Request> batch = std::move(sv->work.front());
		sv->work.pop_front();
		const long long tick = ++sv->ticks;
		lk.unlock();
		std::vector<mfx_serve_result> res;
		for (const Request& r : batch) res.push_back(make_result(r));
		run_tick(sv, batch, res);
		const double t = sv->now();
		lk.lock();
		for (size_t i = 0; i < batch.size(); ++i) {
			res[i].done = t;
			res[i].batch = (int)batch.size();
			res[i].tick = tick;
			Session& s = sv->sessions[batch[i].session];
			s.busy = false;
			sv->heldBytes -= s.need;
			sv->done.push_back(res[i]);
		}
		sv->outstanding -= (long long)batch.size();
		sv->inFlight -= 1;
		sv->cvDispatch.notify_one();
		sv->cvDone.notify_all();
	}
}

// Admission (under the lock): ready requests of idle sessions in arrival
// order, one per session, within maxBatch and the budget.
std::vector<size_t> admit(mfx_server* sv) {
	std::vector<size_t> chosen;
	std::set<int> used;
	long long budget;
	if (sv->opts.memoryLimit > 0) {
		budget = (long long)sv->opts.memoryLimit - (long long)sv->heldBytes;
	} else {
		size_t fr = 0, tot = 0;
		if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
		budget = (long long)fr - (long long)sv->opts.hbmReserve;
	}
	for (size_t i = 0; i < sv->ready.size() && (int)chosen.size() < sv->opts.maxBatch; ++i) {
		const Request& r = sv->ready[i];
		const Session& s = sv->sessions[r.session];
		if (s.busy || used.count(r.session)) continue;
		if ((long long)s.need > budget) continue;
		budget -= (long long)s.need;
		used.insert(r.session);
		chosen.push_back(i);
	}
	return chosen;
}

void dispatcher_loop(mfx_server* sv) {
	(void)hipSetDevice(sv->device);
	std::unique_lock<std::mutex> lk(sv->mu);
	for (;;) {
		if (sv->stop) return;
		const int freeWorkers = sv->opts.workers - sv->inFlight;
		if (freeWorkers > 0 && !sv->ready.empty()) {
			std::vector<size_t> pick = admit(sv);
			if (pick.empty() && sv->inFlight == 0) {
				// nothing in flight will free memory: the first idle request cannot run
				for (size_t i = 0; i < sv->ready.size(); ++i) {
					if (sv->sessions[sv->ready[i].session].busy) continue;
					mfx_serve_result o = make_result(sv->ready[i]);
					o.status = -1;
					o.done = sv->now();
					copy_msg(o.error, sizeof(o.error), "serve: no request fits in free device memory");
					sv->done.push_back(o);
					sv->ready.erase(sv->ready.begin() + (long)i);
					sv->outstanding -= 1;
					sv->cvDone.notify_all();
					break;
				}
				continue;
			}
			if (!pick.empty()) {
				std::vector<Request> batch;
				for (size_t i : pick) batch.push_back(sv->ready[i]);
				for (size_t k = pick.size(); k-- > 0;) sv->ready.erase(sv->ready.begin() + (long)pick[k]);
				for (Request& r : batch) {
					Session& s = sv->sessions[r.session];
					s.busy = true;
					sv->heldBytes += s.need;
					r.g = s.g;
					r.scene = s.scene;
				}
				if ((int)batch.size() <= freeWorkers) {
					for (const Request& r : batch) {
						sv->work.push_back(std::vector<Request>{r});
						sv->inFlight += 1;
					}
				} else {
					sv->work.push_back(std::move(batch));
					sv->inFlight += 1;
				}
				sv->cvWork.notify_all();
				continue;
			}
		}
		sv->cvDispatch.wait(lk);
	}
}

}  // namespace
}  // namespace mfx

using namespace mfx;

#define SV_BEGIN try {
#define SV_END                                  \
	}                                           \
	catch (const std::exception& e) {           \
		::mfx::set_last_error(e.what());        \
		return -1;                              \
	}                                           \
	catch (...) {                               \
		::mfx::set_last_error("unknown error"); \
		return -1;                              \
	}                                           \
	return 0;

extern "C" {

void mfx_serve_opts_default(mfx_serve_opts* o) {
	std::memset(o, 0, sizeof(*o));
	o->maxBatch = 64;
	o->workers = 1;
	o->hbmReserve = (size_t)2 << 30;
	o->memoryLimit = 0;
	mfx_pressure_opts_default(&o->solve);
	o->solve.cgAccuracy = 1e-4f;
	o->solve.preconditioner = MFX_PC_MGSTATIC;
}

int mfx_serve_create(const mfx_serve_opts* opts, mfx_server** out) {
	SV_BEGIN
	MFX_CHECK(out != nullptr, "mfx_serve_create: out is NULL");
	mfx_serve_opts o;
	if (opts)
		o = *opts;
	else
		mfx_serve_opts_default(&o);
	MFX_CHECK(o.maxBatch >= 1, "mfx_serve_create: maxBatch must be >= 1");
	MFX_CHECK(o.workers >= 1 && o.workers <= 16, "mfx_serve_create: workers must be in 1..16");
	mfx_server* sv = new mfx_server();
	sv->opts = o;
	sv->t0 = Clock::now();
	if (hipGetDevice(&sv->device) != hipSuccess) sv->device = 0;
	try {
		sv->dispatcher = std::thread(dispatcher_loop, sv);
		for (int i = 0; i < o.workers; ++i) sv->workers.emplace_back(worker_loop, sv);
	} catch (...) {
		{
			std::lock_guard<std::mutex> g(sv->mu);
			sv->stop = true;
		}
		sv->cvDispatch.notify_all();
		sv->cvWork.notify_all();
		if (sv->dispatcher.joinable()) sv->dispatcher.join();
		for (auto& t : sv->workers) t.join();
		delete sv;
		throw;
	}
	*out = sv;
	SV_END
}

int mfx_serve_add_session(mfx_server* sv, const mfx_serve_session* g, int* id) {
	SV_BEGIN
	MFX_CHECK(sv && g, "mfx_serve_add_session: NULL argument");
	MFX_CHECK(g->flags && g->vel && g->pressure, "mfx_serve_add_session: flags, vel and pressure are required");
	Grid* f = G(g->flags);
	Grid* v = G(g->vel);
	Grid* p = G(g->pressure);
	MFX_CHECK(f->isFlags() && v->isMAC() && p->isReal() && !p->isVec(),
	          "mfx_serve_add_session: flags must be a FlagGrid, vel a MACGrid, pressure a RealGrid");
	Scene* s = f->s;
	MFX_CHECK(v->s == s && p->s == s, "mfx_serve_add_session: the session's grids must belong to one scene");
	const bool stepping = g->density != nullptr;
	if (stepping) {
		MFX_CHECK(g->keep && g->srcval, "mfx_serve_add_session: density needs keep and srcval");
		for (const mfx_grid* h : {(const mfx_grid*)g->density, g->keep, g->srcval}) {
			MFX_CHECK(G(h)->s == s, "mfx_serve_add_session: the session's grids must belong to one scene");
			MFX_CHECK(G(h)->isReal() && !G(h)->isVec(), "mfx_serve_add_session: density, keep and srcval must be RealGrids");
		}
	}
	MFX_CHECK(s->device == sv->device, "mfx_serve_add_session: the session's scene is on another device");
	std::lock_guard<std::mutex> lk(sv->mu);
	Session ss;
	ss.g = *g;
	ss.scene = s;
	ss.need = (size_t)6 * 4 * (size_t)s->d.n;  // rhs and CG vectors of its solve
	sv->sessions.push_back(ss);
	if (id) *id = (int)sv->sessions.size() - 1;
	SV_END
}

int mfx_serve_submit_many(mfx_server* sv, int count, const int* sessions, const int* kinds, const long long* tags) {
	SV_BEGIN
	MFX_CHECK(sv, "mfx_serve_submit: NULL server");
	MFX_CHECK(count >= 0 && (count == 0 || (sessions && kinds && tags)), "mfx_serve_submit: NULL argument");
	std::lock_guard<std::mutex> lk(sv->mu);
	MFX_CHECK(!sv->stop, "mfx_serve_submit: server is shutting down");
	for (int i = 0; i < count; ++i) {
		MFX_CHECK(kinds[i] == MFX_REQ_STEP || kinds[i] == MFX_REQ_SOLVE, "mfx_serve_submit: unknown request kind");
		MFX_CHECK(sessions[i] >= 0 && sessions[i] < (int)sv->sessions.size(), "mfx_serve_submit: unknown session");
		MFX_CHECK(kinds[i] != MFX_REQ_STEP || sv->sessions[sessions[i]].g.density != nullptr,
		          "mfx_serve_submit: step request on a solve-only session");
	}
	const double t = sv->now();
	for (int i = 0; i < count; ++i) sv->ready.push_back(Request{tags[i], sessions[i], kinds[i], t, {}, nullptr});
	sv->outstanding += count;
	sv->cvDispatch.notify_one();
	SV_END
}

int mfx_serve_submit(mfx_server* sv, int session, int kind, long long tag) {
	return mfx_serve_submit_many(sv, 1, &session, &kind, &tag);
}

int mfx_serve_poll(mfx_server* sv, mfx_serve_result* out, int maxOut, int* count, double timeoutSec) {
	SV_BEGIN
	MFX_CHECK(sv && count && (out || maxOut == 0), "mfx_serve_poll: NULL argument");
	std::unique_lock<std::mutex> lk(sv->mu);
	if (sv->done.empty() && sv->outstanding > 0 && timeoutSec > 0)
		sv->cvDone.wait_for(lk, std::chrono::duration<double>(timeoutSec),
		                    [&] { return !sv->done.empty() || sv->outstanding == 0; });
	int k = 0;
	while (k < maxOut && !sv->done.empty()) {
		out[k++] = sv->done.front();
		sv->done.pop_front();
	}
	*count = k;
	SV_END
}

int mfx_serve_drain(mfx_server* sv, double timeoutSec) {
	SV_BEGIN
	MFX_CHECK(sv, "mfx_serve_drain: NULL server");
	std::unique_lock<std::mutex> lk(sv->mu);
	const bool ok = sv->cvDone.wait_for(lk, std::chrono::duration<double>(timeoutSec > 0 ? timeoutSec : 0.0),
	                                    [&] { return sv->outstanding == 0; });
	MFX_CHECK(ok, "mfx_serve_drain: requests still outstanding after the timeout");
	SV_END
}

int mfx_serve_time(const mfx_server* sv, double* seconds) {
	SV_BEGIN
	MFX_CHECK(sv && seconds, "mfx_serve_time: NULL argument");
	*seconds = sv->now();
	SV_END
}

int mfx_serve_destroy(mfx_server* sv) {
	SV_BEGIN
	if (!sv) return 0;
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		sv->stop = true;
		sv->outstanding -= (long long)sv->ready.size();
		sv->ready.clear();
	}
	sv->cvDispatch.notify_all();
	sv->cvWork.notify_all();
	if (sv->dispatcher.joinable()) sv->dispatcher.join();
	for (auto& t : sv->workers) t.join();
	delete sv;
	SV_END
}

}  // extern "C"
