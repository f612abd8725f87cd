// A C++ host program that drives the 3D smoke plume through the C ABI only
// (include/mfx.h), the way reference-side C++ code would: scene and grids,
// the plugins of one step in the scene scripts' order, and the serving
// engine.  Used by tests/test_gpu_cabi.py.
//
//   plume_host <nx> <ny> <nz> <steps> <srcval.f32> <outdir>
//
// srcval.f32: nx*ny*nz float32 source values (drawn by the test).  Writes
// <outdir>/{direct,served}_{density,vel,pressure}.f32: the plume after
// <steps> steps made with direct plugin calls, and the same plume stepped by
// a serving engine (one worker, so the two sessions' requests form batched
// ticks) together with a second, smaller plume.  Prints one line per phase; exits non-zero on any
// error with the library's message.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mfx.h"

#define CK(call)                                                                 \
	do {                                                                         \
		if ((call) != 0) {                                                       \
			std::fprintf(stderr, "%s failed: %s\n", #call, mfx_last_error());   \
			std::exit(1);                                                        \
		}                                                                        \
	} while (0)

struct Plume {
	mfx_scene* s = nullptr;
	mfx_grid *flags = nullptr, *vel = nullptr, *density = nullptr, *pressure = nullptr, *keep = nullptr,
	         *srcval = nullptr;
	int nx = 0, ny = 0, nz = 0;
	long long n() const { return (long long)nx * ny * nz; }
};

static mfx_grid* grid(mfx_scene* s, int type) {
	mfx_grid* g = nullptr;
	CK(mfx_grid_create(s, type, &g));
	return g;
}

// the plume of scenes/simpleplume.py: walls, open top, cylindrical source
static Plume make_plume(int nx, int ny, int nz, const std::vector<float>& val) {
	Plume p;
	p.nx = nx;
	p.ny = ny;
	p.nz = nz;
	CK(mfx_scene_create(nx, ny, nz, 3, &p.s));
	CK(mfx_scene_set_dt(p.s, 1.0f));
	p.flags = grid(p.s, MFX_GRID_FLAGS);
	p.vel = grid(p.s, MFX_GRID_MAC);
	p.density = grid(p.s, MFX_GRID_REAL);
	p.pressure = grid(p.s, MFX_GRID_REAL);
	p.keep = grid(p.s, MFX_GRID_REAL);
	p.srcval = grid(p.s, MFX_GRID_REAL);
	CK(mfx_flags_init_domain(p.flags, 1, "xXyYzZ", "      ", "      ", "      "));
	CK(mfx_flags_fill_grid(p.flags, 1));
	CK(mfx_set_open_bound(p.flags, 1, "Y", 16 | 4));
	mfx_grid* mask = grid(p.s, MFX_GRID_REAL);
	const float center[3] = {(float)(0.5 * nx), (float)(0.1 * ny), (float)(0.5 * nz)};
	const float zax[3] = {0.f, (float)(0.02 * ny), 0.f};
	const float one = 1.0f;
	CK(mfx_apply_shape(mask, MFX_SHAPE_CYLINDER, center, zax, (float)(0.14 * nx), &one, nullptr));
	CK(mfx_grid_upload(p.srcval, val.data(), sizeof(float) * val.size()));
	CK(mfx_grid_mult(p.srcval, mask));
	CK(mfx_grid_set_const(p.keep, &one));
	CK(mfx_grid_sub(p.keep, mask));
	CK(mfx_grid_destroy(mask));
	return p;
}

static mfx_pressure_opts solve_opts() {
	mfx_pressure_opts o;
	mfx_pressure_opts_default(&o);
	o.cgAccuracy = 1e-4f;
	o.preconditioner = MFX_PC_MGSTATIC;
	o.relTol = 1e-4;
	return o;
}

// one plume step as direct plugin calls
static void step(Plume& p) {
	const float g[3] = {0.f, -1e-3f, 0.f};
	CK(mfx_grid_mult(p.density, p.keep));
	CK(mfx_grid_add(p.density, p.srcval));
	CK(mfx_advect_semi_lagrange(p.flags, p.vel, p.density, 2, 1.0f, 1, 2, 1));
	CK(mfx_advect_semi_lagrange(p.flags, p.vel, p.vel, 2, 1.0f, 1, 2, 1));
	CK(mfx_reset_outflow(p.flags, nullptr, p.density));
	CK(mfx_set_wall_bcs(p.flags, p.vel, nullptr));
	CK(mfx_add_buoyancy(p.flags, p.density, p.vel, g, 1.0f, 1));
	const mfx_pressure_opts o = solve_opts();
	mfx_solve_stats st;
	CK(mfx_solve_pressure(p.vel, p.pressure, p.flags, &o, &st, nullptr));
	if (!st.converged) {
		std::fprintf(stderr, "solve did not converge\n");
		std::exit(1);
	}
	CK(mfx_set_wall_bcs(p.flags, p.vel, nullptr));
}

static void dump(const Plume& p, const std::string& dir, const std::string& tag) {
	struct {
		const char* name;
		mfx_grid* g;
		int comp;
	} gs[] = {{"density", p.density, 1}, {"vel", p.vel, 3}, {"pressure", p.pressure, 1}};
	for (auto& e : gs) {
		std::vector<float> h((size_t)p.n() * e.comp);
		CK(mfx_grid_download(e.g, h.data(), sizeof(float) * h.size()));
		const std::string path = dir + "/" + tag + "_" + e.name + ".f32";
		FILE* f = std::fopen(path.c_str(), "wb");
		if (!f || std::fwrite(h.data(), sizeof(float), h.size(), f) != h.size()) {
			std::fprintf(stderr, "cannot write %s\n", path.c_str());
			std::exit(1);
		}
		std::fclose(f);
	}
}

int main(int argc, char** argv) {
	if (argc != 7) {
		std::fprintf(stderr, "usage: %s nx ny nz steps srcval.f32 outdir\n", argv[0]);
		return 2;
	}
	const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]), steps = std::atoi(argv[4]);
	std::vector<float> val((size_t)nx * ny * nz);
	FILE* f = std::fopen(argv[5], "rb");
	if (!f || std::fread(val.data(), sizeof(float), val.size(), f) != val.size()) {
		std::fprintf(stderr, "cannot read %s\n", argv[5]);
		return 1;
	}
	std::fclose(f);
	const std::string out = argv[6];

	// direct plugin calls
	Plume a = make_plume(nx, ny, nz, val);
	for (int i = 0; i < steps; i++) step(a);
	CK(mfx_scene_sync(a.s));
	dump(a, out, "direct");
	std::printf("direct: %d steps\n", steps);

	// the same plume through the serving engine, beside a smaller one
	Plume b = make_plume(nx, ny, nz, val);
	std::vector<float> val2((size_t)24 * 24 * 24, 0.9f);
	Plume c = make_plume(24, 24, 24, val2);
	mfx_serve_opts so;
	mfx_serve_opts_default(&so);
	so.workers = 1;
	so.maxBatch = 8;
	so.solve = solve_opts();
	mfx_server* sv = nullptr;
	CK(mfx_serve_create(&so, &sv));
	int ids[2];
	for (int k = 0; k < 2; k++) {
		Plume& p = k == 0 ? b : c;
		mfx_serve_session g = {p.flags, p.vel, p.pressure, p.density, p.keep, p.srcval, {0.f, -1e-3f, 0.f}};
		CK(mfx_serve_add_session(sv, &g, &ids[k]));
	}
	std::vector<int> sess, kinds;
	std::vector<long long> tags;
	for (int i = 0; i < steps; i++)
		for (int k = 0; k < 2; k++) {
			sess.push_back(ids[k]);
			kinds.push_back(MFX_REQ_STEP);
			tags.push_back((long long)tags.size());
		}
	CK(mfx_serve_submit_many(sv, (int)sess.size(), sess.data(), kinds.data(), tags.data()));
	CK(mfx_serve_drain(sv, 600.0));
	std::vector<mfx_serve_result> res(sess.size());
	int got = 0, maxBatch = 0;
	while (got < (int)res.size()) {
		int cnt = 0;
		CK(mfx_serve_poll(sv, res.data() + got, (int)res.size() - got, &cnt, 1.0));
		if (cnt == 0) break;
		got += cnt;
	}
	for (int i = 0; i < got; i++) {
		if (res[(size_t)i].status != 0 || !res[(size_t)i].converged) {
			std::fprintf(stderr, "request %lld failed: %s\n", res[(size_t)i].tag, res[(size_t)i].error);
			return 1;
		}
		if (res[(size_t)i].batch > maxBatch) maxBatch = res[(size_t)i].batch;
	}
	CK(mfx_serve_destroy(sv));
	dump(b, out, "served");
	std::printf("served: %d requests, largest batch %d\n", got, maxBatch);
	for (Plume* p : {&a, &b, &c}) CK(mfx_scene_destroy(p->s));
	return got == (int)res.size() ? 0 : 1;
}
