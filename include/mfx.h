/*
 * libmfx — C ABI of the MI355X-native per-step grid path of mantaflow:
 * semi-Lagrangian / MacCormack advection, wall boundary conditions and
 * buoyancy, and the pressure projection (RHS, flag-derived Laplacian,
 * conjugate gradient with an optional multigrid preconditioner, velocity
 * correction).
 *
 * Every entry point returns 0 on success and -1 on failure; the message of
 * the last failure on the calling thread is returned by mfx_last_error().
 * Messages reuse the reference's wording where it has one.  Calls are
 * asynchronous on the scene's HIP stream; only downloads, reductions and
 * solve statistics synchronise.  A scene must not be used from two threads at
 * once.  Grids belong to the scene that created them.
 *
 * Host data layout (uploads/downloads) is the reference's: Real grids
 * float[n], Int and Flag grids int32[n], Vec3 and MAC grids float[n][3]
 * (AoS), x fastest (idx = i + nx*j + nx*ny*k).  On the device Vec3/MAC grids
 * are stored SoA and flag grids additionally keep a packed uint8 copy.
 *
 * Reference citations are relative to the mantaflow source tree.
 */
#ifndef MFX_H
#define MFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfx_scene mfx_scene;
typedef struct mfx_grid mfx_grid;

/* grid types: the reference's GridBase::GridType bits (source/grid.h:29) */
enum {
	MFX_GRID_REAL = 1,
	MFX_GRID_INT = 2,
	MFX_GRID_VEC3 = 4,
	MFX_GRID_MAC = 8 | 4,
	MFX_GRID_LEVELSET = 16 | 1,
	MFX_GRID_FLAGS = 32 | 2
};

/* preconditioners (source/plugin/pressure.cpp:27) */
enum { MFX_PC_NONE = 0, MFX_PC_MIC = 1, MFX_PC_MGDYNAMIC = 2, MFX_PC_MGSTATIC = 3 };

/* reductions for mfx_grid_reduce (source/grid.h:184-195, grid.cpp:319-372) */
enum { MFX_RED_MAX = 0, MFX_RED_MIN = 1, MFX_RED_MAXABS = 2, MFX_RED_L1 = 3, MFX_RED_L2 = 4, MFX_RED_SUM = 5 };

/* shapes (source/shapes.h:37-108) */
enum { MFX_SHAPE_BOX = 1, MFX_SHAPE_SPHERE = 2, MFX_SHAPE_CYLINDER = 3 };

/* ---------------------------------------------------------------- errors */
/* message of the last failed call on this thread (replaces the C++ exception
 * the reference's wrappers turn into RuntimeError, source/pwrapper/pclass.cpp:57-61) */
const char* mfx_last_error(void);
int mfx_device_count(int* out);
/* Free and total device memory of the current GPU (hipMemGetInfo); the
 * serving loop admits work by it.  No counterpart in the reference. */
int mfx_device_memory(size_t* freeBytes, size_t* totalBytes);

/* ---------------------------------------------------------------- scenes */
/* FluidSolver(gridSize, dim) (source/fluidsolver.cpp:107-115) */
int mfx_scene_create(int nx, int ny, int nz, int dim, mfx_scene** out);
int mfx_scene_destroy(mfx_scene* s);
/* FluidSolver::mDt, Python property `timestep` (source/fluidsolver.h:59) */
int mfx_scene_set_dt(mfx_scene* s, float dt);
int mfx_scene_get_dt(const mfx_scene* s, float* dt);
/* wait for all work queued on the scene's stream */
int mfx_scene_sync(mfx_scene* s);
/* bytes of device memory held by the scene (grids, temporaries, MG cache) */
int mfx_scene_memory(const mfx_scene* s, size_t* bytes);

/* ----------------------------------------------------------------- grids */
/* Grid<T>::Grid / MACGrid / FlagGrid constructors, zero-initialised
 * (source/grid.cpp:47-59) */
int mfx_grid_create(mfx_scene* s, int type, mfx_grid** out);
int mfx_grid_destroy(mfx_grid* g);
int mfx_grid_type(const mfx_grid* g, int* type);
int mfx_grid_upload(mfx_grid* g, const void* host, size_t bytes);
int mfx_grid_download(const mfx_grid* g, void* host, size_t bytes);
/* Grid<T>::clear, copyFrom, swap (source/grid.cpp:93-110,205-210) */
int mfx_grid_clear(mfx_grid* g);
int mfx_grid_copy(mfx_grid* dst, const mfx_grid* src);
int mfx_grid_swap(mfx_grid* a, mfx_grid* b);
/* element-wise ops (source/grid.h:472-480, grid.cpp:212-280); `v` holds one
 * value for Real/Int grids and three for Vec3/MAC grids */
int mfx_grid_set_const(mfx_grid* g, const float* v);
int mfx_grid_add_const(mfx_grid* g, const float* v);
int mfx_grid_mult_const(mfx_grid* g, const float* v);
int mfx_grid_add(mfx_grid* g, const mfx_grid* a);
int mfx_grid_sub(mfx_grid* g, const mfx_grid* a);
int mfx_grid_mult(mfx_grid* g, const mfx_grid* a);
/* g.mult(a) then g.add(b) on Real grids as one pass, bit-identical to the
 * two calls (grid.cpp:205-280).  No counterpart in the reference; the
 * manta module uses it for a mult directly followed by an add on the same
 * grid (a density source).  b must not be g. */
int mfx_grid_mult_add(mfx_grid* g, const mfx_grid* a, const mfx_grid* b);
int mfx_grid_add_scaled(mfx_grid* g, const mfx_grid* a, const float* factor);
int mfx_grid_clamp(mfx_grid* g, float vmin, float vmax);
/* getMax/getMin/getMaxAbs/getL1/getL2 (source/grid.cpp:315-389); bnd used by L1/L2 */
int mfx_grid_reduce(const mfx_grid* g, int op, int bnd, double* out);
/* setBound / setBoundNeumann / setBoundMAC (source/grid.cpp:585-648) */
int mfx_grid_set_bound(mfx_grid* g, const float* value, int boundaryWidth);
int mfx_grid_set_bound_neumann(mfx_grid* g, int boundaryWidth);
int mfx_grid_set_bound_mac(mfx_grid* g, const float* value, int boundaryWidth, int normalOnly);
/* gridMaxDiff / gridMaxDiffInt / gridMaxDiffVec3 (source/grid.cpp:400-430) */
int mfx_grid_max_diff(const mfx_grid* a, const mfx_grid* b, double* out);
/* getComponent / setComponent / copyMacToVec3 (source/grid.cpp:438,680,685) */
int mfx_get_component(const mfx_grid* vec, mfx_grid* real, int c);
int mfx_set_component(const mfx_grid* real, mfx_grid* vec, int c);
int mfx_copy_mac_to_vec3(const mfx_grid* mac, mfx_grid* vec);

/* ----------------------------------------------------------------- flags */
/* FlagGrid::initDomain (source/grid.cpp:732-842) */
int mfx_flags_init_domain(mfx_grid* flags, int boundaryWidth, const char* wall, const char* open,
                          const char* inflow, const char* outflow);
/* FlagGrid::fillGrid (source/grid.cpp:856-861) */
int mfx_flags_fill_grid(mfx_grid* flags, int type);
/* FlagGrid::countCells (source/grid.cpp:384-398); mask may be NULL */
int mfx_flags_count_cells(const mfx_grid* flags, int flag, int bnd, mfx_grid* mask, long long* out);
/* setOpenBound (source/plugin/extforces.cpp:105-130) */
int mfx_set_open_bound(mfx_grid* flags, int bWidth, const char* openBound, int type);
/* resetOutflow without particles (source/plugin/extforces.cpp:133-160) */
int mfx_reset_outflow(mfx_grid* flags, mfx_grid* phi, mfx_grid* real);

/* ---------------------------------------------------------------- shapes */
/* Shape::applyToGrid for Box/Sphere/Cylinder (source/shapes.cpp:71-84,151-154,240,324-329).
 * BOX: a=p0, b=p1.  SPHERE: a=center, b=scale, r=radius.  CYLINDER: a=center, b=z, r=radius. */
int mfx_apply_shape(mfx_grid* grid, int shape, const float* a, const float* b, float r,
                    const float* value, const mfx_grid* respectFlags);

/* ------------------------------------------------------------- advection */
/* advectSemiLagrange (source/plugin/advection.cpp:442-460); dt is the scene's.
 * orderSpace 1 (trilinear) or 2 (cubic, source/util/interpolHigh.h) and
 * orderTrace 1 (Euler) or 2 (midpoint, advection.cpp:33-37,58-73); other
 * values are rejected with the reference's assertion text. */
int mfx_advect_semi_lagrange(const mfx_grid* flags, const mfx_grid* vel, mfx_grid* grid, int order,
                             float strength, int orderSpace, int clampMode, int orderTrace);

/* ------------------------------------------------- forces / boundaries */
/* setWallBcs without fractions/phiObs (source/plugin/extforces.cpp:307-315); obvel may be NULL */
int mfx_set_wall_bcs(const mfx_grid* flags, mfx_grid* vel, const mfx_grid* obvel);
/* setWallBcs with all its arguments (source/plugin/extforces.cpp:307-315):
 * with both fractions and phiObs, KnSetWallBcsFrac (:220-303: the face
 * velocity loses its component along the obstacle-SDF normal at faces touching
 * an obstacle), otherwise mfx_set_wall_bcs.  obvel, fractions, phiObs may be
 * NULL; boundaryWidth is accepted and unused, as in the reference. */
int mfx_set_wall_bcs_frac(const mfx_grid* flags, mfx_grid* vel, const mfx_grid* obvel, const mfx_grid* fractions,
                          const mfx_grid* phiObs, int boundaryWidth);
/* addBuoyancy (source/plugin/extforces.cpp:83-87) */
int mfx_add_buoyancy(const mfx_grid* flags, const mfx_grid* density, mfx_grid* vel,
                     const float gravity[3], float coefficient, int scale);
/* setWallBcs(flags, vel) without obvel, fractions or phiObs, then
 * addBuoyancy(flags, density, vel, gravity, coefficient, scale), as one pass
 * over the grid with results bit-identical to the two calls.  No counterpart
 * in the reference (it fuses extforces.cpp:307-315 with :83-87, the order in
 * which the plume scenes call them). */
int mfx_set_wall_bcs_buoyancy(const mfx_grid* flags, mfx_grid* vel, const mfx_grid* density,
                              const float gravity[3], float coefficient, int scale);

/* --------------------------------------------------------------- pressure */
typedef struct {
	/* the reference's solvePressure kwargs (source/plugin/pressure.cpp:480-495) */
	float cgAccuracy;          /* 1e-3 */
	float cgMaxIterFac;        /* 1.5 */
	int preconditioner;        /* MFX_PC_MIC */
	int enforceCompatibility;  /* 0 */
	int useL2Norm;             /* 0 */
	int zeroPressureFixing;    /* 0 */
	/* extensions (not in the reference) */
	double relTol;             /* > 0: stop on ||r||_2 <= relTol * ||b||_2 (FP64 norms) */
	int maxIterOverride;       /* > 0: replaces the reference's iteration cap */
	int precision;             /* 0 FP32 path; 1 mixed (BF16 MG smoothing, FP64 scalars) */
	int trueResidualEvery;     /* > 0: recompute r = b - A x with FP64 accumulation every k its */
	int deferCorrection;       /* 1 (mfx_solve_pressure only): vel is left uncorrected; the caller
	                              applies mfx_correct_velocity or mfx_correct_velocity_wall_bcs with
	                              the same grids before anything reads vel (0) */
} mfx_pressure_opts;

typedef struct {
	int iterations;            /* GridCg::mIterations */
	float resNorm;             /* GridCg::mResNorm at exit */
	double relResidual;        /* ||b - A x||_2 / ||b||_2, FP64, at exit */
	int fixPidx;               /* pinned cell (zeroPressureFixing) or -1 */
	int maxIter;               /* iteration cap used */
	int preconditionerUsed;    /* preconditioner actually run on the GPU */
	int converged;             /* 1 if the stopping criterion was met */
	double seconds;            /* wall time of the call */
} mfx_solve_stats;

void mfx_pressure_opts_default(mfx_pressure_opts* o);

/* solvePressure (source/plugin/pressure.cpp:480-521).  retRhs may be NULL.
 * PcMIC (a sequential triangular solve in the reference,
 * conjugategrad.cpp:135-159) runs as multigrid-preconditioned CG in 3D and as
 * plain CG in 2D, where the reference also disables it (conjugategrad.cpp:315-321). */
int mfx_solve_pressure(mfx_grid* vel, mfx_grid* pressure, const mfx_grid* flags,
                       const mfx_pressure_opts* opts, mfx_solve_stats* stats, mfx_grid* retRhs);
/* The optional inputs of solvePressure, computePressureRhs,
 * solvePressureSystem and correctVelocity for second-order obstacle and
 * free-surface boundaries (source/plugin/pressure.cpp:277-495): every grid
 * may be NULL.  fractions (MAC: open face fractions) give MakeLaplaceMatrix's
 * fractions branch (source/conjugategrad.h:172-185) and MakeRhs's weighted
 * divergence, with obvel (MAC) its obstacle term (pressure.cpp:47-58); phi
 * (Real or Levelset: the liquid's level set) adds the ghost-fluid diagonal
 * (:133-149) and velocity correction (:151-219), with curv (Real) and
 * surfTens the surface tension terms; perCellCorr (Real) is added to the
 * divergence.  With fractions or phi the system matrix is stored as four
 * float arrays; the multigrid preconditioner (PcMG*, and PcMIC in 3D) is
 * then built from the flags' plain Laplacian, an SPD approximation of it. */
typedef struct {
	const mfx_grid* phi;
	const mfx_grid* perCellCorr;
	const mfx_grid* fractions;
	const mfx_grid* obvel;
	const mfx_grid* curv;
	float gfClamp;             /* 1e-4 */
	float surfTens;            /* 0 */
} mfx_pressure_bnd;
void mfx_pressure_bnd_default(mfx_pressure_bnd* b);
/* solvePressure with the optional boundary inputs (source/plugin/pressure.cpp:480-521) */
int mfx_solve_pressure_bnd(mfx_grid* vel, mfx_grid* pressure, const mfx_grid* flags, const mfx_pressure_opts* opts,
                           const mfx_pressure_bnd* bnd, mfx_solve_stats* stats, mfx_grid* retRhs);
/* computePressureRhs (pressure.cpp:277-299) with the optional boundary inputs */
int mfx_compute_pressure_rhs_bnd(mfx_grid* rhs, const mfx_grid* vel, const mfx_grid* flags,
                                 const mfx_pressure_opts* opts, const mfx_pressure_bnd* bnd);
/* solvePressureSystem (pressure.cpp:312-452) with the optional boundary inputs */
int mfx_solve_pressure_system_bnd(mfx_grid* rhs, mfx_grid* pressure, const mfx_grid* flags,
                                  const mfx_pressure_opts* opts, const mfx_pressure_bnd* bnd, mfx_solve_stats* stats);
/* correctVelocity (pressure.cpp:455-476) with the optional boundary inputs */
int mfx_correct_velocity_bnd(mfx_grid* vel, const mfx_grid* pressure, const mfx_grid* flags,
                             const mfx_pressure_bnd* bnd);

/* Batched solvePressure: count independent problems (vel[i], pressure[i],
 * flags[i]), each on its own scene (any mix of sizes, 2D and 3D), solved
 * together.  Every CG/multigrid phase is one launch whose work list holds the
 * tiles of all problems; two iterations are captured in a hipGraph and
 * replayed until every problem has met its stopping rule.  Results equal
 * per-problem mfx_solve_pressure calls up to the order of FP64 partial sums.
 * stats may be NULL or point to count entries.  No counterpart in the
 * reference (its callers loop over solvePressure, plugin/pressure.cpp:480). */
int mfx_solve_pressure_batch(int count, mfx_grid* const* vel, mfx_grid* const* pressure,
                             const mfx_grid* const* flags, const mfx_pressure_opts* opts, mfx_solve_stats* stats);

/* solvePressure (without retRhs) that allocates nothing: the same RHS,
 * unpreconditioned CG iterations and velocity correction as mfx_solve_pressure
 * with preconditioner MFX_PC_NONE (or MFX_PC_MIC in 2D), bit-identical to it,
 * with the right-hand side, CG vectors, matrix code and partial sums in one
 * library-held device array instead of pooled allocations.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  Limits: at most 2^20 cells
 * and 65536 partial-sum slots per launch (rejected with an error beyond).
 * Every call on a GPU shares that GPU's array, so calls that overlap on one
 * GPU (two scenes, two threads) race; serialise them.  Other preconditioners
 * are rejected with an error. */
int mfx_solve_pressure_noalloc(mfx_grid* vel, mfx_grid* pressure, const mfx_grid* flags,
                               const mfx_pressure_opts* opts, mfx_solve_stats* stats);

/* cgSolveDiffusion (source/conjugategrad.cpp:347-423): grid (Real, Vec3 or
 * MAC; every component, two in 2D) <- solution of (I + alpha L) u = grid by
 * unpreconditioned CG (maxIter = (int)(cgMaxIterFac * max size), x4 in 2D);
 * stats (may be NULL) gets the last component's solve */
int mfx_cg_solve_diffusion(const mfx_grid* flags, mfx_grid* grid, float alpha, float cgMaxIterFac, float cgAccuracy,
                           mfx_solve_stats* stats);
/* cgSolveWE (source/plugin/waves.cpp:87-141): one implicit wave-equation
 * step with s = dt^2 cSqr / 2 on the scene's dt; out <- solution, then
 * utm1 <- ut (buffer swap) and ut <- out */
int mfx_cg_solve_we(const mfx_grid* flags, mfx_grid* ut, mfx_grid* utm1, mfx_grid* out, int crankNic, float cSqr,
                    float cgMaxIterFac, float cgAccuracy, mfx_solve_stats* stats);

/* computePressureRhs (source/plugin/pressure.cpp:277-299) */
int mfx_compute_pressure_rhs(mfx_grid* rhs, const mfx_grid* vel, const mfx_grid* flags,
                             const mfx_pressure_opts* opts);
/* solvePressureSystem (source/plugin/pressure.cpp:312-452); rhs may be modified
 * by zero-pressure fixing exactly as in the reference */
int mfx_solve_pressure_system(mfx_grid* rhs, mfx_grid* pressure, const mfx_grid* flags,
                              const mfx_pressure_opts* opts, mfx_solve_stats* stats);
/* correctVelocity (source/plugin/pressure.cpp:455-476) */
int mfx_correct_velocity(mfx_grid* vel, const mfx_grid* pressure, const mfx_grid* flags);
/* correctVelocity then setWallBcs(flags, vel) (no obvel, fractions or
 * phiObs) as one pass, bit-identical to the two calls.  No counterpart in
 * the reference (it fuses pressure.cpp:87-109 with extforces.cpp:307-315,
 * the velocity correction that ends solvePressure and the setWallBcs that
 * follows it in the plume scenes); with mfx_pressure_opts.deferCorrection. */
int mfx_correct_velocity_wall_bcs(mfx_grid* vel, const mfx_grid* pressure, const mfx_grid* flags);
/* releaseMG (source/plugin/pressure.cpp:252-266); s == NULL releases all */
int mfx_release_mg(mfx_scene* s);

/* --------------------------------------------------------------- serving */
/* A request loop over scenes that stay resident in HBM.  No counterpart in
 * the reference: its scene scripts run one scene per process and call the
 * plugins of one step in sequence (scenes/simpleplume.py, plume_2d.py); the
 * step a request runs is that loop's body.  A session is one plume scene
 * (its grids created with mfx_grid_create on one scene).  Request kinds:
 *   MFX_REQ_STEP  density <- density*keep + srcval; advectSemiLagrange of
 *                 density and vel (order 2); resetOutflow(real=density);
 *                 setWallBcs; addBuoyancy(gravity); solvePressure;
 *                 setWallBcs
 *   MFX_REQ_SOLVE solvePressure of the session's current vel
 * with the server's solve options.  Continuous batching: the dispatcher
 * admits ready requests of idle sessions (one request per session at a time,
 * arrival order) while slots and device memory allow (6 floats per cell of
 * the session are counted against the free memory less hbmReserve, or
 * against memoryLimit less what admitted requests hold when that is set);
 * while no more are ready than there are free workers each becomes a tick of
 * its own (the plain path: mfx_solve_pressure), otherwise they form one
 * batched tick (their pre-solve phases, one mfx_solve_pressure_batch, their
 * post phases).  Up to `workers` ticks run at once on their sessions'
 * streams.  A session's results do not depend on what it was batched with. */
enum { MFX_REQ_STEP = 0, MFX_REQ_SOLVE = 1 };
typedef struct mfx_server mfx_server;
typedef struct {
	int maxBatch;              /* requests per batched tick (64) */
	int workers;               /* ticks in flight at once, 1..16 (1) */
	size_t hbmReserve;         /* free device memory kept unadmitted (2 GiB) */
	size_t memoryLimit;        /* > 0: admit against this budget instead of the device's free memory */
	mfx_pressure_opts solve;   /* options of every projection (cgAccuracy 1e-4, PcMGStatic) */
} mfx_serve_opts;
typedef struct {
	mfx_grid* flags;           /* FlagGrid */
	mfx_grid* vel;             /* MACGrid */
	mfx_grid* pressure;        /* RealGrid */
	mfx_grid* density;         /* RealGrid; density, keep and srcval may be NULL for solve-only sessions */
	const mfx_grid* keep;      /* RealGrid: 1 outside the source, 0 inside */
	const mfx_grid* srcval;    /* RealGrid: source values inside the source, 0 outside */
	float gravity[3];
} mfx_serve_session;
typedef struct {
	long long tag;             /* the caller's id of the request */
	int session, kind;
	int status;                /* 0 done; -1 failed (error holds the message) */
	int batch;                 /* requests in the tick that served it */
	long long tick;            /* serial number of that tick */
	double submitted, done;    /* seconds on the server's clock (mfx_serve_time) */
	int iterations;            /* CG iterations of its solve */
	int converged;             /* 1 if its solve met the stopping rule */
	double relResidual;        /* ||b - A x||_2 / ||b||_2 of its solve */
	char error[200];
} mfx_serve_result;
void mfx_serve_opts_default(mfx_serve_opts* o);
int mfx_serve_create(const mfx_serve_opts* opts, mfx_server** out);
/* registers a session (its grids must outlive the server); *id = its index */
int mfx_serve_add_session(mfx_server* sv, const mfx_serve_session* session, int* id);
/* queues a request; returns at once */
int mfx_serve_submit(mfx_server* sv, int session, int kind, long long tag);
/* queues count requests that arrive at the same instant (all checked first;
 * none is queued if one is invalid) */
int mfx_serve_submit_many(mfx_server* sv, int count, const int* sessions, const int* kinds, const long long* tags);
/* moves up to maxOut finished requests (completion order) into out; waits
 * up to timeoutSec for the first one while requests are outstanding */
int mfx_serve_poll(mfx_server* sv, mfx_serve_result* out, int maxOut, int* count, double timeoutSec);
/* waits until every submitted request has finished (error after timeoutSec) */
int mfx_serve_drain(mfx_server* sv, double timeoutSec);
/* seconds since the server was created (the clock of submitted/done) */
int mfx_serve_time(const mfx_server* sv, double* seconds);
/* waits for the ticks in flight, drops requests not yet admitted, joins the
 * server's threads; the sessions' grids are not freed */
int mfx_serve_destroy(mfx_server* sv);

/* ------------------------------------------------------------- profiling */
/* optional HIP-event timing of the main kernels on their launching stream
 * (names: "cg_matvec", "cg_update", "mg_l0_smooth", "mg_l0_residual",
 * "advect_sl", "advect_maccormack"); off by default, no counterpart in the
 * reference (its TimingData times whole plugin calls, timing.cpp:23-89) */
int mfx_kernel_timing(int enable);
int mfx_kernel_timing_reset(void);
int mfx_kernel_timing_get(const char* name, double* totalMs, long long* count);

#ifdef __cplusplus
}
#endif
#endif
