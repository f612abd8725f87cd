/*
 * CPU oracle for the mantaflow per-step grid path (advection + pressure
 * projection).  TEST INFRASTRUCTURE ONLY: this library is the checker that
 * tests/, the smoke entry point and the bench's CPU-baseline leg compare the
 * HIP path against.  It is never linked into, loaded by or called from the
 * product library (libmfx).
 *
 * It is a plain-C restatement of the reference's CPU plugins, written from
 * the reference sources (citations per function in mfx_oracle.c).  Data uses
 * the reference's host layout:
 *   - Real grids : og_real[n]                 (x fastest, idx = i + nx*j + sz*k)
 *   - Vec3 / MAC : og_real[3*n], AoS (x,y,z)  (grid.h: Grid<Vec3>)
 *   - Flag grids : int32[n]                    (grid.h:284-365 bit values)
 * with sz = nx*ny in 3D and sz = 0 in 2D (grid.cpp:55).
 *
 * og_real is float by default (the reference's FLOATINGPOINT_PRECISION=1) and
 * double when built with -DOG_F64 (the reference's DOUBLEPRECISION build,
 * used as the FP64 error yardstick).
 */
#ifndef MFX_ORACLE_H
#define MFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef OG_F64
typedef double og_real;
#else
typedef float og_real;
#endif

/* cell types, grid.h:292-304 */
enum {
	OG_FLUID = 1, OG_OBSTACLE = 2, OG_EMPTY = 4, OG_INFLOW = 8,
	OG_OUTFLOW = 16, OG_OPEN = 32, OG_STICK = 64, OG_RESERVED = 256
};

/* grid type codes, grid.h:29 */
enum { OG_TYPE_REAL = 1, OG_TYPE_INT = 2, OG_TYPE_VEC3 = 4, OG_TYPE_MAC = 8 };

/* preconditioners, pressure.cpp:27 */
enum { OG_PC_NONE = 0, OG_PC_MIC = 1, OG_PC_MGDYN = 2, OG_PC_MGSTATIC = 3 };

typedef struct {
	int nx, ny, nz;
	int is3d;
} og_dims;

typedef struct {
	og_real cgAccuracy;          /* default 1e-3 */
	og_real cgMaxIterFac;        /* default 1.5 */
	int preconditioner;          /* default OG_PC_MIC */
	int enforceCompatibility;    /* default 0 */
	int useL2Norm;               /* default 0 */
	int zeroPressureFixing;      /* default 0 */
	/* extension (not in the reference): stop on ||r||_2/||b||_2 <= relTol
	 * (FP64 norms over all cells) instead of the reference criterion, when > 0 */
	double relTol;
	int maxIterOverride;         /* > 0 replaces the reference maxIter rule */
} og_pressure_opts;

typedef struct {
	int iterations;
	og_real resNorm;             /* reference's mResNorm at exit */
	double relResidual;          /* ||b - A x||_2 / ||b||_2 (FP64) at exit */
	int fixPidx;                 /* pinned cell or -1 */
	int diverged;                /* 1 if the reference would throw the divergence error */
	int maxIter;
} og_solve_stats;

/* thread count for the data-parallel loops (1 = serial); reductions are
 * chunked so results do not depend on it */
void og_set_threads(int n);
int og_get_threads(void);

/* ---- flags (grid.cpp:732-861, extforces.cpp:105-160) ---- */
void og_init_domain(og_dims d, int* flags, int boundaryWidth, const char* wall,
                    const char* open, const char* inflow, const char* outflow);
void og_fill_grid(og_dims d, int* flags, int type);
void og_set_open_bound(og_dims d, int* flags, int bWidth, const char* openBound, int type);
void og_reset_outflow(og_dims d, int* flags, og_real* phi, og_real* real);

/* ---- interpolation (util/interpol.h:50-157) ---- */
og_real og_interpol(og_dims d, const og_real* data, og_real px, og_real py, og_real pz);
og_real og_interpol_component(og_dims d, const og_real* vec3data, int c,
                              og_real px, og_real py, og_real pz);
/* cubic lookup, util/interpolHigh.h:22-170: Real rules / Vec3 rules per component */
og_real og_interpol_cubic(og_dims d, const og_real* data, og_real px, og_real py, og_real pz);
og_real og_interpol_cubic_component(og_dims d, const og_real* vec3data, int c, og_real px, og_real py,
                                    og_real pz);
/* interpolMAC (util/interpol.h:121-157): all three components of a MAC grid, out[3] */
void og_interpol_mac(og_dims d, const og_real* vec3data, og_real px, og_real py, og_real pz, og_real* out);

/* ---- advection (advection.cpp:25-460) ---- */
/* gridType: OG_TYPE_REAL, OG_TYPE_VEC3 or OG_TYPE_MAC; grid may alias vel */
int og_advect_semi_lagrange(og_dims d, og_real dt, const int* flags, const og_real* vel,
                            og_real* grid, int gridType, int order, og_real strength,
                            int clampMode);
/* the same with orderSpace (1 trilinear, 2 cubic) and orderTrace (1 Euler, 2
 * midpoint); returns -2 / -3 for other orders */
int og_advect_semi_lagrange_hi(og_dims d, og_real dt, const int* flags, const og_real* vel, og_real* grid,
                               int gridType, int order, og_real strength, int clampMode, int orderSpace,
                               int orderTrace);
/* individual steps, exposed for unit tests */
void og_sl_scalar(og_dims d, const og_real* vel, og_real* dst, const og_real* src,
                  og_real dt, int ncomp);
void og_sl_mac(og_dims d, const og_real* vel, og_real* dst, const og_real* src, og_real dt);
void og_sl_scalar_hi(og_dims d, const og_real* vel, og_real* dst, const og_real* src, og_real dt, int ncomp,
                     int orderSpace, int orderTrace);
void og_sl_mac_hi(og_dims d, const og_real* vel, og_real* dst, const og_real* src, og_real dt, int orderSpace,
                  int orderTrace);
void og_mac_cormack_correct(og_dims d, const int* flags, og_real* dst, const og_real* old,
                            const og_real* fwd, const og_real* bwd, og_real strength,
                            int gridType);
void og_mac_cormack_clamp(og_dims d, const int* flags, const og_real* vel, og_real* dst,
                          const og_real* orig, const og_real* fwd, og_real dt,
                          int clampMode, int ncomp);
void og_mac_cormack_clamp_mac(og_dims d, const int* flags, const og_real* vel, og_real* dst,
                              const og_real* orig, const og_real* fwd, og_real dt,
                              int clampMode);
void og_apply_outflow_bc(og_dims d, const int* flags, og_real* vel, const og_real* velPrev,
                         og_real dt);

/* ---- forces / boundary conditions (extforces.cpp:72-87,186-217) ---- */
void og_set_wall_bcs(og_dims d, const int* flags, og_real* vel, const og_real* obvel);
void og_add_buoyancy(og_dims d, og_real dt, const int* flags, const og_real* density,
                     og_real* vel, og_real gx, og_real gy, og_real gz,
                     og_real coefficient, int scale);

/* ---- pressure (pressure.cpp, conjugategrad.h/.cpp) ---- */
/* returns cnt, *sum gets the double divergence sum */
int og_make_rhs(og_dims d, const int* flags, og_real* rhs, const og_real* vel, double* sum);
void og_make_laplace(og_dims d, const int* flags, og_real* A0, og_real* Ai, og_real* Aj,
                     og_real* Ak);
void og_apply_matrix(og_dims d, const int* flags, og_real* dst, const og_real* src,
                     const og_real* A0, const og_real* Ai, const og_real* Aj,
                     const og_real* Ak);
int og_find_fix_cell(og_dims d, const int* flags);
void og_fix_pressure(og_dims d, int fixPidx, og_real value, og_real* rhs, og_real* A0,
                     og_real* Ai, og_real* Aj, og_real* Ak);
void og_correct_velocity(og_dims d, const int* flags, og_real* vel, const og_real* pressure);
/* z = M^-1 r with the modified incomplete Cholesky preconditioner of A
 * (conjugategrad.cpp:66-97,135-159); z must be zero-initialised */
void og_mic_precondition(og_dims d, const int* flags, const og_real* A0, const og_real* Ai,
                         const og_real* Aj, const og_real* Ak, const og_real* r, og_real* z);
/* full solvePressure: rhs -> solve -> correct.  rhsOut (optional) gets the rhs.
 * Returns 0, or 1 if the reference would raise the divergence error. */
int og_solve_pressure(og_dims d, og_real* vel, og_real* pressure, const int* flags,
                      const og_pressure_opts* opts, og_solve_stats* st, og_real* rhsOut);
/* the system solve only (rhs given), pressure is overwritten */
int og_solve_pressure_system(og_dims d, og_real* rhs, og_real* pressure, const int* flags,
                             const og_pressure_opts* opts, og_solve_stats* st);
void og_pressure_opts_default(og_pressure_opts* o);

/* ---- second-order obstacle and free-surface boundaries (pressure.cpp:32-84,
 *      111-214,455-476, conjugategrad.h:154-187, extforces.cpp:220-315) ---- */
/* the optional inputs of solvePressure (pressure.cpp:480-495); a NULL
 * pointer is an absent argument */
typedef struct {
	const og_real* phi;          /* Real: fluid level set (ghost fluid) */
	const og_real* perCellCorr;  /* Real: per-cell divergence correction */
	const og_real* fractions;    /* MAC (AoS): open face fractions */
	const og_real* obvel;        /* MAC (AoS): obstacle velocity, used with fractions */
	const og_real* curv;         /* Real: curvature, used with phi (surface tension) */
	og_real gfClamp;             /* 1e-4 */
	og_real surfTens;            /* 0 */
} og_pressure_bnd;
void og_pressure_bnd_default(og_pressure_bnd* b);
/* ghostFluidHelper (pressure.cpp:114-126) for the cell value phi_in and its
 * neighbour's phi_out */
og_real og_ghost_fluid_factor(og_real phi_in, og_real phi_out, og_real gfClamp);
int og_make_rhs_bnd(og_dims d, const int* flags, og_real* rhs, const og_real* vel, const og_pressure_bnd* b,
                    double* sum);
/* MakeLaplaceMatrix with fractions, then ApplyGhostFluidDiagonal with phi */
void og_make_laplace_bnd(og_dims d, const int* flags, og_real* A0, og_real* Ai, og_real* Aj, og_real* Ak,
                         const og_pressure_bnd* b);
void og_correct_velocity_bnd(og_dims d, const int* flags, og_real* vel, const og_real* pressure,
                             const og_pressure_bnd* b);
int og_solve_pressure_system_bnd(og_dims d, og_real* rhs, og_real* pressure, const int* flags,
                                 const og_pressure_opts* opts, const og_pressure_bnd* b, og_solve_stats* st);
int og_solve_pressure_bnd(og_dims d, og_real* vel, og_real* pressure, const int* flags,
                          const og_pressure_opts* opts, const og_pressure_bnd* b, og_solve_stats* st,
                          og_real* rhsOut);
/* setWallBcs with fractions and phiObs: KnSetWallBcsFrac into a copy, then swap */
void og_set_wall_bcs_frac(og_dims d, const int* flags, og_real* vel, const og_real* phiObs);

/* ---- other GridCg clients (conjugategrad.cpp:347-423, plugin/waves.cpp:69-141) ---- */
void og_diffusion_matrix(og_dims d, const int* flags, og_real alpha, og_real* A0, og_real* Ai, og_real* Aj,
                         og_real* Ak);
/* ncomp 1: Real grid; 3: Vec3/MAC grid (AoS) solved per component */
void og_cg_solve_diffusion(og_dims d, const int* flags, og_real* grid, int ncomp, og_real alpha,
                           og_real cgMaxIterFac, og_real cgAccuracy, og_solve_stats* st);
void og_make_rhs_we(og_dims d, const int* flags, og_real* rhs, const og_real* ut, const og_real* utm1, og_real s,
                    int crankNic);
void og_cg_solve_we(og_dims d, const int* flags, og_real* ut, og_real* utm1, og_real* out, int crankNic,
                    og_real cSqr, og_real cgMaxIterFac, og_real cgAccuracy, og_real dt, og_solve_stats* st);

/* ---- multigrid preconditioner after Dick et al. (multigrid.cpp) ---- */
typedef struct og_mg og_mg;
og_mg* og_mg_create(og_dims d);
void og_mg_destroy(og_mg* mg);
void og_mg_set_a(og_mg* mg, const og_real* A0, const og_real* Ai, const og_real* Aj,
                 const og_real* Ak);
void og_mg_set_coarsest_accuracy(og_mg* mg, og_real acc);
/* one V-cycle on A x = rhs from x = 0; returns the level-0 residual norm */
og_real og_mg_vcycle(og_mg* mg, const og_real* rhs, og_real* x);
int og_mg_num_levels(const og_mg* mg);
void og_mg_level_size(const og_mg* mg, int l, int* sx, int* sy, int* sz);
/* vertex types of a level (0 inactive, 1 active, 2 active trivial); returns count */
int og_mg_level_types(const og_mg* mg, int l, int8_t* out);
/* copies level l's stencil storage (n*4 / n*3 on level 0, n*14 / n*5 on
 * coarser levels, row-major per vertex); returns the entries per vertex */
int og_mg_level_matrix(const og_mg* mg, int l, og_real* out);

/* ---- shapes (shapes.cpp:41-84,151-154,240,324-329) ---- */
/* shape: 1 box (a=p0, b=p1), 2 sphere (a=center, b=scale, r=radius),
 *        3 cylinder (a=center, b=z, r=radius)
 * gridType OG_TYPE_INT: value[0] cast to int; REAL: value[0]; VEC3/MAC: value[0..2] */
void og_apply_shape(og_dims d, int shape, const og_real* a, const og_real* b, og_real r,
                    void* grid, int gridType, const og_real* value, const int* respectFlags);

#ifdef __cplusplus
}
#endif
#endif
