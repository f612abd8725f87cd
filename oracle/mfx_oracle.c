/*
 * CPU oracle: plain-C restatement of the reference's CPU plugins for the
 * per-step grid path.  TEST INFRASTRUCTURE ONLY (see mfx_oracle.h): it is the
 * checker for the HIP path and never part of the product.
 *
 * Arithmetic follows the reference's C++ expressions operation by operation,
 * including its mixed float/double promotions (a float lvalue "+=" a double
 * expression is evaluated in double and rounded once, Vec3*double narrows per
 * component, ...), so that with the same summation order the results are
 * bit-identical to the reference's serial FP32 build.  Build with
 * -ffp-contract=off: the reference is compiled for baseline x86-64 (SSE2, no
 * FMA), so no multiply-add is ever fused there.
 */
#include "mfx_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef OG_F64
#define OG_MAX DBL_MAX
#define OG_SQRT sqrt
#define OG_FABS fabs
#define OG_VEC_EPS 1e-10
#else
#define OG_MAX FLT_MAX
#define OG_SQRT sqrtf
#define OG_FABS fabsf
#define OG_VEC_EPS 1e-6f
#endif

/* Planted mistakes for the oracle's own tests (tests/test_oracle_mutants.py):
 * built with -DOG_MUTANT=k, the library carries mistake k of the list below,
 * and the pinning tests must reject it.  The normal builds have OG_MUTANT 0.
 *   1  getAtMACX: y neighbour taken at i+1 instead of i-1
 *   2  getAtMACZ: x neighbour taken at k+1 instead of k-1
 *   3  MacCormackCorrectMAC: no skip for a non-fluid lower x neighbour
 *   4  doClampComponentMAC, clampMode 1: min and max bounds swapped
 *   5  doClampComponent (centered), clampMode 1: min and max bounds swapped
 *   6  SemiLagrangeMAC: component lookup shifted back by half a cell
 *   7  extrapolateVelConvectiveBC: adds vel at the fluid cell, not at low/up
 *   8  extrapolateVelConvectiveBC: bulk x velocity used for every axis
 *   9  KnSetWallBcs: obstacle cell above a fluid cell not set
 *  10  KnSetWallBcs: a stick x neighbour zeroes x instead of y and z
 *  11  getCentered: y average with the lower neighbour
 *  12  doClampComponentMAC, clampMode 2: flag check on the upper neighbour
 *  13  cubicInterp<Real>: a2 with +d1 instead of -d1
 *  14  interpolCubic: support starting at x1 instead of x1-1
 *  15  midpoint trace of centred grids: full step to the midpoint
 *  16  interpolCubicMAC: no half-cell shift of the component's lookup
 *  17  MAC midpoint trace: velocity instead of the advected field
 *  18  interpolMAC: x component without its shifted index
 *  19  cgSolveDiffusion matrix: diagonal scaled by -alpha
 *  20  MakeRhsWE (Crank-Nicolson): x+1 neighbour read at x-1
 *  21  MakeRhs with fractions: the +y face's fraction read at the cell
 *  22  ghostFluidHelper: 1 - alpha instead of 1 - 1/alpha
 *  23  KnSetWallBcsFrac, x face: y gradient with the sides swapped
 *  24  knCorrectVelocityGhostFluid: empty cell's x face scaled by its own pressure */
#ifndef OG_MUTANT
#define OG_MUTANT 0
#endif
#define OG_MUT(k) (OG_MUTANT == (k))

static int g_threads = 1;
void og_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int og_get_threads(void) { return g_threads; }

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for num_threads(g_threads) if(g_threads > 1) schedule(static)")
#else
#define OMP_FOR
#endif

/* ---------------------------------------------------------------------------
 * indexing: idx = i + nx*j + sz*k with sz = 0 in 2D (grid.h:70, grid.cpp:55)
 * ------------------------------------------------------------------------- */
#define SZ(d) ((d).is3d ? (long)(d).nx * (long)(d).ny : 0L)
#define NCELLS(d) ((long)(d).nx * (long)(d).ny * (long)(d).nz)
static inline long IX(og_dims d, int i, int j, int k) { return (long)i + (long)d.nx * j + SZ(d) * k; }

/* KERNEL(bnd=b) ijk domain (kernel.cpp:21-48, codegen_kernel.cpp:117-140):
 * k in [b, nz-b) in 3D, k = 0 in 2D; j in [b, ny-b); i in [b, nx-b). */
#define KMIN(d, b) ((d).is3d ? (b) : 0)
#define KMAX(d, b) ((d).is3d ? (d).nz - (b) : 1)

static inline int isFluid(const int* f, long i) { return f[i] & OG_FLUID; }
static inline int isObs(const int* f, long i) { return f[i] & OG_OBSTACLE; }
static inline int isEmpty(const int* f, long i) { return f[i] & OG_EMPTY; }
static inline int isOutflow(const int* f, long i) { return f[i] & OG_OUTFLOW; }
static inline int isStick(const int* f, long i) { return f[i] & OG_STICK; }

static inline int clampi(int v, int lo, int hi) { /* general.h:137-141 */
	if (v < lo) return lo;
	if (v > hi) return hi;
	return v;
}
static inline og_real clampr(og_real v, og_real lo, og_real hi) {
	if (v < lo) return lo;
	if (v > hi) return hi;
	return v;
}

/* deterministic chunked double reduction helper: chunk = one k-slab (3D) or
 * one j-row (2D); partial sums combined in order. */
typedef double (*og_slab_fn)(og_dims d, long lo, long hi, const void* ctx);

static double reduce_sum(og_dims d, og_slab_fn fn, const void* ctx) {
	long n = NCELLS(d);
	long chunk = d.is3d ? (long)d.nx * d.ny : (long)d.nx;
	long nch = (n + chunk - 1) / chunk;
	double* part = (double*)malloc(sizeof(double) * (size_t)nch);
	OMP_FOR
	for (long c = 0; c < nch; c++) {
		long lo = c * chunk, hi = lo + chunk < n ? lo + chunk : n;
		part[c] = fn(d, lo, hi, ctx);
	}
	double s = 0.0;
	for (long c = 0; c < nch; c++) s += part[c];
	free(part);
	return s;
}

/* ===========================================================================
 * Flags
 * ========================================================================= */

/* FlagGrid::initDomain + initBoundaries (grid.cpp:732-842) */
void og_init_domain(og_dims d, int* flags, int boundaryWidth, const char* wallIn,
                    const char* openIn, const char* inflowIn, const char* outflowIn) {
	int types[6] = {0, 0, 0, 0, 0, 0};
	int set[6] = {0, 0, 0, 0, 0, 0};
	char wall[64], open[64], inflow[64], outflow[64];
	/* "make sure we have at least 6 entries": append six blanks */
	snprintf(wall, sizeof wall, "%s      ", wallIn ? wallIn : "");
	snprintf(open, sizeof open, "%s      ", openIn ? openIn : "");
	snprintf(inflow, sizeof inflow, "%s      ", inflowIn ? inflowIn : "");
	snprintf(outflow, sizeof outflow, "%s      ", outflowIn ? outflowIn : "");
	const char dirs[6] = {'x', 'X', 'y', 'Y', 'z', 'Z'};
	for (int i = 0; i < 6; ++i) {
		for (int s = 0; s < 6; ++s) {
			if (s >= 4 && !d.is3d) continue;
			if (set[s]) continue;
			char c = dirs[s];
			if (open[i] == c) { types[s] = OG_OPEN; set[s] = 1; }
			else if (inflow[i] == c) { types[s] = OG_INFLOW; set[s] = 1; }
			else if (outflow[i] == c) { types[s] = OG_OUTFLOW; set[s] = 1; }
			else if (wall[i] == c) { types[s] = OG_OBSTACLE; set[s] = 1; }
		}
	}
	const int w = boundaryWidth;
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long ix = IX(d, i, j, k);
				flags[ix] = OG_EMPTY; /* setConst(TypeEmpty) */
				if (i <= w) flags[ix] = types[0];
				if (i >= d.nx - 1 - w) flags[ix] = types[1];
				if (j <= w) flags[ix] = types[2];
				if (j >= d.ny - 1 - w) flags[ix] = types[3];
				if (d.is3d) {
					if (k <= w) flags[ix] = types[4];
					if (k >= d.nz - 1 - w) flags[ix] = types[5];
				}
			}
}

/* FlagGrid::fillGrid (grid.cpp:856-861) */
void og_fill_grid(og_dims d, int* flags, int type) {
	long n = NCELLS(d);
	for (long idx = 0; idx < n; idx++) {
		int f = flags[idx];
		if ((f & OG_OBSTACLE) == 0 && (f & OG_INFLOW) == 0 && (f & OG_OUTFLOW) == 0 &&
		    (f & OG_OPEN) == 0)
			flags[idx] = (f & ~(OG_EMPTY | OG_FLUID)) | type;
	}
}

/* setOpenBound (extforces.cpp:90-130); returns void, invalid characters are
 * rejected by the caller (convertDescToVec raises there). */
void og_set_open_bound(og_dims d, int* flags, int bWidth, const char* openBound, int type) {
	if (!openBound || openBound[0] == 0) return;
	int lox = 0, loy = 0, loz = 0, upx = 0, upy = 0, upz = 0;
	for (const char* p = openBound; *p; p++) {
		switch (*p) {
		case 'x': lox = 1; break;
		case 'y': loy = 1; break;
		case 'z': loz = 1; break;
		case 'X': upx = 1; break;
		case 'Y': upy = 1; break;
		case 'Z': upz = 1; break;
		default: return;
		}
	}
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long ix = IX(d, i, j, k);
				int loX = lox && i <= bWidth;
				int loY = loy && j <= bWidth;
				int upX = upx && i >= d.nx - bWidth - 1;
				int upY = upy && j >= d.ny - bWidth - 1;
				int innerI = i > bWidth && i < d.nx - bWidth - 1;
				int innerJ = j > bWidth && j < d.ny - bWidth - 1;
				if ((!d.is3d) && (loX || upX || loY || upY)) {
					if ((loX || upX || innerI) && (loY || upY || innerJ) && isObs(flags, ix))
						flags[ix] = type;
				} else {
					int loZ = loz && k <= bWidth;
					int upZ = upz && k >= d.nz - bWidth - 1;
					int innerK = k > bWidth && k < d.nz - bWidth - 1;
					if (loX || upX || loY || upY || loZ || upZ) {
						if ((loX || upX || innerI) && (loY || upY || innerJ) &&
						    (loZ || upZ || innerK) && isObs(flags, ix))
							flags[ix] = type;
					}
				}
			}
}

/* resetOutflow without particles (extforces.cpp:133-160) */
void og_reset_outflow(og_dims d, int* flags, og_real* phi, og_real* real) {
	long n = NCELLS(d);
	for (long ix = 0; ix < n; ix++) {
		if (isOutflow(flags, ix)) {
			flags[ix] = (flags[ix] | OG_EMPTY) & ~OG_FLUID;
			if (phi) phi[ix] = (og_real)0.5;
			if (real) real[ix] = 0;
		}
	}
}

/* ===========================================================================
 * Interpolation (util/interpol.h:50-90).  `es` is the element stride (1 for
 * Real grids, 3 for one component of an AoS Vec3 grid).
 * ========================================================================= */
static inline og_real interp_core(og_dims d, const og_real* data, long es, og_real posx,
                                  og_real posy, og_real posz) {
	/* BUILD_INDEX */
	og_real px = posx - 0.5f, py = posy - 0.5f, pz = posz - 0.5f;
	int xi = (int)px;
	int yi = (int)py;
	int zi = (int)pz;
	og_real s1 = px - (og_real)xi, s0 = 1. - s1;
	og_real t1 = py - (og_real)yi, t0 = 1. - t1;
	og_real f1 = pz - (og_real)zi, f0 = 1. - f1;
	if (px < 0.) { xi = 0; s0 = 1.0; s1 = 0.0; }
	if (py < 0.) { yi = 0; t0 = 1.0; t1 = 0.0; }
	if (pz < 0.) { zi = 0; f0 = 1.0; f1 = 0.0; }
	if (xi >= d.nx - 1) { xi = d.nx - 2; s0 = 0.0; s1 = 1.0; }
	if (yi >= d.ny - 1) { yi = d.ny - 2; t0 = 0.0; t1 = 1.0; }
	if (d.nz > 1) { if (zi >= d.nz - 1) { zi = d.nz - 2; f0 = 0.0; f1 = 1.0; } }
	const long X = es, Y = (long)d.nx * es, Z = SZ(d) * es;
	const long idx = ((long)xi + (long)d.nx * yi + SZ(d) * zi) * es;
	return ((data[idx] * t0 + data[idx + Y] * t1) * s0 +
	        (data[idx + X] * t0 + data[idx + X + Y] * t1) * s1) * f0 +
	       ((data[idx + Z] * t0 + data[idx + Y + Z] * t1) * s0 +
	        (data[idx + X + Z] * t0 + data[idx + X + Y + Z] * t1) * s1) * f1;
}

og_real og_interpol(og_dims d, const og_real* data, og_real px, og_real py, og_real pz) {
	return interp_core(d, data, 1, px, py, pz);
}
og_real og_interpol_component(og_dims d, const og_real* v, int c, og_real px, og_real py,
                              og_real pz) {
	return interp_core(d, v + c, 3, px, py, pz);
}

/* cubicInterp<T> (util/interpolHigh.h:22-39), the Catmull-Rom segment
 * between points[1] and points[2].  For T = Real the coefficient
 * expressions are evaluated in double and stored as Real (cubic_real); for
 * T = Vec3 every operator rounds its component (Vec3 * double narrows,
 * Vec3 +- Vec3 is a Real operation, vectorbase.h:235-279; cubic_vec).  The
 * polynomial itself is evaluated in Real either way. */
static inline og_real cubic_real(og_real t, const og_real* p) {
	const og_real d0 = (og_real)((p[2] - p[0]) * 0.5);
	const og_real d1 = (og_real)((p[3] - p[1]) * 0.5);
	const og_real dk = p[2] - p[1];
	const og_real a0 = p[1], a1 = d0;
	const og_real a2 = (og_real)(3.0 * dk - 2.0 * d0 + (OG_MUT(13) ? d1 : -d1));
	const og_real a3 = (og_real)(-2.0 * dk + d0 + d1);
	const og_real sq = t * t, cu = sq * t;
	return a3 * cu + a2 * sq + a1 * t + a0;
}
static inline og_real cubic_vec(og_real t, const og_real* p) {
	const og_real d0 = (og_real)((og_real)(p[2] - p[0]) * 0.5);
	const og_real d1 = (og_real)((og_real)(p[3] - p[1]) * 0.5);
	const og_real dk = p[2] - p[1];
	const og_real a0 = p[1], a1 = d0;
	og_real a2 = (og_real)(3.0 * dk) - (og_real)(2.0 * d0);
	a2 = a2 - d1;
	og_real a3 = (og_real)(-2.0 * dk) + d0;
	a3 = a3 + d1;
	const og_real sq = t * t, cu = sq * t;
	return a3 * cu + a2 * sq + a1 * t + a0;
}

/* interpolCubic<T> / interpolCubic2D<T> (util/interpolHigh.h:41-170) on one
 * component (element stride es); positions whose 4x4(x4) support leaves the
 * grid fall back to the trilinear interpol<T> (:58-60, :97-99). */
static og_real cubic_core(og_dims d, const og_real* data, long es, og_real posx, og_real posy, og_real posz,
                          int vecRules) {
	og_real (*cub)(og_real, const og_real*) = vecRules ? cubic_vec : cubic_real;
	const og_real px = posx - 0.5f, py = posy - 0.5f;
	const int x1 = (int)px, y1 = (int)py;
	const int x0 = x1 - (OG_MUT(14) ? 0 : 1), x3 = x0 + 3, y0 = y1 - 1, y3 = y1 + 2;
	if (!d.is3d) {
		if (x0 < 0 || y0 < 0 || x3 >= d.nx || y3 >= d.ny) return interp_core(d, data, es, posx, posy, posz);
		const og_real xt = px - (og_real)x1, yt = py - (og_real)y1;
		og_real rows[4];
		for (int r = 0; r < 4; r++) {
			og_real pts[4];
			for (int q = 0; q < 4; q++) pts[q] = data[es * ((long)(x0 + q) + (long)d.nx * (y0 + r))];
			rows[r] = cub(xt, pts);
		}
		return cub(yt, rows);
	}
	const og_real pz = posz - 0.5f;
	const int z1 = (int)pz, z0 = z1 - 1, z3 = z1 + 2;
	if (x0 < 0 || y0 < 0 || z0 < 0 || x3 >= d.nx || y3 >= d.ny || z3 >= d.nz)
		return interp_core(d, data, es, posx, posy, posz);
	const og_real xt = px - (og_real)x1, yt = py - (og_real)y1, zt = pz - (og_real)z1;
	og_real slabs[4];
	for (int c = 0; c < 4; c++) {
		og_real rows[4];
		for (int r = 0; r < 4; r++) {
			og_real pts[4];
			for (int q = 0; q < 4; q++) pts[q] = data[es * IX(d, x0 + q, y0 + r, z0 + c)];
			rows[r] = cub(xt, pts);
		}
		slabs[c] = cub(yt, rows);
	}
	return cub(zt, slabs);
}

og_real og_interpol_cubic(og_dims d, const og_real* data, og_real px, og_real py, og_real pz) {
	return cubic_core(d, data, 1, px, py, pz, 0);
}
og_real og_interpol_cubic_component(og_dims d, const og_real* v, int c, og_real px, og_real py, og_real pz) {
	return cubic_core(d, v + c, 3, px, py, pz, 1);
}

/* interpolMAC (util/interpol.h:121-157): every component of a MAC grid at
 * pos, each with the lookup shifted by half a cell along its own axis
 * (BUILD_INDEX_SHIFT); out[3] */
void og_interpol_mac(og_dims d, const og_real* v, og_real px0, og_real py0, og_real pz0, og_real* out) {
	/* BUILD_INDEX */
	og_real px = px0 - 0.5f, py = py0 - 0.5f, pz = pz0 - 0.5f;
	int xi = (int)px, yi = (int)py, zi = (int)pz;
	og_real s1 = px - (og_real)xi, s0 = 1. - s1;
	og_real t1 = py - (og_real)yi, t0 = 1. - t1;
	og_real f1 = pz - (og_real)zi, f0 = 1. - f1;
	if (px < 0.) { xi = 0; s0 = 1.0; s1 = 0.0; }
	if (py < 0.) { yi = 0; t0 = 1.0; t1 = 0.0; }
	if (pz < 0.) { zi = 0; f0 = 1.0; f1 = 0.0; }
	if (xi >= d.nx - 1) { xi = d.nx - 2; s0 = 0.0; s1 = 1.0; }
	if (yi >= d.ny - 1) { yi = d.ny - 2; t0 = 0.0; t1 = 1.0; }
	if (d.nz > 1) { if (zi >= d.nz - 1) { zi = d.nz - 2; f0 = 0.0; f1 = 1.0; } }
	/* shifted coordinates */
	int sxi = (int)px0, syi = (int)py0, szi = (int)pz0;
	og_real ss1 = px0 - (og_real)sxi, ss0 = 1. - ss1;
	og_real st1 = py0 - (og_real)syi, st0 = 1. - st1;
	og_real sf1 = pz0 - (og_real)szi, sf0 = 1. - sf1;
	if (px0 < 0) { sxi = 0; ss0 = 1.0; ss1 = 0.0; }
	if (py0 < 0) { syi = 0; st0 = 1.0; st1 = 0.0; }
	if (pz0 < 0) { szi = 0; sf0 = 1.0; sf1 = 0.0; }
	if (sxi >= d.nx - 1) { sxi = d.nx - 2; ss0 = 0.0; ss1 = 1.0; }
	if (syi >= d.ny - 1) { syi = d.ny - 2; st0 = 0.0; st1 = 1.0; }
	if (d.nz > 1) { if (szi >= d.nz - 1) { szi = d.nz - 2; sf0 = 0.0; sf1 = 1.0; } }
	const long X = 3, Y = 3L * d.nx, Z = 3 * SZ(d);
	{
		const og_real* r = v + 3 * (((long)zi * d.ny + yi) * d.nx + (OG_MUT(18) ? xi : sxi));
		out[0] = f0 * ((r[0] * t0 + r[Y] * t1) * ss0 + (r[X] * t0 + r[X + Y] * t1) * ss1) +
		         f1 * ((r[Z] * t0 + r[Z + Y] * t1) * ss0 + (r[X + Z] * t0 + r[X + Y + Z] * t1) * ss1);
	}
	{
		const og_real* r = v + 3 * (((long)zi * d.ny + syi) * d.nx + xi) + 1;
		out[1] = f0 * ((r[0] * st0 + r[Y] * st1) * s0 + (r[X] * st0 + r[X + Y] * st1) * s1) +
		         f1 * ((r[Z] * st0 + r[Z + Y] * st1) * s0 + (r[X + Z] * st0 + r[X + Y + Z] * st1) * s1);
	}
	{
		const og_real* r = v + 3 * (((long)szi * d.ny + yi) * d.nx + xi) + 2;
		out[2] = sf0 * ((r[0] * t0 + r[Y] * t1) * s0 + (r[X] * t0 + r[X + Y] * t1) * s1) +
		         sf1 * ((r[Z] * t0 + r[Z + Y] * t1) * s0 + (r[X + Z] * t0 + r[X + Y + Z] * t1) * s1);
	}
}

/* MACGrid accessors (grid.h:424-470).  0.5*(float) is evaluated in double and
 * narrowed, which equals the float product (scaling by a power of two). */
static inline void getCentered(og_dims d, const og_real* v, int i, int j, int k, og_real* o) {
	long ix = IX(d, i, j, k);
	o[0] = (og_real)(0.5 * (v[3 * ix] + v[3 * (ix + 1)]));
	o[1] = (og_real)(0.5 * (v[3 * ix + 1] + v[3 * (ix + (OG_MUT(11) ? -d.nx : d.nx)) + 1]));
	o[2] = 0.;
	if (d.is3d) o[2] = (og_real)(0.5 * (v[3 * ix + 2] + v[3 * (ix + SZ(d)) + 2]));
}
static inline void getAtMACX(og_dims d, const og_real* v, int i, int j, int k, og_real* o) {
	long ix = IX(d, i, j, k), Y = d.nx, Z = SZ(d);
	const long m1 = OG_MUT(1) ? 1 : -1;
	o[0] = v[3 * ix];
	o[1] = (og_real)(0.25 * (v[3 * ix + 1] + v[3 * (ix + m1) + 1] + v[3 * (ix + Y) + 1] +
	                         v[3 * (ix + Y + m1) + 1]));
	o[2] = 0.;
	if (d.is3d)
		o[2] = (og_real)(0.25 * (v[3 * ix + 2] + v[3 * (ix - 1) + 2] + v[3 * (ix + Z) + 2] +
		                         v[3 * (ix + Z - 1) + 2]));
}
static inline void getAtMACY(og_dims d, const og_real* v, int i, int j, int k, og_real* o) {
	long ix = IX(d, i, j, k), Y = d.nx, Z = SZ(d);
	o[0] = (og_real)(0.25 * (v[3 * ix] + v[3 * (ix - Y)] + v[3 * (ix + 1)] + v[3 * (ix + 1 - Y)]));
	o[1] = v[3 * ix + 1];
	o[2] = 0.;
	if (d.is3d)
		o[2] = (og_real)(0.25 * (v[3 * ix + 2] + v[3 * (ix - Y) + 2] + v[3 * (ix + Z) + 2] +
		                         v[3 * (ix + Z - Y) + 2]));
}
/* note: not guarded by is3D in the reference, Z = 0 in 2D */
static inline void getAtMACZ(og_dims d, const og_real* v, int i, int j, int k, og_real* o) {
	long ix = IX(d, i, j, k), Y = d.nx, Z = SZ(d);
	const long mz = OG_MUT(2) ? Z : -Z;
	o[0] = (og_real)(0.25 * (v[3 * ix] + v[3 * (ix + mz)] + v[3 * (ix + 1)] + v[3 * (ix + 1 + mz)]));
	o[1] = (og_real)(0.25 * (v[3 * ix + 1] + v[3 * (ix - Z) + 1] + v[3 * (ix + Y) + 1] +
	                         v[3 * (ix + Y - Z) + 1]));
	o[2] = v[3 * ix + 2];
}

/* ===========================================================================
 * Semi-Lagrangian advection
 * ========================================================================= */

/* one value of a centred grid at pos: interpol (orderSpace 1) or
 * interpolCubic (orderSpace 2) with Real (ncomp 1) or Vec3 (ncomp 3) rules */
static inline og_real lookup_centered(og_dims d, const og_real* src, int ncomp, int c, int orderSpace, og_real px,
                                      og_real py, og_real pz) {
	if (orderSpace == 2) return cubic_core(d, src + c, ncomp, px, py, pz, ncomp == 3);
	return interp_core(d, src + c, ncomp, px, py, pz);
}

/* SemiLagrange<T> with orderTrace 1 or 2 (explicit midpoint) and orderSpace
 * 1 or 2 (advection.cpp:25-41; grid.h:146-151 getInterpolatedHi) */
void og_sl_scalar_hi(og_dims d, const og_real* vel, og_real* dst, const og_real* src, og_real dt, int ncomp,
                     int orderSpace, int orderTrace) {
	const int b = 1;
	OMP_FOR
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				og_real v[3], p0[3] = {(og_real)((float)i + 0.5f), (og_real)((float)j + 0.5f),
				                       (og_real)((float)k + 0.5f)};
				og_real pos[3];
				getCentered(d, vel, i, j, k, v);
				if (orderTrace == 2) {
					/* p1 = p0 - getCentered*dt*0.5; p2 = p0 - getInterpolated(p1)*dt */
					og_real p1[3], u[3];
					for (int a = 0; a < 3; a++)
						p1[a] = p0[a] - (og_real)((og_real)(v[a] * dt) * (OG_MUT(15) ? 1.0 : 0.5));
					og_interpol_mac(d, vel, p1[0], p1[1], p1[2], u);
					for (int a = 0; a < 3; a++) pos[a] = p0[a] - u[a] * dt;
				} else {
					for (int a = 0; a < 3; a++) pos[a] = p0[a] - v[a] * dt;
				}
				long ix = IX(d, i, j, k);
				for (int c = 0; c < ncomp; c++)
					dst[ncomp * ix + c] = lookup_centered(d, src, ncomp, c, orderSpace, pos[0], pos[1], pos[2]);
			}
}

/* component c of a MAC grid at pos: interpolComponent<c> (orderSpace 1) or
 * interpolCubicMAC(pos)[c] = interpolCubic<Vec3>(pos + 0.5 e_c)[c], which is 0
 * for c = 2 in 2D (grid.h:268-273, interpolHigh.h:172-179) */
static inline og_real lookup_mac(og_dims d, const og_real* src, int c, int orderSpace, const og_real* pos) {
	if (orderSpace == 2) {
		if (c == 2 && !d.is3d) return 0;
		og_real q[3] = {pos[0], pos[1], pos[2]};
		if (!OG_MUT(16)) q[c] = q[c] + 0.5f;
		return cubic_core(d, src + c, 3, q[0], q[1], q[2], 1);
	}
	return interp_core(d, src + c, 3, pos[0] - ((OG_MUT(6) && c == 0) ? 0.5f : 0.f), pos[1], pos[2]);
}

/* SemiLagrangeMAC with orderTrace 1 or 2 and orderSpace 1 or 2
 * (advection.cpp:44-77).  The midpoint trace takes the face velocity and
 * the interpolated velocity from src, the field being advected (:59-73). */
void og_sl_mac_hi(og_dims d, const og_real* vel, og_real* dst, const og_real* src, og_real dt, int orderSpace,
                  int orderTrace) {
	const int b = 1;
	OMP_FOR
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				const og_real hi = (og_real)((float)i + 0.5f), hj = (og_real)((float)j + 0.5f),
				              hk = (og_real)((float)k + 0.5f);
				og_real out[3];
				for (int c = 0; c < 3; c++) {
					og_real v[3], pos[3];
					const og_real* fld = (orderTrace == 2 && !OG_MUT(17)) ? src : vel;
					if (c == 0) getAtMACX(d, fld, i, j, k, v);
					else if (c == 1) getAtMACY(d, fld, i, j, k, v);
					else getAtMACZ(d, fld, i, j, k, v);
					if (orderTrace == 2) {
						og_real f0[3] = {hi, hj, hk};
						f0[c] = (og_real)(c == 0 ? i : (c == 1 ? j : k));
						og_real p1[3], u[3];
						for (int a = 0; a < 3; a++) p1[a] = f0[a] - (og_real)((og_real)(v[a] * dt) * 0.5);
						og_interpol_mac(d, OG_MUT(17) ? vel : src, p1[0], p1[1], p1[2], u);
						pos[0] = hi - u[0] * dt;
						pos[1] = hj - u[1] * dt;
						pos[2] = hk - u[2] * dt;
					} else {
						pos[0] = hi - v[0] * dt;
						pos[1] = hj - v[1] * dt;
						pos[2] = hk - v[2] * dt;
					}
					out[c] = lookup_mac(d, src, c, orderSpace, pos);
				}
				long ix = IX(d, i, j, k);
				dst[3 * ix] = out[0];
				dst[3 * ix + 1] = out[1];
				dst[3 * ix + 2] = out[2];
			}
}

/* SemiLagrange<T> / SemiLagrangeMAC with orderTrace 1 and orderSpace 1 */
void og_sl_scalar(og_dims d, const og_real* vel, og_real* dst, const og_real* src, og_real dt, int ncomp) {
	og_sl_scalar_hi(d, vel, dst, src, dt, ncomp, 1, 1);
}
void og_sl_mac(og_dims d, const og_real* vel, og_real* dst, const og_real* src, og_real dt) {
	og_sl_mac_hi(d, vel, dst, src, dt, 1, 1);
}

/* MacCormackCorrect<T> (advection.cpp:81-91, all idx) and MacCormackCorrectMAC
 * (advection.cpp:94-115, all ijk, isMAC=true).  Real and MAC: the float
 * destination "+=" a double expression; centered Vec3: the double product is
 * narrowed per component first (Vec3 * double), then added in float. */
void og_mac_cormack_correct(og_dims d, const int* flags, og_real* dst, const og_real* old,
                            const og_real* fwd, const og_real* bwd, og_real strength,
                            int gridType) {
	long n = NCELLS(d);
	if (gridType == OG_TYPE_REAL) {
		for (long ix = 0; ix < n; ix++) {
			dst[ix] = fwd[ix];
			if (isFluid(flags, ix)) dst[ix] += strength * 0.5 * (old[ix] - bwd[ix]);
		}
	} else if (gridType == OG_TYPE_VEC3) {
		for (long ix = 0; ix < n; ix++) {
			for (int c = 0; c < 3; c++) dst[3 * ix + c] = fwd[3 * ix + c];
			if (isFluid(flags, ix))
				for (int c = 0; c < 3; c++) {
					og_real t = (og_real)(strength * 0.5 * (og_real)(old[3 * ix + c] - bwd[3 * ix + c]));
					dst[3 * ix + c] += t;
				}
		}
	} else {
		for (int k = 0; k < d.nz; k++)
			for (int j = 0; j < d.ny; j++)
				for (int i = 0; i < d.nx; i++) {
					long ix = IX(d, i, j, k);
					int skip[3] = {0, 0, 0};
					if (!isFluid(flags, ix)) skip[0] = skip[1] = skip[2] = 1;
					if (!OG_MUT(3) && (i > 0) && (!isFluid(flags, IX(d, i - 1, j, k)))) skip[0] = 1;
					if ((j > 0) && (!isFluid(flags, IX(d, i, j - 1, k)))) skip[1] = 1;
					if ((k > 0) && (!isFluid(flags, IX(d, i, j, k - 1)))) skip[2] = 1;
					for (int c = 0; c < 3; ++c) {
						if (skip[c]) dst[3 * ix + c] = fwd[3 * ix + c];
						else
							dst[3 * ix + c] = fwd[3 * ix + c] +
							                  strength * 0.5 * (old[3 * ix + c] - bwd[3 * ix + c]);
					}
				}
	}
}

#define CHECKFLAG(fl, ix) ((fl)[ix] & (OG_FLUID | OG_EMPTY))

/* MacCormackClamp<T> + doClampComponent<T> (advection.cpp:144-186,241-267);
 * ncomp 1 (Real) or 3 (centered Vec3, component-wise min/max, any component
 * out of range reverts the whole vector in clampMode 2). */
void og_mac_cormack_clamp(og_dims d, const int* flags, const og_real* vel, og_real* dst,
                          const og_real* orig, const og_real* fwd, og_real dt, int clampMode,
                          int ncomp) {
	const int b = 1;
	const int gux = d.nx - 1, guy = d.ny - 1, guz = d.nz - 1;
	OMP_FOR
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				long ix = IX(d, i, j, k);
				og_real dval[3], fv[3];
				for (int c = 0; c < ncomp; c++) {
					dval[c] = dst[ncomp * ix + c];
					fv[c] = fwd[ncomp * ix + c];
				}
				og_real vc[3];
				getCentered(d, vel, i, j, k, vc);
				og_real vdt[3] = {vc[0] * dt, vc[1] * dt, vc[2] * dt};
				og_real pos[3] = {(og_real)i, (og_real)j, (og_real)k};
				/* doClampComponent */
				og_real minv[3] = {OG_MAX, OG_MAX, OG_MAX}, maxv[3] = {-OG_MAX, -OG_MAX, -OG_MAX};
				int haveFl = 0;
				int P[2][3];
				int numPos = 1;
				for (int a = 0; a < 3; a++) P[0][a] = (int)(pos[a] - vdt[a]);
				if (clampMode == 1) {
					numPos = 2;
					for (int a = 0; a < 3; a++) P[1][a] = (int)(pos[a] + vdt[a]);
				}
				for (int l = 0; l < numPos; ++l) {
					const int i0 = clampi(P[l][0], 0, gux - 1);
					const int j0 = clampi(P[l][1], 0, guy - 1);
					const int k0 = clampi(P[l][2], 0, (d.is3d ? (guz - 1) : 1));
					const int i1 = i0 + 1, j1 = j0 + 1, k1 = (d.is3d ? (k0 + 1) : k0);
					int cs[8][3] = {{i0, j0, k0}, {i1, j0, k0}, {i0, j1, k0}, {i1, j1, k0},
					                {i0, j0, k1}, {i1, j0, k1}, {i0, j1, k1}, {i1, j1, k1}};
					int nc = d.is3d ? 8 : 4;
					for (int q = 0; q < nc; q++) {
						long cx = IX(d, cs[q][0], cs[q][1], cs[q][2]);
						if (CHECKFLAG(flags, cx)) {
							for (int c = 0; c < ncomp; c++) {
								og_real val = orig[ncomp * cx + c];
								if (val < minv[c]) minv[c] = val;
								if (val > maxv[c]) maxv[c] = val;
							}
							haveFl = 1;
						}
					}
				}
				og_real res[3];
				if (!haveFl) {
					for (int c = 0; c < ncomp; c++) res[c] = fv[c];
				} else if (clampMode == 1) {
					for (int c = 0; c < ncomp; c++)
						res[c] = OG_MUT(5) ? clampr(dval[c], maxv[c], minv[c]) : clampr(dval[c], minv[c], maxv[c]);
				} else {
					int out = 0;
					for (int c = 0; c < ncomp; c++) out |= (dval[c] < minv[c]) | (dval[c] > maxv[c]);
					for (int c = 0; c < ncomp; c++) res[c] = out ? fv[c] : dval[c];
				}
				if (clampMode == 1) {
					int pf[3], pb[3];
					for (int a = 0; a < 3; a++) {
						og_real base = (og_real)((float)pos[a] + 0.5f);
						pf[a] = (int)(base - vdt[a]);
						pb[a] = (int)(base + vdt[a]);
					}
					if (pf[0] < 0 || pf[1] < 0 || pf[2] < 0 || pb[0] < 0 || pb[1] < 0 || pb[2] < 0 ||
					    pf[0] > gux || pf[1] > guy || ((pf[2] > guz) && d.is3d) || pb[0] > gux ||
					    pb[1] > guy || ((pb[2] > guz) && d.is3d) ||
					    isObs(flags, IX(d, pf[0], pf[1], pf[2])) ||
					    isObs(flags, IX(d, pb[0], pb[1], pb[2]))) {
						for (int c = 0; c < ncomp; c++) res[c] = fv[c];
					}
				}
				for (int c = 0; c < ncomp; c++) dst[ncomp * ix + c] = res[c];
			}
}

/* doClampComponentMAC<c> (advection.cpp:191-235) */
static og_real clamp_comp_mac(og_dims d, const int* flags, int c, og_real dst, const og_real* orig,
                              og_real fwd, const og_real* pos, const og_real* vel, int clampMode) {
	const int gux = d.nx - 1, guy = d.ny - 1, guz = d.nz - 1;
	og_real minv = OG_MAX, maxv = -OG_MAX;
	int P[2][3];
	int numPos = 1;
	for (int a = 0; a < 3; a++) P[0][a] = (int)(pos[a] - vel[a]);
	if (clampMode == 1) {
		numPos = 2;
		for (int a = 0; a < 3; a++) P[1][a] = (int)(pos[a] + vel[a]);
	}
	int o[3] = {(int)pos[0], (int)pos[1], (int)pos[2]};
	int nb[3] = {o[0], o[1], o[2]};
	nb[c] += OG_MUT(12) ? 1 : -1;
	if (clampMode == 2 &&
	    (!(CHECKFLAG(flags, IX(d, o[0], o[1], o[2])) && CHECKFLAG(flags, IX(d, nb[0], nb[1], nb[2])))))
		return fwd;
	for (int l = 0; l < numPos; ++l) {
		const int i0 = clampi(P[l][0], 0, gux - 1);
		const int j0 = clampi(P[l][1], 0, guy - 1);
		const int k0 = clampi(P[l][2], 0, (d.is3d ? (guz - 1) : 0));
		const int i1 = i0 + 1, j1 = j0 + 1, k1 = (d.is3d ? (k0 + 1) : k0);
		int cs[8][3] = {{i0, j0, k0}, {i1, j0, k0}, {i0, j1, k0}, {i1, j1, k0},
		                {i0, j0, k1}, {i1, j0, k1}, {i0, j1, k1}, {i1, j1, k1}};
		int nc = d.is3d ? 8 : 4;
		for (int q = 0; q < nc; q++) {
			og_real val = orig[3 * IX(d, cs[q][0], cs[q][1], cs[q][2]) + c];
			if (val < minv) minv = val;
			if (val > maxv) maxv = val;
		}
	}
	if (clampMode == 1) {
		dst = OG_MUT(4) ? clampr(dst, maxv, minv) : clampr(dst, minv, maxv);
	} else {
		if (dst < minv || dst > maxv) dst = fwd;
	}
	return dst;
}

/* MacCormackClampMAC (advection.cpp:270-287) */
void og_mac_cormack_clamp_mac(og_dims d, const int* flags, const og_real* vel, og_real* dst,
                              const og_real* orig, const og_real* fwd, og_real dt, int clampMode) {
	const int b = 1;
	OMP_FOR
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				long ix = IX(d, i, j, k);
				og_real pos[3] = {(og_real)i, (og_real)j, (og_real)k};
				og_real v[3], vdt[3];
				og_real dx = dst[3 * ix], dy = dst[3 * ix + 1], dz = dst[3 * ix + 2];
				getAtMACX(d, vel, i, j, k, v);
				for (int a = 0; a < 3; a++) vdt[a] = v[a] * dt;
				dx = clamp_comp_mac(d, flags, 0, dx, orig, fwd[3 * ix], pos, vdt, clampMode);
				getAtMACY(d, vel, i, j, k, v);
				for (int a = 0; a < 3; a++) vdt[a] = v[a] * dt;
				dy = clamp_comp_mac(d, flags, 1, dy, orig, fwd[3 * ix + 1], pos, vdt, clampMode);
				if (d.is3d) {
					getAtMACZ(d, vel, i, j, k, v);
					for (int a = 0; a < 3; a++) vdt[a] = v[a] * dt;
					dz = clamp_comp_mac(d, flags, 2, dz, orig, fwd[3 * ix + 2], pos, vdt, clampMode);
				}
				dst[3 * ix] = dx;
				dst[3 * ix + 1] = dy;
				dst[3 * ix + 2] = dz;
			}
}

static inline int inBounds(og_dims d, int i, int j, int k) {
	return i >= 0 && j >= 0 && k >= 0 && i < d.nx && j < d.ny && k < d.nz;
}

/* applyOutflowBC = extrapolateVelConvectiveBC + copyChangedVels
 * (advection.cpp:326-391) */
void og_apply_outflow_bc(og_dims d, const int* flags, og_real* vel, const og_real* velPrev,
                         og_real dt) {
	long n = NCELLS(d);
	og_real* velDst = (og_real*)calloc((size_t)(3 * n), sizeof(og_real));
	double tsd = (double)dt * 4;
	const og_real timeStep = (og_real)((1.0 < tsd) ? tsd : 1.0); /* std::max(1.0, timeStep*4) */
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long ix = IX(d, i, j, k);
				if (!isOutflow(flags, ix)) continue;
				/* getBulkVel */
				og_real avg[3] = {0, 0, 0};
				int count = 0;
				int nmax = d.is3d ? 1 : 0;
				for (int nn = -nmax; nn <= nmax; nn++)
					for (int m = -1; m <= 1; m++)
						for (int l = -1; l <= 1; l++) {
							if (inBounds(d, i + l, j + m, k + nn)) {
								long q = IX(d, i + l, j + m, k + nn);
								if (isFluid(flags, q) || isOutflow(flags, q)) {
									for (int c = 0; c < 3; c++) avg[c] += vel[3 * q + c];
									count++;
								}
							}
						}
				og_real bulk[3];
				for (int c = 0; c < 3; c++) bulk[c] = count > 0 ? avg[c] / (og_real)count : avg[c];
				int dim = d.is3d ? 3 : 2;
				int cur[3] = {i, j, k};
				int cnt = 0;
				og_real* vd = velDst + 3 * ix;
				for (int c = 0; c < dim; c++) {
					int low[3], up[3], flLow[3], flUp[3];
					for (int a = 0; a < 3; a++) low[a] = up[a] = flLow[a] = flUp[a] = cur[a];
					const int bc = OG_MUT(8) ? 0 : c;
					og_real mb = bulk[bc] > (og_real)1.0 ? bulk[bc] : (og_real)1.0; /* max((Real)1.0, bulk) */
					og_real factor = timeStep * mb;
					low[c] = flLow[c] = cur[c] - 1;
					up[c] = flUp[c] = cur[c] + 1;
					for (int dd = 0; dd < 2; dd++) {
						int fromLower = inBounds(d, flLow[0], flLow[1], flLow[2]) &&
						                isFluid(flags, IX(d, flLow[0], flLow[1], flLow[2]));
						int fromUpper = inBounds(d, flUp[0], flUp[1], flUp[2]) &&
						                isFluid(flags, IX(d, flUp[0], flUp[1], flUp[2]));
						if (fromLower || fromUpper) {
							if (fromLower) {
								long q = OG_MUT(7) ? IX(d, flLow[0], flLow[1], flLow[2]) : IX(d, low[0], low[1], low[2]);
								for (int a = 0; a < 3; a++)
									vd[a] += ((vel[3 * ix + a] - velPrev[3 * ix + a]) / factor) + vel[3 * q + a];
								cnt++;
							}
							if (fromUpper) {
								long q = OG_MUT(7) ? IX(d, flUp[0], flUp[1], flUp[2]) : IX(d, up[0], up[1], up[2]);
								for (int a = 0; a < 3; a++)
									vd[a] += ((vel[3 * ix + a] - velPrev[3 * ix + a]) / factor) + vel[3 * q + a];
								cnt++;
							}
							break;
						}
						flLow[c]--;
						flUp[c]++;
					}
				}
				if (cnt > 0)
					for (int a = 0; a < 3; a++) vd[a] /= (og_real)cnt;
			}
	for (long ix = 0; ix < n; ix++)
		if (isOutflow(flags, ix))
			for (int a = 0; a < 3; a++) vel[3 * ix + a] = velDst[3 * ix + a];
	free(velDst);
}

/* advectSemiLagrange -> fnAdvectSemiLagrange (advection.cpp:292-321,406-460)
 * with orderSpace / orderTrace 1 or 2.  Temporaries start at zero like the
 * reference's pooled temp grids (Grid constructor clears). */
int og_advect_semi_lagrange_hi(og_dims d, og_real dt, const int* flags, const og_real* vel, og_real* grid,
                               int gridType, int order, og_real strength, int clampMode, int orderSpace,
                               int orderTrace) {
	if (order != 1 && order != 2) return -1;
	if (orderSpace != 1 && orderSpace != 2) return -2;
	if (orderTrace != 1 && orderTrace != 2) return -3;
	long n = NCELLS(d);
	int nc = (gridType == OG_TYPE_REAL) ? 1 : 3;
	size_t bytes = sizeof(og_real) * (size_t)(nc * n);
	og_real* fwd = (og_real*)calloc(1, bytes);
	if (gridType == OG_TYPE_MAC) og_sl_mac_hi(d, vel, fwd, grid, dt, orderSpace, orderTrace);
	else og_sl_scalar_hi(d, vel, fwd, grid, dt, nc, orderSpace, orderTrace);
	if (order == 1) {
		if (gridType == OG_TYPE_MAC) og_apply_outflow_bc(d, flags, fwd, grid, dt);
		memcpy(grid, fwd, bytes);
		free(fwd);
		return 0;
	}
	og_real* bwd = (og_real*)calloc(1, bytes);
	og_real* ng = (og_real*)calloc(1, bytes);
	if (gridType == OG_TYPE_MAC) {
		og_sl_mac_hi(d, vel, bwd, fwd, -dt, orderSpace, orderTrace);
		og_mac_cormack_correct(d, flags, ng, grid, fwd, bwd, strength, OG_TYPE_MAC);
		og_mac_cormack_clamp_mac(d, flags, vel, ng, grid, fwd, dt, clampMode);
		og_apply_outflow_bc(d, flags, ng, grid, dt);
	} else {
		og_sl_scalar_hi(d, vel, bwd, fwd, -dt, nc, orderSpace, orderTrace);
		og_mac_cormack_correct(d, flags, ng, grid, fwd, bwd, strength, gridType);
		og_mac_cormack_clamp(d, flags, vel, ng, grid, fwd, dt, clampMode, nc);
	}
	memcpy(grid, ng, bytes);
	free(fwd);
	free(bwd);
	free(ng);
	return 0;
}

int og_advect_semi_lagrange(og_dims d, og_real dt, const int* flags, const og_real* vel,
                            og_real* grid, int gridType, int order, og_real strength,
                            int clampMode) {
	return og_advect_semi_lagrange_hi(d, dt, flags, vel, grid, gridType, order, strength, clampMode, 1, 1);
}

/* ===========================================================================
 * Forces and boundary conditions
 * ========================================================================= */

/* KnSetWallBcs (extforces.cpp:186-217), all ijk */
void og_set_wall_bcs(og_dims d, const int* flags, og_real* vel, const og_real* obvel) {
	OMP_FOR
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long ix = IX(d, i, j, k);
				int curFluid = isFluid(flags, ix);
				int curObs = isObs(flags, ix);
				og_real bc[3] = {0., 0., 0.};
				if (!curFluid && !curObs) continue;
				if (obvel) {
					bc[0] = obvel[3 * ix];
					bc[1] = obvel[3 * ix + 1];
					if (d.is3d) bc[2] = obvel[3 * ix + 2];
				}
				og_real* v = vel + 3 * ix;
				if (i > 0 && isObs(flags, IX(d, i - 1, j, k))) v[0] = bc[0];
				if (!OG_MUT(9) && i > 0 && curObs && isFluid(flags, IX(d, i - 1, j, k))) v[0] = bc[0];
				if (j > 0 && isObs(flags, IX(d, i, j - 1, k))) v[1] = bc[1];
				if (j > 0 && curObs && isFluid(flags, IX(d, i, j - 1, k))) v[1] = bc[1];
				if (!d.is3d) {
					v[2] = 0;
				} else {
					if (k > 0 && isObs(flags, IX(d, i, j, k - 1))) v[2] = bc[2];
					if (k > 0 && curObs && isFluid(flags, IX(d, i, j, k - 1))) v[2] = bc[2];
				}
				if (curFluid) {
					if ((i > 0 && isStick(flags, IX(d, i - 1, j, k))) ||
					    (i < d.nx - 1 && isStick(flags, IX(d, i + 1, j, k)))) {
						if (OG_MUT(10)) v[0] = 0;
						else v[1] = v[2] = 0;
					}
					if ((j > 0 && isStick(flags, IX(d, i, j - 1, k))) ||
					    (j < d.ny - 1 && isStick(flags, IX(d, i, j + 1, k))))
						v[0] = v[2] = 0;
					if (d.is3d && ((k > 0 && isStick(flags, IX(d, i, j, k - 1))) ||
					               (k < d.nz - 1 && isStick(flags, IX(d, i, j, k + 1)))))
						v[0] = v[1] = 0;
				}
			}
}

/* addBuoyancy + KnAddBuoyancy (extforces.cpp:72-87).  gridScale is a float
 * even in double builds; the face update is evaluated in double. */
void og_add_buoyancy(og_dims d, og_real dt, const int* flags, const og_real* density, og_real* vel,
                     og_real gx, og_real gy, og_real gz, og_real coefficient, int scale) {
	int mx = d.nx > d.ny ? d.nx : d.ny;
	if (d.nz > mx) mx = d.nz;
	og_real dx = (og_real)(1.0 / mx); /* GridBase::mDx */
	float gridScale = scale ? (float)dx : 1;
	og_real f[3] = {-gx, -gy, -gz};
	for (int c = 0; c < 3; c++) f[c] = f[c] * dt / gridScale * coefficient;
	const int b = 1;
	OMP_FOR
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				long ix = IX(d, i, j, k);
				if (!isFluid(flags, ix)) continue;
				if (isFluid(flags, IX(d, i - 1, j, k)))
					vel[3 * ix] += (0.5 * f[0]) * (density[ix] + density[IX(d, i - 1, j, k)]);
				if (isFluid(flags, IX(d, i, j - 1, k)))
					vel[3 * ix + 1] += (0.5 * f[1]) * (density[ix] + density[IX(d, i, j - 1, k)]);
				if (d.is3d && isFluid(flags, IX(d, i, j, k - 1)))
					vel[3 * ix + 2] += (0.5 * f[2]) * (density[ix] + density[IX(d, i, j, k - 1)]);
			}
}

/* ===========================================================================
 * Pressure
 * ========================================================================= */

void og_pressure_opts_default(og_pressure_opts* o) {
	o->cgAccuracy = (og_real)1e-3;
	o->cgMaxIterFac = (og_real)1.5;
	o->preconditioner = OG_PC_MIC;
	o->enforceCompatibility = 0;
	o->useL2Norm = 0;
	o->zeroPressureFixing = 0;
	o->relTol = 0;
	o->maxIterOverride = 0;
}

/* MakeRhs without fractions/obvel/phi/perCellCorr (pressure.cpp:32-84) */
int og_make_rhs(og_dims d, const int* flags, og_real* rhs, const og_real* vel, double* sumOut) {
	const int b = 1;
	long n = NCELLS(d);
	memset(rhs, 0, sizeof(og_real) * (size_t)n); /* fresh temp grid */
	int cnt = 0;
	double sum = 0;
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				long ix = IX(d, i, j, k);
				if (!isFluid(flags, ix)) { rhs[ix] = 0; continue; }
				og_real set = vel[3 * ix] - vel[3 * IX(d, i + 1, j, k)] + vel[3 * ix + 1] -
				              vel[3 * IX(d, i, j + 1, k) + 1];
				if (d.is3d) set += vel[3 * ix + 2] - vel[3 * IX(d, i, j, k + 1) + 2];
				sum += set;
				cnt++;
				rhs[ix] = set;
			}
	if (sumOut) *sumOut = sum;
	return cnt;
}

/* MakeLaplaceMatrix without fractions (conjugategrad.h:154-171) */
void og_make_laplace(og_dims d, const int* flags, og_real* A0, og_real* Ai, og_real* Aj,
                     og_real* Ak) {
	const int b = 1;
	long n = NCELLS(d);
	memset(A0, 0, sizeof(og_real) * (size_t)n);
	memset(Ai, 0, sizeof(og_real) * (size_t)n);
	memset(Aj, 0, sizeof(og_real) * (size_t)n);
	memset(Ak, 0, sizeof(og_real) * (size_t)n);
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				long ix = IX(d, i, j, k);
				if (!isFluid(flags, ix)) continue;
				if (!isObs(flags, IX(d, i - 1, j, k))) A0[ix] += 1.;
				if (!isObs(flags, IX(d, i + 1, j, k))) A0[ix] += 1.;
				if (!isObs(flags, IX(d, i, j - 1, k))) A0[ix] += 1.;
				if (!isObs(flags, IX(d, i, j + 1, k))) A0[ix] += 1.;
				if (d.is3d && !isObs(flags, IX(d, i, j, k - 1))) A0[ix] += 1.;
				if (d.is3d && !isObs(flags, IX(d, i, j, k + 1))) A0[ix] += 1.;
				if (isFluid(flags, IX(d, i + 1, j, k))) Ai[ix] = -1.;
				if (isFluid(flags, IX(d, i, j + 1, k))) Aj[ix] = -1.;
				if (d.is3d && isFluid(flags, IX(d, i, j, k + 1))) Ak[ix] = -1.;
			}
}

/* ApplyMatrix / ApplyMatrix2D (conjugategrad.h:118-151), all idx.  A fluid
 * cell on the outermost layer would read outside the grid in the reference;
 * such out-of-range neighbour terms are taken as 0 here. */
void og_apply_matrix(og_dims d, const int* flags, og_real* dst, const og_real* src,
                     const og_real* A0, const og_real* Ai, const og_real* Aj, const og_real* Ak) {
	long n = NCELLS(d);
	const long X = 1, Y = d.nx, Z = SZ(d);
	OMP_FOR
	for (long ix = 0; ix < n; ix++) {
		if (!isFluid(flags, ix)) { dst[ix] = src[ix]; continue; }
#define SRC(o) ((ix + (o)) >= 0 && (ix + (o)) < n ? src[ix + (o)] : (og_real)0)
#define COEF(A, o) ((ix + (o)) >= 0 && (ix + (o)) < n ? (A)[ix + (o)] : (og_real)0)
		if (d.is3d)
			dst[ix] = src[ix] * A0[ix] + SRC(-X) * COEF(Ai, -X) + SRC(X) * Ai[ix] +
			          SRC(-Y) * COEF(Aj, -Y) + SRC(Y) * Aj[ix] + SRC(-Z) * COEF(Ak, -Z) +
			          SRC(Z) * Ak[ix];
		else
			dst[ix] = src[ix] * A0[ix] + SRC(-X) * COEF(Ai, -X) + SRC(X) * Ai[ix] +
			          SRC(-Y) * COEF(Aj, -Y) + SRC(Y) * Aj[ix];
#undef SRC
#undef COEF
	}
}

/* zero-pressure cell selection (pressure.cpp:349-381); -1 if empty cells exist */
int og_find_fix_cell(og_dims d, const int* flags) {
	long n = NCELLS(d);
	long numEmpty = 0;
	for (long ix = 0; ix < n; ix++)
		if (isEmpty(flags, ix)) numEmpty++;
	if (numEmpty != 0) return -1;
	int tcx = d.nx / 2, tcy = d.ny - 1, tcz = d.is3d ? d.nz / 2 : 0;
	for (int o = 0; o < 3; o++) {
		int y = tcy - o;
		if (y < 0) continue;
		long ix = IX(d, tcx, y, tcz);
		if (isFluid(flags, ix)) return (int)ix;
	}
	for (int k = KMIN(d, 1); k < KMAX(d, 1); k++)
		for (int j = 1; j < d.ny - 1; j++)
			for (int i = 1; i < d.nx - 1; i++)
				if (isFluid(flags, IX(d, i, j, k))) return (int)IX(d, i, j, k);
	return -1;
}

/* fixPressure (pressure.cpp:226-245) */
void og_fix_pressure(og_dims d, int fixPidx, og_real value, og_real* rhs, og_real* A0, og_real* Ai,
                     og_real* Aj, og_real* Ak) {
	const long X = 1, Y = d.nx, Z = SZ(d);
	long f = fixPidx;
	rhs[f + X] -= Ai[f] * value;
	rhs[f + Y] -= Aj[f] * value;
	rhs[f - X] -= Ai[f - X] * value;
	rhs[f - Y] -= Aj[f - Y] * value;
	if (d.is3d) {
		rhs[f + Z] -= Ak[f] * value;
		rhs[f - Z] -= Ak[f - Z] * value;
	}
	rhs[f] = value;
	A0[f] = (og_real)1;
	Ai[f] = Aj[f] = Ak[f] = (og_real)0;
	Ai[f - X] = (og_real)0;
	Aj[f - Y] = (og_real)0;
	if (d.is3d) Ak[f - Z] = (og_real)0;
}

/* knCorrectVelocity (pressure.cpp:87-109), bnd=1 */
void og_correct_velocity(og_dims d, const int* flags, og_real* vel, const og_real* p) {
	const int b = 1;
	OMP_FOR
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				long ix = IX(d, i, j, k);
				long xm = IX(d, i - 1, j, k), ym = IX(d, i, j - 1, k), zm = IX(d, i, j, k - 1);
				og_real* v = vel + 3 * ix;
				if (isFluid(flags, ix)) {
					if (isFluid(flags, xm)) v[0] -= (p[ix] - p[xm]);
					if (isFluid(flags, ym)) v[1] -= (p[ix] - p[ym]);
					if (d.is3d && isFluid(flags, zm)) v[2] -= (p[ix] - p[zm]);
					if (isEmpty(flags, xm)) v[0] -= p[ix];
					if (isEmpty(flags, ym)) v[1] -= p[ix];
					if (d.is3d && isEmpty(flags, zm)) v[2] -= p[ix];
				} else if (isEmpty(flags, ix) && !isOutflow(flags, ix)) {
					if (isFluid(flags, xm)) v[0] += p[xm];
					else v[0] = 0.f;
					if (isFluid(flags, ym)) v[1] += p[ym];
					else v[1] = 0.f;
					if (d.is3d) {
						if (isFluid(flags, zm)) v[2] += p[zm];
						else v[2] = 0.f;
					}
				}
			}
}

/* ---- CG building blocks (conjugategrad.cpp:175-196, grid.h:478, grid.cpp:162-323) ---- */
typedef struct { const og_real *a, *b; } dot_ctx;
static double dot_slab(og_dims d, long lo, long hi, const void* c) {
	(void)d;
	const dot_ctx* x = (const dot_ctx*)c;
	double r = 0.0;
	for (long i = lo; i < hi; i++) r += (x->a[i] * x->b[i]); /* float product, double sum */
	return r;
}
static double grid_dot(og_dims d, const og_real* a, const og_real* b) {
	dot_ctx c = {a, b};
	return reduce_sum(d, dot_slab, &c);
}
static double sumsq_slab(og_dims d, long lo, long hi, const void* c) {
	(void)d;
	const og_real* a = (const og_real*)c;
	double r = 0.0;
	for (long i = lo; i < hi; i++) r += (double)a[i] * (double)a[i];
	return r;
}
static og_real grid_max_abs(og_dims d, const og_real* a) {
	long n = NCELLS(d);
	og_real mn = OG_MAX, mx = -OG_MAX;
	for (long i = 0; i < n; i++) {
		if (a[i] < mn) mn = a[i];
		if (a[i] > mx) mx = a[i];
	}
	og_real amn = OG_FABS(mn), amx = OG_FABS(mx);
	return amn > amx ? amn : amx; /* max(fabs(amin), fabs(amax)) */
}

/* InitPreconditionModifiedIncompCholesky2 (conjugategrad.cpp:66-97);
 * sqrt(float) resolves to the float overload there (using namespace std). */
static void mic_init(og_dims d, const int* flags, og_real* P, const og_real* A0,
                     const og_real* Ai, const og_real* Aj, const og_real* Ak) {
	long n = NCELLS(d);
	memset(P, 0, sizeof(og_real) * (size_t)n);
	const long X = 1, Y = d.nx, Z = SZ(d);
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long c = IX(d, i, j, k);
				if (!isFluid(flags, c)) continue;
				const og_real tau = 0.97;
				const og_real sigma = 0.25;
				og_real e = 0.;
				e = A0[c] - (Ai[c - X] * P[c - X]) * (Ai[c - X] * P[c - X]) -
				    (Aj[c - Y] * P[c - Y]) * (Aj[c - Y] * P[c - Y]) -
				    (Ak[c - Z] * P[c - Z]) * (Ak[c - Z] * P[c - Z]);
				e -= tau * (Ai[c - X] * (Aj[c - X] + Ak[c - X]) * (P[c - X] * P[c - X]) +
				            Aj[c - Y] * (Ai[c - Y] + Ak[c - Y]) * (P[c - Y] * P[c - Y]) +
				            Ak[c - Z] * (Ai[c - Z] + Aj[c - Z]) * (P[c - Z] * P[c - Z]) + 0.);
				if (e < sigma * A0[c]) e = A0[c];
				P[c] = (og_real)(1. / (double)OG_SQRT(e));
			}
}

/* ApplyPreconditionModifiedIncompCholesky2 (conjugategrad.cpp:135-159) */
static void mic_apply(og_dims d, const int* flags, og_real* dst, const og_real* r, const og_real* P,
                      const og_real* Ai, const og_real* Aj, const og_real* Ak) {
	const long X = 1, Y = d.nx, Z = SZ(d);
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long c = IX(d, i, j, k);
				if (!isFluid(flags, c)) continue;
				const og_real p = P[c];
				dst[c] = p * (r[c] - dst[c - X] * Ai[c - X] * P[c - X] -
				              dst[c - Y] * Aj[c - Y] * P[c - Y] - dst[c - Z] * Ak[c - Z] * P[c - Z]);
			}
	for (int k = d.nz - 1; k >= 0; k--)
		for (int j = d.ny - 1; j >= 0; j--)
			for (int i = d.nx - 1; i >= 0; i--) {
				long c = IX(d, i, j, k);
				if (!isFluid(flags, c)) continue;
				const og_real p = P[c];
				dst[c] = p * (dst[c] - dst[c + X] * Ai[c] * p - dst[c + Y] * Aj[c] * p -
				              dst[c + Z] * Ak[c] * p);
			}
}

/* z = M^-1 r for the mIC preconditioner built from A (init + apply) */
void og_mic_precondition(og_dims d, const int* flags, const og_real* A0, const og_real* Ai,
                         const og_real* Aj, const og_real* Ak, const og_real* r, og_real* z) {
	long n = NCELLS(d);
	og_real* P = (og_real*)calloc((size_t)n, sizeof(og_real));
	mic_init(d, flags, P, A0, Ai, Aj, Ak);
	mic_apply(d, flags, z, r, P, Ai, Aj, Ak);
	free(P);
}

/* ||b - A x||_2 / ||b||_2 in FP64 (extension used for the relative criterion
 * and for reporting); A applied as in ApplyMatrix with double accumulation */
static double rel_residual(og_dims d, const int* flags, const og_real* b, const og_real* x,
                           const og_real* A0, const og_real* Ai, const og_real* Aj,
                           const og_real* Ak) {
	long n = NCELLS(d);
	const long X = 1, Y = d.nx, Z = SZ(d);
	double rr = 0, bb = 0;
	for (long ix = 0; ix < n; ix++) {
		double ax;
		if (!isFluid(flags, ix)) ax = x[ix];
		else {
			ax = (double)x[ix] * A0[ix];
			if (ix - X >= 0) ax += (double)x[ix - X] * Ai[ix - X];
			if (ix + X < n) ax += (double)x[ix + X] * Ai[ix];
			if (ix - Y >= 0) ax += (double)x[ix - Y] * Aj[ix - Y];
			if (ix + Y < n) ax += (double)x[ix + Y] * Aj[ix];
			if (d.is3d) {
				if (ix - Z >= 0) ax += (double)x[ix - Z] * Ak[ix - Z];
				if (ix + Z < n) ax += (double)x[ix + Z] * Ak[ix];
			}
		}
		double r = (double)b[ix] - ax;
		rr += r * r;
		bb += (double)b[ix] * (double)b[ix];
	}
	return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* solvePressureSystem (pressure.cpp:312-452) with GridCg (conjugategrad.cpp:198-299) */
int og_solve_pressure_system(og_dims d, og_real* rhs, og_real* pressure, const int* flags,
                             const og_pressure_opts* opts, og_solve_stats* st) {
	return og_solve_pressure_system_bnd(d, rhs, pressure, flags, opts, NULL, st);
}

/* solvePressureSystem (pressure.cpp:312-452): matrix (with fractions and the
 * ghost-fluid diagonal when given), optional pinned cell, preconditioned CG */
int og_solve_pressure_system_bnd(og_dims d, og_real* rhs, og_real* pressure, const int* flags,
                                 const og_pressure_opts* opts, const og_pressure_bnd* bnd, og_solve_stats* st) {
	og_pressure_opts o;
	if (opts) o = *opts;
	else og_pressure_opts_default(&o);
	long n = NCELLS(d);
	size_t nb = sizeof(og_real) * (size_t)n;
	og_real *res = (og_real*)calloc(1, nb), *search = (og_real*)calloc(1, nb),
	        *A0 = (og_real*)calloc(1, nb), *Ai = (og_real*)calloc(1, nb),
	        *Aj = (og_real*)calloc(1, nb), *Ak = (og_real*)calloc(1, nb),
	        *tmp = (og_real*)calloc(1, nb), *P = NULL;
	og_make_laplace_bnd(d, flags, A0, Ai, Aj, Ak, bnd);
	int fix = -1;
	if (o.zeroPressureFixing || o.cgAccuracy < 1e-07) {
		fix = og_find_fix_cell(d, flags);
		if (fix >= 0) og_fix_pressure(d, fix, (og_real)0, rhs, A0, Ai, Aj, Ak);
	}
	int pc = o.preconditioner;
	int maxIter;
	int mx = d.nx > d.ny ? d.nx : d.ny;
	if (d.nz > mx) mx = d.nz;
	og_mg* mg = NULL;
	if (pc == OG_PC_NONE || pc == OG_PC_MIC) {
		maxIter = (int)(o.cgMaxIterFac * mx) * (d.is3d ? 1 : 4);
		if (!d.is3d) pc = OG_PC_NONE; /* setICPreconditioner: mIC is 3D only */
	} else {
		maxIter = 100;
		mg = og_mg_create(d);
	}
	if (o.maxIterOverride > 0) maxIter = o.maxIterOverride;

	/* doInit */
	memset(pressure, 0, nb);
	memcpy(res, rhs, nb);
	if (pc == OG_PC_MIC) {
		P = (og_real*)calloc(1, nb);
		mic_init(d, flags, P, A0, Ai, Aj, Ak);
		mic_apply(d, flags, tmp, res, P, Ai, Aj, Ak);
	} else if (mg) {
		og_mg_set_a(mg, A0, Ai, Aj, Ak);
		og_mg_set_coarsest_accuracy(mg, o.cgAccuracy * 1E-4);
		og_mg_vcycle(mg, res, tmp);
	} else {
		memcpy(tmp, res, nb);
	}
	memcpy(search, tmp, nb);
	og_real sigma = (og_real)grid_dot(d, tmp, res);
	og_real resNorm = 1e20;
	int iters = 0, diverged = 0;
	double bnorm2 = 0;
	if (o.relTol > 0) bnorm2 = reduce_sum(d, sumsq_slab, rhs);
	for (int iter = 0; iter < maxIter; iter++) {
		/* iterate() */
		iters++;
		og_apply_matrix(d, flags, tmp, search, A0, Ai, Aj, Ak);
		og_real dp = (og_real)grid_dot(d, tmp, search);
		og_real alpha = 0.;
		if (OG_FABS(dp) > 0.) alpha = sigma / (og_real)dp;
		OMP_FOR
		for (long i = 0; i < n; i++) pressure[i] += alpha * search[i];
		OMP_FOR
		for (long i = 0; i < n; i++) res[i] += (-alpha) * tmp[i];
		if (pc == OG_PC_MIC) mic_apply(d, flags, tmp, res, P, Ai, Aj, Ak);
		else if (mg) og_mg_vcycle(mg, res, tmp);
		else memcpy(tmp, res, nb);
		int done;
		if (o.relTol > 0) {
			double rr = reduce_sum(d, sumsq_slab, res);
			resNorm = (og_real)sqrt(rr / (bnorm2 > 0 ? bnorm2 : 1));
			done = sqrt(rr) <= o.relTol * sqrt(bnorm2);
		} else {
			if (o.useL2Norm) resNorm = (og_real)reduce_sum(d, sumsq_slab, res);
			else resNorm = grid_max_abs(d, res);
			done = resNorm < o.cgAccuracy;
		}
		if (done) {
			sigma = resNorm;
			break;
		}
		og_real sigmaNew = (og_real)grid_dot(d, tmp, res);
		og_real beta = sigmaNew / sigma;
		OMP_FOR
		for (long i = 0; i < n; i++) search[i] = tmp[i] + beta * search[i];
		sigma = sigmaNew;
		if (!(resNorm < 1e35)) {
			diverged = 1;
			break;
		}
	}
	if (st) {
		st->iterations = iters;
		st->resNorm = resNorm;
		st->relResidual = rel_residual(d, flags, rhs, pressure, A0, Ai, Aj, Ak);
		st->fixPidx = fix;
		st->diverged = diverged;
		st->maxIter = maxIter;
	}
	if (mg) og_mg_destroy(mg);
	free(res); free(search); free(A0); free(Ai); free(Aj); free(Ak); free(tmp); free(P);
	return diverged;
}

/* solvePressure (pressure.cpp:480-521) */
int og_solve_pressure(og_dims d, og_real* vel, og_real* pressure, const int* flags,
                      const og_pressure_opts* opts, og_solve_stats* st, og_real* rhsOut) {
	return og_solve_pressure_bnd(d, vel, pressure, flags, opts, NULL, st, rhsOut);
}

int og_solve_pressure_bnd(og_dims d, og_real* vel, og_real* pressure, const int* flags,
                          const og_pressure_opts* opts, const og_pressure_bnd* bnd, og_solve_stats* st,
                          og_real* rhsOut) {
	og_pressure_opts o;
	if (opts) o = *opts;
	else og_pressure_opts_default(&o);
	long n = NCELLS(d);
	og_real* rhs = (og_real*)calloc((size_t)n, sizeof(og_real));
	double sum = 0;
	int cnt = og_make_rhs_bnd(d, flags, rhs, vel, bnd, &sum);
	if (o.enforceCompatibility) {
		og_real corr = (og_real)(-sum / (og_real)cnt);
		for (long i = 0; i < n; i++) rhs[i] += corr;
	}
	int rc = og_solve_pressure_system_bnd(d, rhs, pressure, flags, &o, bnd, st);
	if (rc == 0) og_correct_velocity_bnd(d, flags, vel, pressure, bnd);
	if (rhsOut) memcpy(rhsOut, rhs, sizeof(og_real) * (size_t)n);
	free(rhs);
	return rc;
}

/* ===========================================================================
 * Second-order obstacle and free-surface boundaries
 * ========================================================================= */
void og_pressure_bnd_default(og_pressure_bnd* b) {
	memset(b, 0, sizeof(*b));
	b->gfClamp = (og_real)1e-04;
	b->surfTens = (og_real)0;
}

/* thetaHelper (pressure.cpp:114-120): fraction of the way from the cell to
 * the interface; denom is Real, the guard compares against a double */
static og_real gf_theta(og_real inside, og_real outside) {
	const og_real denom = inside - outside;
	if ((double)denom > -1e-04) return (og_real)0.5;
	og_real q = inside / denom;
	if (q > (og_real)1) q = (og_real)1;  /* std::min(Real(1), q) */
	if ((og_real)0 > q) q = (og_real)0;  /* std::max(Real(0), .) */
	return q;
}

/* ghostFluidHelper (pressure.cpp:122-127): returns gfClamp itself when the
 * fraction falls below it (the reference's "return alpha = gfClamp"), else
 * 1 - 1/alpha evaluated in double and stored as Real */
og_real og_ghost_fluid_factor(og_real phi_in, og_real phi_out, og_real gfClamp) {
	og_real alpha = gf_theta(phi_in, phi_out);
	if (alpha < gfClamp) return gfClamp;
	if (OG_MUT(22)) return (og_real)(1. - (double)alpha);
	return (og_real)(1. - (1. / (double)alpha));
}
static og_real gf_helper(const og_real* phi, long idx, long off, og_real gfClamp) {
	return og_ghost_fluid_factor(phi[idx], phi[idx + off], gfClamp);
}
/* surfTensHelper (pressure.cpp:129-131) */
static og_real st_helper(const og_real* phi, const og_real* curv, long idx, long off, og_real surfTens,
                         og_real gfClamp) {
	return surfTens * (curv[idx + off] - gf_helper(phi, idx, off, gfClamp) * curv[idx]);
}
/* ghostFluidWasClamped (pressure.cpp:194-199) */
static int gf_clamped(const og_real* phi, long idx, long off, og_real gfClamp) {
	return gf_theta(phi[idx], phi[idx + off]) < gfClamp;
}

/* MakeRhs (pressure.cpp:32-84), bnd=1, with fractions (+obvel), surface
 * tension (phi and curv) and perCellCorr */
int og_make_rhs_bnd(og_dims d, const int* flags, og_real* rhs, const og_real* vel, const og_pressure_bnd* b,
                    double* sumOut) {
	if (!b) return og_make_rhs(d, flags, rhs, vel, sumOut);
	const int bw = 1;
	long n = NCELLS(d);
	memset(rhs, 0, sizeof(og_real) * (size_t)n);
	const long X = 1, Y = d.nx, Z = SZ(d);
	const og_real* fr = b->fractions;
	const og_real* ob = b->obvel;
	int cnt = 0;
	double sum = 0;
	for (int k = KMIN(d, bw); k < KMAX(d, bw); k++)
		for (int j = bw; j < d.ny - bw; j++)
			for (int i = bw; i < d.nx - bw; i++) {
				long ix = IX(d, i, j, k);
				if (!isFluid(flags, ix)) { rhs[ix] = 0; continue; }
				og_real set = 0;
				if (!fr) {
					set = vel[3 * ix] - vel[3 * (ix + X)] + vel[3 * ix + 1] - vel[3 * (ix + Y) + 1];
					if (d.is3d) set += vel[3 * ix + 2] - vel[3 * (ix + Z) + 2];
				} else {
					set = fr[3 * ix] * vel[3 * ix] - fr[3 * (ix + X)] * vel[3 * (ix + X)] +
					      fr[3 * ix + 1] * vel[3 * ix + 1] -
					      fr[3 * (OG_MUT(21) ? ix : ix + Y) + 1] * vel[3 * (ix + Y) + 1];
					if (d.is3d) set += fr[3 * ix + 2] * vel[3 * ix + 2] - fr[3 * (ix + Z) + 2] * vel[3 * (ix + Z) + 2];
					if (ob) {
						set += (1 - fr[3 * ix]) * ob[3 * ix] - (1 - fr[3 * (ix + X)]) * ob[3 * (ix + X)] +
						       (1 - fr[3 * ix + 1]) * ob[3 * ix + 1] - (1 - fr[3 * (ix + Y) + 1]) * ob[3 * (ix + Y) + 1];
						if (d.is3d)
							set += (1 - fr[3 * ix + 2]) * ob[3 * ix + 2] - (1 - fr[3 * (ix + Z) + 2]) * ob[3 * (ix + Z) + 2];
					}
				}
				if (b->phi && b->curv) {
					const og_real *ph = b->phi, *cv = b->curv, st = b->surfTens, gc = b->gfClamp;
					if (isEmpty(flags, ix - X)) set += st_helper(ph, cv, ix, -X, st, gc);
					if (isEmpty(flags, ix + X)) set += st_helper(ph, cv, ix, +X, st, gc);
					if (isEmpty(flags, ix - Y)) set += st_helper(ph, cv, ix, -Y, st, gc);
					if (isEmpty(flags, ix + Y)) set += st_helper(ph, cv, ix, +Y, st, gc);
					if (d.is3d) {
						if (isEmpty(flags, ix - Z)) set += st_helper(ph, cv, ix, -Z, st, gc);
						if (isEmpty(flags, ix + Z)) set += st_helper(ph, cv, ix, +Z, st, gc);
					}
				}
				if (b->perCellCorr) set += b->perCellCorr[ix];
				sum += set;
				cnt++;
				rhs[ix] = set;
			}
	if (sumOut) *sumOut = sum;
	return cnt;
}

/* MakeLaplaceMatrix (conjugategrad.h:154-187), fractions branch, then
 * ApplyGhostFluidDiagonal (pressure.cpp:133-149) when phi is given */
void og_make_laplace_bnd(og_dims d, const int* flags, og_real* A0, og_real* Ai, og_real* Aj, og_real* Ak,
                         const og_pressure_bnd* b) {
	const og_real* fr = b ? b->fractions : NULL;
	if (!fr) {
		og_make_laplace(d, flags, A0, Ai, Aj, Ak);
	} else {
		const int bw = 1;
		long n = NCELLS(d);
		memset(A0, 0, sizeof(og_real) * (size_t)n);
		memset(Ai, 0, sizeof(og_real) * (size_t)n);
		memset(Aj, 0, sizeof(og_real) * (size_t)n);
		memset(Ak, 0, sizeof(og_real) * (size_t)n);
		const long X = 1, Y = d.nx, Z = SZ(d);
		for (int k = KMIN(d, bw); k < KMAX(d, bw); k++)
			for (int j = bw; j < d.ny - bw; j++)
				for (int i = bw; i < d.nx - bw; i++) {
					long ix = IX(d, i, j, k);
					if (!isFluid(flags, ix)) continue;
					A0[ix] += fr[3 * ix];
					A0[ix] += fr[3 * (ix + X)];
					A0[ix] += fr[3 * ix + 1];
					A0[ix] += fr[3 * (ix + Y) + 1];
					if (d.is3d) A0[ix] += fr[3 * ix + 2];
					if (d.is3d) A0[ix] += fr[3 * (ix + Z) + 2];
					if (isFluid(flags, ix + X)) Ai[ix] = -fr[3 * (ix + X)];
					if (isFluid(flags, ix + Y)) Aj[ix] = -fr[3 * (ix + Y) + 1];
					if (d.is3d && isFluid(flags, ix + Z)) Ak[ix] = -fr[3 * (ix + Z) + 2];
				}
	}
	if (b && b->phi) {
		const int bw = 1;
		const long X = 1, Y = d.nx, Z = SZ(d);
		const og_real* ph = b->phi;
		const og_real gc = b->gfClamp;
		for (int k = KMIN(d, bw); k < KMAX(d, bw); k++)
			for (int j = bw; j < d.ny - bw; j++)
				for (int i = bw; i < d.nx - bw; i++) {
					long ix = IX(d, i, j, k);
					if (!isFluid(flags, ix)) continue;
					if (isEmpty(flags, ix - X)) A0[ix] -= gf_helper(ph, ix, -X, gc);
					if (isEmpty(flags, ix + X)) A0[ix] -= gf_helper(ph, ix, +X, gc);
					if (isEmpty(flags, ix - Y)) A0[ix] -= gf_helper(ph, ix, -Y, gc);
					if (isEmpty(flags, ix + Y)) A0[ix] -= gf_helper(ph, ix, +Y, gc);
					if (d.is3d) {
						if (isEmpty(flags, ix - Z)) A0[ix] -= gf_helper(ph, ix, -Z, gc);
						if (isEmpty(flags, ix + Z)) A0[ix] -= gf_helper(ph, ix, +Z, gc);
					}
				}
	}
}

/* correctVelocity (pressure.cpp:455-476): knCorrectVelocity, then with phi
 * knCorrectVelocityGhostFluid (:151-190) and knReplaceClampedGhostFluidVels
 * (:201-219), all bnd=1 */
void og_correct_velocity_bnd(og_dims d, const int* flags, og_real* vel, const og_real* p, const og_pressure_bnd* b) {
	og_correct_velocity(d, flags, vel, p);
	if (!b || !b->phi) return;
	const int bw = 1;
	const long X = 1, Y = d.nx, Z = SZ(d);
	const og_real* ph = b->phi;
	const og_real gc = b->gfClamp;
	for (int k = KMIN(d, bw); k < KMAX(d, bw); k++)
		for (int j = bw; j < d.ny - bw; j++)
			for (int i = bw; i < d.nx - bw; i++) {
				long ix = IX(d, i, j, k);
				og_real* v = vel + 3 * ix;
				if (isFluid(flags, ix)) {
					if (isEmpty(flags, ix - X)) v[0] += p[ix] * gf_helper(ph, ix, -X, gc);
					if (isEmpty(flags, ix - Y)) v[1] += p[ix] * gf_helper(ph, ix, -Y, gc);
					if (d.is3d && isEmpty(flags, ix - Z)) v[2] += p[ix] * gf_helper(ph, ix, -Z, gc);
				} else if (isEmpty(flags, ix) && !isOutflow(flags, ix)) {
					if (isFluid(flags, ix - X)) v[0] -= p[OG_MUT(24) ? ix : ix - X] * gf_helper(ph, ix - X, +X, gc);
					else v[0] = 0.f;
					if (isFluid(flags, ix - Y)) v[1] -= p[ix - Y] * gf_helper(ph, ix - Y, +Y, gc);
					else v[1] = 0.f;
					if (d.is3d) {
						if (isFluid(flags, ix - Z)) v[2] -= p[ix - Z] * gf_helper(ph, ix - Z, +Z, gc);
						else v[2] = 0.f;
					}
				}
				if (b->curv) {
					const og_real* cv = b->curv;
					const og_real st = b->surfTens;
					if (isFluid(flags, ix)) {
						if (isEmpty(flags, ix - X)) v[0] += st_helper(ph, cv, ix, -X, st, gc);
						if (isEmpty(flags, ix - Y)) v[1] += st_helper(ph, cv, ix, -Y, st, gc);
						if (d.is3d && isEmpty(flags, ix - Z)) v[2] += st_helper(ph, cv, ix, -Z, st, gc);
					} else if (isEmpty(flags, ix) && !isOutflow(flags, ix)) {
						v[0] -= isFluid(flags, ix - X) ? st_helper(ph, cv, ix - X, +X, st, gc) : (og_real)0;
						v[1] -= isFluid(flags, ix - Y) ? st_helper(ph, cv, ix - Y, +Y, st, gc) : (og_real)0;
						if (d.is3d) v[2] -= isFluid(flags, ix - Z) ? st_helper(ph, cv, ix - Z, +Z, st, gc) : (og_real)0;
					}
				}
			}
	/* knReplaceClampedGhostFluidVels: empty cells take the fluid neighbour's
	 * face value where that face's fraction was clamped (reads fluid cells'
	 * faces only, which this pass leaves alone) */
	for (int k = KMIN(d, bw); k < KMAX(d, bw); k++)
		for (int j = bw; j < d.ny - bw; j++)
			for (int i = bw; i < d.nx - bw; i++) {
				long ix = IX(d, i, j, k);
				if (!isEmpty(flags, ix)) continue;
				og_real* v = vel + 3 * ix;
				if (isFluid(flags, ix - X) && gf_clamped(ph, ix - X, +X, gc)) v[0] = vel[3 * (ix - X)];
				if (isFluid(flags, ix - Y) && gf_clamped(ph, ix - Y, +Y, gc)) v[1] = vel[3 * (ix - Y) + 1];
				if (d.is3d && isFluid(flags, ix - Z) && gf_clamped(ph, ix - Z, +Z, gc)) v[2] = vel[3 * (ix - Z) + 2];
				if (isFluid(flags, ix + X) && gf_clamped(ph, ix + X, -X, gc)) v[0] = vel[3 * (ix + X)];
				if (isFluid(flags, ix + Y) && gf_clamped(ph, ix + Y, -Y, gc)) v[1] = vel[3 * (ix + Y) + 1];
				if (d.is3d && isFluid(flags, ix + Z) && gf_clamped(ph, ix + Z, -Z, gc)) v[2] = vel[3 * (ix + Z) + 2];
			}
}

/* normalize(Vec3&) (util/vectorbase.h:416-429): unchanged when |v|^2 is
 * within eps^2 of 1, scaled by (Real)(1/norm) when above eps^2, else 0 */
static void vec_normalize(og_real* v) {
	const og_real l = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
	const og_real eps2 = OG_VEC_EPS * OG_VEC_EPS;
	if (fabs((double)l - 1.) < (double)eps2) return;
	if (l > eps2) {
		const og_real norm = (og_real)sqrt((double)l);
		const og_real s = (og_real)(1. / (double)norm);
		v[0] *= s;
		v[1] *= s;
		v[2] *= s;
	} else {
		v[0] = v[1] = v[2] = 0;
	}
}

/* KnSetWallBcsFrac (extforces.cpp:220-303) into a copy of vel, then
 * vel.swap(copy) (:307-315): at interior fluid or obstacle cells, a face
 * that touches an obstacle loses the component of the face velocity
 * (getAtMAC) along the normalised obstacle-SDF gradient */
void og_set_wall_bcs_frac(og_dims d, const int* flags, og_real* vel, const og_real* phiObs) {
	long n = NCELLS(d);
	og_real* tgt = (og_real*)malloc(sizeof(og_real) * 3 * (size_t)n);
	memcpy(tgt, vel, sizeof(og_real) * 3 * (size_t)n);
	const long X = 1, Y = d.nx, Z = SZ(d);
	const og_real* P = phiObs;
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long ix = IX(d, i, j, k);
				const int curFluid = isFluid(flags, ix) != 0, curObs = isObs(flags, ix) != 0;
				if (!curFluid && !curObs) continue;
				const int inb = i >= 1 && j >= 1 && i < d.nx - 1 && j < d.ny - 1 && (!d.is3d || (k >= 1 && k < d.nz - 1));
				if (!inb) continue;
				if (curObs | (isObs(flags, ix - X) != 0)) {
					og_real dphi[3] = {0, 0, 0};
					const long my = OG_MUT(23) ? -Y : Y;
					const og_real tmp1 = (og_real)((P[ix] + P[ix - X]) * .5);
					og_real tmp2 = (og_real)((P[ix + my] + P[ix - X + my]) * .5);
					og_real phi1 = (og_real)((tmp1 + tmp2) * .5);
					tmp2 = (og_real)((P[ix - my] + P[ix - X - my]) * .5);
					og_real phi2 = (og_real)((tmp1 + tmp2) * .5);
					dphi[0] = P[ix] - P[ix - X];
					dphi[1] = phi1 - phi2;
					if (d.is3d) {
						tmp2 = (og_real)((P[ix + Z] + P[ix - X + Z]) * .5);
						phi1 = (og_real)((tmp1 + tmp2) * .5);
						tmp2 = (og_real)((P[ix - Z] + P[ix - X - Z]) * .5);
						phi2 = (og_real)((tmp1 + tmp2) * .5);
						dphi[2] = phi1 - phi2;
					}
					vec_normalize(dphi);
					og_real vm[3];
					getAtMACX(d, vel, i, j, k, vm);
					const og_real dt = dphi[0] * vm[0] + dphi[1] * vm[1] + dphi[2] * vm[2];
					tgt[3 * ix] = vm[0] - dt * dphi[0];
				}
				if (curObs | (isObs(flags, ix - Y) != 0)) {
					og_real dphi[3] = {0, 0, 0};
					const og_real tmp1 = (og_real)((P[ix] + P[ix - Y]) * .5);
					og_real tmp2 = (og_real)((P[ix + X] + P[ix + X - Y]) * .5);
					og_real phi1 = (og_real)((tmp1 + tmp2) * .5);
					tmp2 = (og_real)((P[ix - X] + P[ix - X - Y]) * .5);
					og_real phi2 = (og_real)((tmp1 + tmp2) * .5);
					dphi[0] = phi1 - phi2;
					dphi[1] = P[ix] - P[ix - Y];
					if (d.is3d) {
						tmp2 = (og_real)((P[ix + Z] + P[ix - Y + Z]) * .5);
						phi1 = (og_real)((tmp1 + tmp2) * .5);
						tmp2 = (og_real)((P[ix - Z] + P[ix - Y - Z]) * .5);
						phi2 = (og_real)((tmp1 + tmp2) * .5);
						dphi[2] = phi1 - phi2;
					}
					vec_normalize(dphi);
					og_real vm[3];
					getAtMACY(d, vel, i, j, k, vm);
					const og_real dt = dphi[0] * vm[0] + dphi[1] * vm[1] + dphi[2] * vm[2];
					tgt[3 * ix + 1] = vm[1] - dt * dphi[1];
				}
				if (d.is3d && (curObs | (isObs(flags, ix - Z) != 0))) {
					og_real dphi[3] = {0, 0, 0};
					const og_real tmp1 = (og_real)((P[ix] + P[ix - Z]) * .5);
					og_real tmp2 = (og_real)((P[ix + X] + P[ix + X - Z]) * .5);
					og_real phi1 = (og_real)((tmp1 + tmp2) * .5);
					tmp2 = (og_real)((P[ix - X] + P[ix - X - Z]) * .5);
					og_real phi2 = (og_real)((tmp1 + tmp2) * .5);
					dphi[0] = phi1 - phi2;
					tmp2 = (og_real)((P[ix + Y] + P[ix + Y - Z]) * .5);
					phi1 = (og_real)((tmp1 + tmp2) * .5);
					tmp2 = (og_real)((P[ix - Y] + P[ix - Y - Z]) * .5);
					phi2 = (og_real)((tmp1 + tmp2) * .5);
					dphi[1] = phi1 - phi2;
					dphi[2] = P[ix] - P[ix - Z];
					vec_normalize(dphi);
					og_real vm[3];
					getAtMACZ(d, vel, i, j, k, vm);
					const og_real dt = dphi[0] * vm[0] + dphi[1] * vm[1] + dphi[2] * vm[2];
					tgt[3 * ix + 2] = vm[2] - dt * dphi[2];
				}
			}
	memcpy(vel, tgt, sizeof(og_real) * 3 * (size_t)n);
	free(tgt);
}

/* ===========================================================================
 * Other GridCg clients: cgSolveDiffusion (conjugategrad.cpp:347-423) and
 * cgSolveWE (plugin/waves.cpp:69-141)
 * ========================================================================= */

/* GridCg<ApplyMatrix(2D)> without preconditioner: doInit clears x, r = rhs,
 * then iterate() until ||r||_inf < acc or maxIter (conjugategrad.cpp:210-310) */
static void plain_gridcg(og_dims d, const int* flags, og_real* x, const og_real* rhs, const og_real* A0,
                         const og_real* Ai, const og_real* Aj, const og_real* Ak, og_real acc, int maxIter,
                         og_solve_stats* st) {
	long n = NCELLS(d);
	size_t nb = sizeof(og_real) * (size_t)n;
	og_real *res = (og_real*)calloc(1, nb), *search = (og_real*)calloc(1, nb), *tmp = (og_real*)calloc(1, nb);
	memset(x, 0, nb);
	memcpy(res, rhs, nb);
	memcpy(tmp, res, nb);
	memcpy(search, tmp, nb);
	og_real sigma = (og_real)grid_dot(d, tmp, res);
	og_real resNorm = 1e20;
	int iters = 0, diverged = 0;
	for (int iter = 0; iter < maxIter; iter++) {
		iters++;
		og_apply_matrix(d, flags, tmp, search, A0, Ai, Aj, Ak);
		og_real dp = (og_real)grid_dot(d, tmp, search);
		og_real alpha = 0.;
		if (OG_FABS(dp) > 0.) alpha = sigma / (og_real)dp;
		for (long i = 0; i < n; i++) x[i] += alpha * search[i];
		for (long i = 0; i < n; i++) res[i] += (-alpha) * tmp[i];
		memcpy(tmp, res, nb);
		resNorm = grid_max_abs(d, res);
		if (resNorm < acc) {
			sigma = resNorm;
			break;
		}
		og_real sigmaNew = (og_real)grid_dot(d, tmp, res);
		og_real beta = sigmaNew / sigma;
		for (long i = 0; i < n; i++) search[i] = tmp[i] + beta * search[i];
		sigma = sigmaNew;
		if (!(resNorm < 1e35)) {
			diverged = 1;
			break;
		}
	}
	if (st) {
		st->iterations = iters;
		st->resNorm = resNorm;
		st->relResidual = rel_residual(d, flags, rhs, x, A0, Ai, Aj, Ak);
		st->fixPidx = -1;
		st->diverged = diverged;
		st->maxIter = maxIter;
	}
	free(res);
	free(search);
	free(tmp);
}

/* the diffusion matrix of cgSolveDiffusion (conjugategrad.cpp:355-373):
 * MakeLaplaceMatrix on an all-fluid flag grid, then obstacle rows of the
 * real flags become identity (A0 = 1) and every other row is scaled by alpha
 * with 1 added to the diagonal */
void og_diffusion_matrix(og_dims d, const int* flags, og_real alpha, og_real* A0, og_real* Ai, og_real* Aj,
                         og_real* Ak) {
	long n = NCELLS(d);
	int* dummy = (int*)calloc((size_t)n, sizeof(int));
	for (long i = 0; i < n; i++) dummy[i] = OG_FLUID;
	og_make_laplace(d, dummy, A0, Ai, Aj, Ak);
	free(dummy);
	for (long ix = 0; ix < n; ix++) {
		if (isObs(flags, ix)) {
			Ai[ix] = Aj[ix] = Ak[ix] = 0.0;
			A0[ix] = 1.0;
		} else {
			Ai[ix] *= alpha;
			Aj[ix] *= alpha;
			Ak[ix] *= alpha;
			A0[ix] *= (OG_MUT(19) ? -alpha : alpha);
			A0[ix] += 1.;
		}
	}
}

/* cgSolveDiffusion (conjugategrad.cpp:347-423).  ncomp 1: Real grid; ncomp
 * 3: Vec3 or MAC grid (AoS), whose components 0..2 (0..1 in 2D) are solved
 * one after another with the same matrix.  st (optional) gets the last
 * component's statistics.  maxIter = (int)(cgMaxIterFac * max size) * (4 in 2D). */
void og_cg_solve_diffusion(og_dims d, const int* flags, og_real* grid, int ncomp, og_real alpha,
                           og_real cgMaxIterFac, og_real cgAccuracy, og_solve_stats* st) {
	long n = NCELLS(d);
	size_t nb = sizeof(og_real) * (size_t)n;
	og_real *A0 = (og_real*)calloc(1, nb), *Ai = (og_real*)calloc(1, nb), *Aj = (og_real*)calloc(1, nb),
	        *Ak = (og_real*)calloc(1, nb), *rhs = (og_real*)calloc(1, nb), *u = (og_real*)calloc(1, nb);
	og_diffusion_matrix(d, flags, alpha, A0, Ai, Aj, Ak);
	int mx = d.nx > d.ny ? d.nx : d.ny;
	if (d.nz > mx) mx = d.nz;
	const int maxIter = (int)(cgMaxIterFac * (og_real)mx) * (d.is3d ? 1 : 4);
	if (ncomp == 1) {
		memcpy(rhs, grid, nb);
		plain_gridcg(d, flags, grid, rhs, A0, Ai, Aj, Ak, cgAccuracy, maxIter, st);
	} else {
		for (int c = 0; c < (d.is3d ? 3 : 2); c++) {
			for (long i = 0; i < n; i++) u[i] = grid[3 * i + c]; /* getComponent */
			memcpy(rhs, u, nb);
			plain_gridcg(d, flags, u, rhs, A0, Ai, Aj, Ak, cgAccuracy, maxIter, st);
			for (long i = 0; i < n; i++) grid[3 * i + c] = u[i]; /* setComponent */
		}
	}
	free(A0); free(Ai); free(Aj); free(Ak); free(rhs); free(u);
}

/* MakeRhsWE (waves.cpp:69-79), bnd=1, rhs cleared first */
void og_make_rhs_we(og_dims d, const int* flags, og_real* rhs, const og_real* ut, const og_real* utm1, og_real s,
                    int crankNic) {
	(void)flags;
	const int b = 1;
	memset(rhs, 0, sizeof(og_real) * (size_t)NCELLS(d));
	const long X = 1, Y = d.nx;
	for (int k = KMIN(d, b); k < KMAX(d, b); k++)
		for (int j = b; j < d.ny - b; j++)
			for (int i = b; i < d.nx - b; i++) {
				long ix = IX(d, i, j, k);
				rhs[ix] = (og_real)(2. * ut[ix] - utm1[ix]);
				if (crankNic)
					rhs[ix] += s * (-4. * ut[ix] + 1. * ut[ix - X] + 1. * ut[ix + (OG_MUT(20) ? -X : X)] +
					                1. * ut[ix - Y] + 1. * ut[ix + Y]);
			}
}

/* cgSolveWE (waves.cpp:87-141): s = dt*dt*cSqr*0.5; MakeLaplaceMatrix of the
 * flags scaled by s with 1 added to every diagonal; rhs = MakeRhsWE; GridCg
 * into out; then utm1 <- ut and ut <- out */
void og_cg_solve_we(og_dims d, const int* flags, og_real* ut, og_real* utm1, og_real* out, int crankNic,
                    og_real cSqr, og_real cgMaxIterFac, og_real cgAccuracy, og_real dt, og_solve_stats* st) {
	long n = NCELLS(d);
	size_t nb = sizeof(og_real) * (size_t)n;
	og_real *A0 = (og_real*)calloc(1, nb), *Ai = (og_real*)calloc(1, nb), *Aj = (og_real*)calloc(1, nb),
	        *Ak = (og_real*)calloc(1, nb), *rhs = (og_real*)calloc(1, nb);
	memset(out, 0, nb);
	og_make_laplace(d, flags, A0, Ai, Aj, Ak);
	const og_real s = (og_real)(dt * dt * cSqr * 0.5);
	for (long ix = 0; ix < n; ix++) {
		Ai[ix] *= s;
		Aj[ix] *= s;
		Ak[ix] *= s;
		A0[ix] *= s;
		A0[ix] += 1.;
	}
	og_make_rhs_we(d, flags, rhs, ut, utm1, s, crankNic);
	int mx = d.nx > d.ny ? d.nx : d.ny;
	if (d.nz > mx) mx = d.nz;
	const int maxIter = (int)(cgMaxIterFac * (og_real)mx) * (d.is3d ? 1 : 4);
	plain_gridcg(d, flags, out, rhs, A0, Ai, Aj, Ak, cgAccuracy, maxIter, st);
	memcpy(utm1, ut, nb); /* utm1.swap(ut); ut.copyFrom(out) */
	memcpy(ut, out, nb);
	free(A0); free(Ai); free(Aj); free(Ak); free(rhs);
}

/* ===========================================================================
 * Shapes (shapes.cpp:41-84,151-154,240,324-329; vectorbase.h:351-430)
 * ========================================================================= */
static int shape_inside(int shape, const og_real* a, const og_real* b, og_real r, og_real zlen,
                        og_real x, og_real y, og_real z) {
	if (shape == 1) { /* Box: inclusive bounds */
		return (x >= a[0] && y >= a[1] && z >= a[2] && x <= b[0] && y <= b[1] && z <= b[2]);
	} else if (shape == 2) { /* Sphere */
		og_real dx = (x - a[0]) / b[0], dy = (y - a[1]) / b[1], dz = (z - a[2]) / b[2];
		return dx * dx + dy * dy + dz * dz <= r * r;
	} else { /* Cylinder: b is the normalized axis, zlen its original length */
		og_real px = x - a[0], py = y - a[1], pz = z - a[2];
		og_real zz = px * b[0] + py * b[1] + pz * b[2];
		if (OG_FABS(zz) > zlen) return 0;
		og_real r2 = (px * px + py * py + pz * pz) - zz * zz;
		return r2 < r * r;
	}
}

void og_apply_shape(og_dims d, int shape, const og_real* a, const og_real* bIn, og_real r,
                    void* grid, int gridType, const og_real* value, const int* respectFlags) {
	og_real b[3] = {bIn[0], bIn[1], bIn[2]};
	og_real zlen = 0;
	if (shape == 3) { /* Cylinder ctor: mZ = normalize(mZDir) */
		og_real l = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
		og_real nrm;
		if (OG_FABS(l - 1.) < OG_VEC_EPS * OG_VEC_EPS) nrm = 1.;
		else if (l > OG_VEC_EPS * OG_VEC_EPS) {
			nrm = OG_SQRT(l);
			og_real s = (og_real)(1. / nrm);
			b[0] *= s; b[1] *= s; b[2] *= s;
		} else {
			b[0] = b[1] = b[2] = 0;
			nrm = 0.;
		}
		zlen = nrm;
	}
	for (int k = 0; k < d.nz; k++)
		for (int j = 0; j < d.ny; j++)
			for (int i = 0; i < d.nx; i++) {
				long ix = IX(d, i, j, k);
				if (respectFlags && isObs(respectFlags, ix)) continue;
				if (gridType == OG_TYPE_MAC) {
					og_real* v = (og_real*)grid + 3 * ix;
					if (shape_inside(shape, a, b, r, zlen, (og_real)i, (og_real)(j + 0.5), (og_real)(k + 0.5)))
						v[0] = value[0];
					if (shape_inside(shape, a, b, r, zlen, (og_real)(i + 0.5), (og_real)j, (og_real)(k + 0.5)))
						v[1] = value[1];
					if (shape_inside(shape, a, b, r, zlen, (og_real)(i + 0.5), (og_real)(j + 0.5), (og_real)k))
						v[2] = value[2];
					continue;
				}
				if (!shape_inside(shape, a, b, r, zlen, (og_real)(i + 0.5), (og_real)(j + 0.5),
				                  (og_real)(k + 0.5)))
					continue;
				if (gridType == OG_TYPE_INT) ((int*)grid)[ix] = (int)value[0];
				else if (gridType == OG_TYPE_REAL) ((og_real*)grid)[ix] = value[0];
				else
					for (int c = 0; c < 3; c++) ((og_real*)grid)[3 * ix + c] = value[c];
			}
}
