/*
 * Sanitizer driver for the CPU oracle (test infrastructure only).  Built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by `make sanitize-run`; it
 * runs every restated plugin of the oracle on small 2D and 3D problems (flags
 * with obstacles, stick cells and an open top, all advection types, orders
 * and clamp modes, outflow extrapolation, wall BCs with obstacle velocities,
 * buoyancy, the pressure solve with each preconditioner and zero-pressure
 * fixing, the multigrid hierarchy and shapes) so that an out-of-bounds access,
 * a use after free or undefined arithmetic in the restatement aborts the run.
 * It checks no values: the pinning tests do that.  Exit status 0 = clean.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mfx_oracle.h"

static unsigned g_seed = 12345u;
static og_real urand(og_real lo, og_real hi) {
	g_seed = g_seed * 1664525u + 1013904223u;
	return lo + (hi - lo) * (og_real)((g_seed >> 8) * (1.0 / 16777216.0));
}

static void fill(og_real* a, long n, og_real lo, og_real hi) {
	for (long i = 0; i < n; i++) a[i] = urand(lo, hi);
}

static void run_shape(int nx, int ny, int nz) {
	og_dims d = {nx, ny, nz, nz > 1};
	long n = (long)nx * ny * nz;
	int* flags = (int*)calloc((size_t)n, sizeof(int));
	og_real* vel = (og_real*)calloc((size_t)(3 * n), sizeof(og_real));
	og_real* vel2 = (og_real*)calloc((size_t)(3 * n), sizeof(og_real));
	og_real* obvel = (og_real*)calloc((size_t)(3 * n), sizeof(og_real));
	og_real* dens = (og_real*)calloc((size_t)n, sizeof(og_real));
	og_real* p = (og_real*)calloc((size_t)n, sizeof(og_real));
	og_real* rhs = (og_real*)calloc((size_t)n, sizeof(og_real));
	og_real* A[4];
	for (int a = 0; a < 4; a++) A[a] = (og_real*)calloc((size_t)n, sizeof(og_real));

	og_init_domain(d, flags, 1, "xXyYzZ", "      ", "      ", "      ");
	og_fill_grid(d, flags, OG_FLUID);
	og_set_open_bound(d, flags, 1, "Y", OG_OUTFLOW | OG_EMPTY);
	/* an obstacle block and a stick cell in the interior */
	for (int k = (d.is3d ? nz / 3 : 0); k < (d.is3d ? nz / 3 + 2 : 1); k++)
		for (int j = ny / 3; j < ny / 3 + 2; j++)
			for (int i = nx / 3; i < nx / 3 + 3; i++) flags[i + (long)nx * j + (d.is3d ? (long)nx * ny * k : 0)] = OG_OBSTACLE;
	flags[nx / 2 + (long)nx * (ny / 2) + (d.is3d ? (long)nx * ny * (nz / 2) : 0)] = OG_OBSTACLE | OG_STICK;
	fill(dens, n, 0, 1);
	og_reset_outflow(d, flags, NULL, dens);

	const int types[3] = {OG_TYPE_REAL, OG_TYPE_VEC3, OG_TYPE_MAC};
	for (int t = 0; t < 3; t++)
		for (int order = 1; order <= 2; order++)
			for (int cm = 0; cm <= 2; cm++) {
				fill(vel, 3 * n, -1.5f, 1.5f);
				og_real* g = types[t] == OG_TYPE_REAL ? dens : vel2;
				if (types[t] == OG_TYPE_REAL) fill(dens, n, 0, 1);
				else fill(vel2, 3 * n, -1, 1);
				if (types[t] == OG_TYPE_MAC && (t + order + cm) % 2 == 0) g = vel; /* self-advection */
				og_advect_semi_lagrange(d, (og_real)0.9, flags, vel, g, types[t], order, (og_real)1.0, cm);
			}
	og_interpol(d, dens, (og_real)-3.0, (og_real)(ny + 4), (og_real)0.25);
	og_interpol_component(d, vel, 2, (og_real)(nx + 1), (og_real)-1.0, (og_real)(nz + 2));
	memcpy(vel2, vel, sizeof(og_real) * 3 * n);
	og_apply_outflow_bc(d, flags, vel, vel2, (og_real)0.7);

	fill(obvel, 3 * n, -1, 1);
	og_set_wall_bcs(d, flags, vel, obvel);
	og_set_wall_bcs(d, flags, vel, NULL);
	og_add_buoyancy(d, (og_real)1.0, flags, dens, vel, 0, (og_real)-1e-3, 0, (og_real)1.0, 1);

	double sum = 0;
	og_make_rhs(d, flags, rhs, vel, &sum);
	og_make_laplace(d, flags, A[0], A[1], A[2], A[3]);
	og_apply_matrix(d, flags, p, rhs, A[0], A[1], A[2], A[3]);
	og_mic_precondition(d, flags, A[0], A[1], A[2], A[3], rhs, p);
	og_find_fix_cell(d, flags);

	const int pcs[4] = {OG_PC_NONE, OG_PC_MIC, OG_PC_MGDYN, OG_PC_MGSTATIC};
	for (int q = 0; q < 4; q++)
		for (int variant = 0; variant < 2; variant++) {
			og_pressure_opts o;
			og_solve_stats st;
			og_pressure_opts_default(&o);
			o.preconditioner = pcs[q];
			o.cgAccuracy = (og_real)1e-4;
			o.enforceCompatibility = variant;
			o.useL2Norm = variant;
			memset(p, 0, sizeof(og_real) * n);
			memcpy(vel2, vel, sizeof(og_real) * 3 * n);
			og_solve_pressure(d, vel2, p, flags, &o, &st, rhs);
		}
	/* closed domain: zero-pressure fixing engages (no empty cells) */
	og_init_domain(d, flags, 0, "xXyYzZ", "      ", "      ", "      ");
	og_fill_grid(d, flags, OG_FLUID);
	{
		og_pressure_opts o;
		og_solve_stats st;
		og_pressure_opts_default(&o);
		o.cgAccuracy = (og_real)1e-8;
		o.preconditioner = OG_PC_MGSTATIC;
		memset(p, 0, sizeof(og_real) * n);
		og_solve_pressure(d, vel, p, flags, &o, &st, NULL);
	}
	og_correct_velocity(d, flags, vel, p);

	og_make_laplace(d, flags, A[0], A[1], A[2], A[3]);
	og_mg* mg = og_mg_create(d);
	og_mg_set_a(mg, A[0], A[1], A[2], A[3]);
	og_mg_set_coarsest_accuracy(mg, (og_real)1e-6);
	og_mg_vcycle(mg, rhs, p);
	int nl = og_mg_num_levels(mg);
	for (int l = 0; l < nl; l++) {
		int sx, sy, sz;
		og_mg_level_size(mg, l, &sx, &sy, &sz);
		long ln = (long)sx * sy * sz;
		int8_t* ty = (int8_t*)malloc((size_t)ln);
		og_real* m = (og_real*)malloc(sizeof(og_real) * (size_t)ln * 14);
		og_mg_level_types(mg, l, ty);
		og_mg_level_matrix(mg, l, m);
		free(ty);
		free(m);
	}
	og_mg_destroy(mg);

	const og_real a[3] = {(og_real)(nx * 0.3), (og_real)(ny * 0.2), (og_real)(nz * 0.3)};
	const og_real b[3] = {(og_real)(nx * 0.7), (og_real)(ny * 0.6), (og_real)(nz * 0.7)};
	const og_real val[3] = {1, 2, 3};
	og_apply_shape(d, 1, a, b, 0, dens, OG_TYPE_REAL, val, NULL);
	og_apply_shape(d, 2, a, b, (og_real)(nx * 0.25), vel, OG_TYPE_MAC, val, flags);
	og_apply_shape(d, 3, a, b, (og_real)(nx * 0.2), flags, OG_TYPE_INT, val, NULL);

	for (int q = 0; q < 4; q++) free(A[q]);
	free(flags); free(vel); free(vel2); free(obvel); free(dens); free(p); free(rhs);
}

int main(void) {
	og_set_threads(2);
	run_shape(17, 14, 1);
	run_shape(12, 11, 10);
	run_shape(34, 33, 1);
	printf("oracle sanitizer run: clean\n");
	return 0;
}
