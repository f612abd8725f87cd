// Pressure projection on the GPU: RHS (MakeRhs), the Laplace matrix derived
// from cell flags, conjugate gradient (GridCg) with device-side scalars and
// convergence control, optional multigrid preconditioning, and the velocity
// correction (knCorrectVelocity).
//
// Reference: plugin/pressure.cpp:32-521, conjugategrad.h:118-187,
// conjugategrad.cpp:175-299.
//
// Matrix.  The reference stores A0/Ai/Aj/Ak as four float grids built from
// the flags every solve (MakeLaplaceMatrix, then fixPressure).  Here the same
// matrix is packed into one byte per cell ("code"): bits 0-2 A0 (0..6),
// bit 3/4/5 Ai/Aj/Ak == -1, bit 6 fluid row (else identity row), bit 7 trivial
// row (A0 == 1, no off-diagonals; used by the multigrid).  The matvec reads
// the code of the cell and of its three lower neighbours.
//
// CG.  One iteration is two streaming kernels plus two single-block
// finalisers:
//   A: p = z + beta*p (on the fly, also for the 6 neighbours) ; q = A p ;
//      partial sums of fl(q*p)                            -> 17 B/cell
//   FA: alpha = sigma / (float)dot
//   B: x += alpha p ; r -= alpha q ; min/max(r), sum fl(z*r), sum r^2
//                                                         -> 24 B/cell
//   FB: resNorm, convergence, sigma/beta, divergence, iteration count
// With PcNone z aliases r.  Dots are accumulated in double from float
// products, scalars are rounded to float, exactly as GridCg does, so the
// iterates follow the reference's serial FP32 build up to the summation
// order of the double dot products.  Kernels read a device "done" flag and
// return at once after convergence, so the host can queue iterations in
// batches and only synchronise between batches.
#include <algorithm>
#include <type_traits>
#include <array>
#include <map>
#include <set>
#include <mutex>
#include <chrono>
#include <atomic>
#include <exception>
#include <thread>
#include <deque>
#include <cmath>
#include <cstring>

#include "mfx_batch.h"
#include "mfx_internal.h"
#include "mfx_mg.h"
#include "mfx_prof.h"

namespace mfx {

__device__ inline bool interior_c(const Dims& d, int i, int j, int k) {
	return i >= 1 && i < d.nx - 1 && j >= 1 && j < d.ny - 1 && (d.is3d ? (k >= 1 && k < d.nz - 1) : true);
}

// ---------------------------------------------------------------------------
// matrix code (MakeLaplaceMatrix conjugategrad.h:154-171 + fixPressure
// pressure.cpp:226-245)
// ---------------------------------------------------------------------------
__device__ inline int off_coef(const Dims& d, const uint8_t* f8, long long c, int ci, int cj, int ck,
                               long long nb, long long fix) {
	// coefficient stored at c towards its upper neighbour nb: -1 iff c is an
	// interior fluid cell, nb is fluid, and neither is the pinned cell
	if (!interior_c(d, ci, cj, ck)) return 0;
	if (!(f8[c] & kFluid) || !(f8[nb] & kFluid)) return 0;
	if (c == fix || nb == fix) return 0;
	return 1;
}

// The matrix code byte of cell c = (i,j,k) from its flags and its six
// neighbours' (used only where they exist) and the pinned cell fix (-1 for
// none): the rules as 0/1 integer predicates (vector ALU, no branch per
// condition), the same values as MakeLaplaceMatrix + fixPressure give.
__device__ inline uint8_t code_byte(const Dims& d, int i, int j, int k, long long c, long long fix, int f, int fxm,
                                   int fxp, int fym, int fyp, int fzm, int fzp) {
	const int is3 = d.is3d ? 1 : 0;
	const int gi = min(i, 1), li = min(d.nx - 1 - i, 1), g2i = min(max(i - 1, 0), 1);  // i>=1, i<nx-1, i>=2
	const int gj = min(j, 1), lj = min(d.ny - 1 - j, 1), g2j = min(max(j - 1, 0), 1);
	const int gk = min(k, 1), lk = min(d.nz - 1 - k, 1), g2k = min(max(k - 1, 0), 1);
	const int kin = is3 ? (gk & lk) : 1;  // k inside the interior (always in 2D)
	const int fl = f & 1;                  // kFluid = 1, kObstacle = 2
	const int inner = gi & li & gj & lj & kin;
	const int isfix = c == fix ? 1 : 0, nfix = isfix ^ 1;
	const int fi = fl & inner;
	const int nobs = (((fxm >> 1) & 1) ^ 1) + (((fxp >> 1) & 1) ^ 1) + (((fym >> 1) & 1) ^ 1) +
	                 (((fyp >> 1) & 1) ^ 1) + is3 * ((((fzm >> 1) & 1) ^ 1) + (((fzp >> 1) & 1) ^ 1));
	const int A0 = fi * (isfix + nfix * nobs);
	// off_coef(c -> upper neighbour): c interior fluid, neighbour fluid, neither pinned
	const int bx = fi & nfix & fxp & (c + 1 != fix ? 1 : 0);
	const int by = fi & nfix & fyp & (c + d.nx != fix ? 1 : 0);
	const int bz = fi & nfix & is3 & fzp & (c + d.sz != fix ? 1 : 0);
	// any lower off-diagonal into c: off_coef(lower neighbour -> c)
	const int lox = gi & g2i & gj & lj & kin & fxm & (c - 1 != fix ? 1 : 0);
	const int loy = gj & g2j & gi & li & kin & fym & (c - d.nx != fix ? 1 : 0);
	const int loz = is3 & gk & g2k & gi & li & gj & lj & fzm & (c - d.sz != fix ? 1 : 0);
	const int lo = fl & nfix & (lox | loy | loz) & 1;
	const int trivial = (A0 == 1 ? 1 : 0) & ((bx | by | bz | lo) & 1 ^ 1);
	return (uint8_t)(A0 | ((bx & 1) << 3) | ((by & 1) << 4) | ((bz & 1) << 5) | (fl << 6) | (trivial << 7));
}

// kLightKZ planes per thread; the seven flags of every plane are loaded
// first (offsets 0 where the neighbour is outside the grid: those flags are
// then never used).
__global__ void __launch_bounds__(kBlock) k_build_code(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                       uint8_t* __restrict__ code, const CgState* __restrict__ st) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const long long fix = st->fix;
	const int i = p.i, j = p.j;
	constexpr int KZ = kLightKZ;
	const long long oxm = i > 0 ? 1 : 0, oxp = i < d.nx - 1 ? 1 : 0;
	const long long oym = j > 0 ? d.nx : 0, oyp = j < d.ny - 1 ? d.nx : 0;
	uint8_t f[KZ], fxm[KZ], fxp[KZ], fym[KZ], fyp[KZ], fzm[KZ], fzp[KZ];
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = min(p.k + q, d.nz - 1);
		const long long c = d.idx(i, j, k);
		const long long ozm = (d.is3d && k > 0) ? d.sz : 0, ozp = (d.is3d && k < d.nz - 1) ? d.sz : 0;
		f[q] = f8[c];
		fxm[q] = f8[c - oxm];
		fxp[q] = f8[c + oxp];
		fym[q] = f8[c - oym];
		fyp[q] = f8[c + oyp];
		fzm[q] = f8[c - ozm];
		fzp[q] = f8[c + ozp];
	}
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = p.k + q;
		if (k >= d.nz) break;
		const long long c = d.idx(i, j, k);
		code[c] = code_byte(d, i, j, k, c, fix, f[q], fxm[q], fxp[q], fym[q], fyp[q], fzm[q], fzp[q]);
	}
}

__device__ inline float cA0(uint8_t c) { return (float)(c & 7); }
__device__ inline float cAi(uint8_t c) { return (c & 8) ? -1.0f : 0.0f; }
__device__ inline float cAj(uint8_t c) { return (c & 16) ? -1.0f : 0.0f; }
__device__ inline float cAk(uint8_t c) { return (c & 32) ? -1.0f : 0.0f; }
// Scaled matrices (cgSolveDiffusion, conjugategrad.cpp:355-373; cgSolveWE,
// waves.cpp:106-115): A0 (= cnt, a float) *= scale; A0 += 1. (a double sum
// stored as float); off-diagonals -1 * scale
__device__ inline float sA0(uint8_t c, float scale) {
	const float a = (float)(c & 7) * scale;
	return (float)((double)a + 1.0);
}
__device__ inline float sOff(uint8_t c, int bit, float scale) { return (c & bit) ? -scale : 0.0f; }

// ApplyMatrix / ApplyMatrix2D at cell c (conjugategrad.h:118-151) with
// src values given by a functor; out-of-range linear neighbours contribute 0
template <typename SRC>
__device__ inline float apply_row(const Dims& d, const uint8_t* __restrict__ code, long long c, uint8_t cc,
                                  SRC src) {
	const long long X = 1, Y = d.nx, Z = d.sz, n = d.n;
	const float s0 = src(c);
	if (!(cc & 64)) return s0;
	const float sxm = c - X >= 0 ? src(c - X) : 0.f, cxm = c - X >= 0 ? cAi(code[c - X]) : 0.f;
	const float sxp = c + X < n ? src(c + X) : 0.f;
	const float sym = c - Y >= 0 ? src(c - Y) : 0.f, cym = c - Y >= 0 ? cAj(code[c - Y]) : 0.f;
	const float syp = c + Y < n ? src(c + Y) : 0.f;
	if (d.is3d) {
		const float szm = c - Z >= 0 ? src(c - Z) : 0.f, czm = c - Z >= 0 ? cAk(code[c - Z]) : 0.f;
		const float szp = c + Z < n ? src(c + Z) : 0.f;
		return s0 * cA0(cc) + sxm * cxm + sxp * cAi(cc) + sym * cym + syp * cAj(cc) + szm * czm + szp * cAk(cc);
	}
	return s0 * cA0(cc) + sxm * cxm + sxp * cAi(cc) + sym * cym + syp * cAj(cc);
}

// ---------------------------------------------------------------------------
// RHS: MakeRhs (pressure.cpp:32-84) without fractions/obvel/surface tension
// ---------------------------------------------------------------------------
// kLightKZ planes per thread, all loads issued first (upper-neighbour offsets
// are 0 on the last layer, where the value is not used).  CODE: the same
// pass also writes the matrix code (k_build_code's values with no pinned
// cell), from the six neighbours' flags.
template <bool CODE>
__global__ void __launch_bounds__(kBlock) k_make_rhs(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                     const float* __restrict__ u, const float* __restrict__ v,
                                                     const float* __restrict__ w, float* __restrict__ rhs,
                                                     double* __restrict__ part, uint8_t* __restrict__ code) {
	__shared__ double sh[8];
	const CellPos p = tile_cell(t);
	double sum = 0, cnt = 0;
	if (p.i < d.nx && p.j < d.ny) {
		constexpr int KZ = kLightKZ;
		const int i = p.i, j = p.j;
		const long long ox = i < d.nx - 1 ? 1 : 0, oy = j < d.ny - 1 ? d.nx : 0;
		const long long oxm = i > 0 ? 1 : 0, oym = j > 0 ? d.nx : 0;
		uint8_t f[KZ], fxm[KZ], fxp[KZ], fym[KZ], fyp[KZ], fzm[KZ], fzp[KZ];
		float u0[KZ], u1[KZ], v0[KZ], v1[KZ], w0[KZ], w1[KZ];
#pragma unroll
		for (int q = 0; q < KZ; q++) {
			const int k = min(p.k + q, d.nz - 1);
			const long long c = d.idx(i, j, k);
			f[q] = f8[c];
			if (CODE) {
				fxm[q] = f8[c - oxm];
				fxp[q] = f8[c + ox];
				fym[q] = f8[c - oym];
				fyp[q] = f8[c + oy];
			}
			u0[q] = u[c];
			u1[q] = u[c + ox];
			v0[q] = v[c];
			v1[q] = v[c + oy];
			w0[q] = w[c];
		}
		// z-neighbours from the thread's adjacent planes (clamped planes
		// repeat the last one, as the zero offsets there do); the planes
		// below the first and above the last are loaded
		{
			const int kl = min(p.k + KZ - 1, d.nz - 1), k0 = min(p.k, d.nz - 1);
			const long long cl = d.idx(i, j, kl) + ((d.is3d && kl < d.nz - 1) ? d.sz : 0);
			w1[KZ - 1] = w[cl];
			if (CODE) {
				fzp[KZ - 1] = f8[cl];
				fzm[0] = f8[d.idx(i, j, k0) - ((d.is3d && k0 > 0) ? d.sz : 0)];
			}
		}
#pragma unroll
		for (int q = 0; q < KZ - 1; q++) {
			w1[q] = w0[q + 1];
			if (CODE) {
				fzp[q] = f[q + 1];
				fzm[q + 1] = f[q];
			}
		}
#pragma unroll
		for (int q = 0; q < KZ; q++) {
			const int k = p.k + q;
			if (k >= d.nz) break;
			if (CODE)
				code[d.idx(i, j, k)] =
				    code_byte(d, i, j, k, d.idx(i, j, k), -1, f[q], fxm[q], fxp[q], fym[q], fyp[q], fzm[q], fzp[q]);
			float set = 0.f;
			if (interior_c(d, i, j, k) && (f[q] & kFluid)) {
				set = u0[q] - u1[q] + v0[q] - v1[q];
				if (d.is3d) set += w0[q] - w1[q];
				sum += (double)set;
				cnt += 1;
			}
			rhs[d.idx(i, j, k)] = set;
		}
	}
	const double s = block_sum(sum, sh);
	const double n = block_sum(cnt, sh);
	if (threadIdx.x == 0) {
		part[2 * blockIdx.x] = s;
		part[2 * blockIdx.x + 1] = n;
	}
}

__global__ void k_add_const(float* a, long long n, float v) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
		a[i] += v;
}

// knCorrectVelocity (pressure.cpp:87-109), bnd=1.  kLightKZ planes per
// thread with all loads issued first; each face is updated in a register in
// the reference's order and written once.
template <bool WALL>
__device__ __forceinline__ void correct_velocity_cells(const CorrVelP& P, long long blk) {
	const Dims& d = P.d;
	const uint8_t* __restrict__ f8 = P.f8;
	const float* __restrict__ pr = P.pr;
	float* __restrict__ u = P.u;
	float* __restrict__ v = P.v;
	float* __restrict__ w = P.w;
	const CellPos p = tile_cell_of(P.t, (unsigned)blk);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const int i = p.i, j = p.j;
	constexpr int KZ = kLightKZ;
	// neighbour offsets, valid for every cell (interior cells use the true ones)
	const long long ox = i > 0 ? 1 : 0, oy = j > 0 ? d.nx : 0;
	const long long oxp = i < d.nx - 1 ? 1 : 0, oyp = j < d.ny - 1 ? d.nx : 0;
	uint8_t f[KZ], gx[KZ], gy[KZ], gz[KZ], hx[KZ], hy[KZ], hz[KZ];
	float pc[KZ], px[KZ], py[KZ], pz[KZ], uc[KZ], vc[KZ], wc[KZ];
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = min(p.k + q, d.nz - 1);
		const long long c = d.idx(i, j, k);
		f[q] = f8[c];
		gx[q] = f8[c - ox];
		gy[q] = f8[c - oy];
		if (WALL) {
			hx[q] = f8[c + oxp];
			hy[q] = f8[c + oyp];
		}
		pc[q] = pr[c];
		px[q] = pr[c - ox];
		py[q] = pr[c - oy];
		uc[q] = u[c];
		vc[q] = v[c];
		wc[q] = w[c];
	}
	// z-neighbours: the plane below the first is loaded; for the others it is
	// the thread's previous plane (planes past nz are clamped and never used);
	// the plane above the last one likewise
	{
		const int k0 = min(p.k, d.nz - 1);
		const long long c0 = d.idx(i, j, k0) - (k0 > 0 ? d.sz : 0);
		gz[0] = f8[c0];
		pz[0] = pr[c0];
	}
#pragma unroll
	for (int q = 1; q < KZ; q++) {
		gz[q] = f[q - 1];
		pz[q] = pc[q - 1];
	}
	if (WALL) {
#pragma unroll
		for (int q = 0; q < KZ - 1; q++) hz[q] = f[q + 1];
		const int k3 = min(p.k + KZ - 1, d.nz - 1);
		hz[KZ - 1] = f8[d.idx(i, j, k3) + ((d.is3d && k3 < d.nz - 1) ? d.sz : 0)];
	}
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = p.k + q;
		if (k >= d.nz) break;
		const long long c = d.idx(i, j, k);
		const uint8_t fc = f[q];
		float nu = uc[q], nv = vc[q], nw = wc[q];
		bool cor = interior_c(d, i, j, k);
		if (cor) {
			if (fc & kFluid) {
				const float pcc = pc[q];
				if (gx[q] & kFluid) nu -= (pcc - px[q]);
				if (gy[q] & kFluid) nv -= (pcc - py[q]);
				if (d.is3d && (gz[q] & kFluid)) nw -= (pcc - pz[q]);
				if (gx[q] & kEmpty) nu -= pcc;
				if (gy[q] & kEmpty) nv -= pcc;
				if (d.is3d && (gz[q] & kEmpty)) nw -= pcc;
			} else if ((fc & kEmpty) && !(fc & kOutflow)) {
				if (gx[q] & kFluid) nu += px[q];
				else nu = 0.f;
				if (gy[q] & kFluid) nv += py[q];
				else nv = 0.f;
				if (d.is3d) {
					if (gz[q] & kFluid) nw += pz[q];
					else nw = 0.f;
				}
			} else {
				cor = false;
			}
		}
		if (WALL) {
			// then setWallBcs (no obstacle velocity): written faces take 0
			const WallRule r = wall_rule(d, i, j, k, fc, gx[q], hx[q], gy[q], hy[q], gz[q], hz[q]);
			if (r.wu) nu = 0.f;
			if (r.wv) nv = 0.f;
			if (r.ww) nw = 0.f;
			if (cor || r.wu) u[c] = nu;
			if (cor || r.wv) v[c] = nv;
			if ((cor && d.is3d) || r.ww) w[c] = nw;
		} else if (cor) {
			u[c] = nu;
			v[c] = nv;
			if (d.is3d) w[c] = nw;
		}
	}
}

__device__ __forceinline__ void body_correct_velocity(const CorrVelP& P, long long blk, long long) {
	correct_velocity_cells<false>(P, blk);
}
// correctVelocity then setWallBcs(flags, vel) in one pass: each face's
// corrected value is zeroed where KnSetWallBcs writes it (no obstacle
// velocity), and written once
__device__ __forceinline__ void body_correct_velocity_wall(const CorrVelP& P, long long blk, long long) {
	correct_velocity_cells<true>(P, blk);
}

// ---------------------------------------------------------------------------
// CG kernels
// ---------------------------------------------------------------------------
// doInit (conjugategrad.cpp:210-235): x = 0, r = b, partial sums of
// fl(b*b) (GridDotProduct of the unpreconditioned tmp = r) and of b^2 in
// FP64; float4 groups with a grid stride of per*256 (as body_update), the
// n % 4 trailing cells spread the same way; one partial per block
__global__ void __launch_bounds__(kBlock) k_cg_init(Dims d, long long per, const float* __restrict__ rhs,
                                                    float* __restrict__ x, float* __restrict__ r, int withDot,
                                                    double* __restrict__ part) {
	__shared__ double sh[8];
	double zr = 0, bb = 0;
	const long long n4 = d.n >> 2;
	const long long stride = per * kBlock;
	const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
	for (long long v = blockIdx.x * (long long)kBlock + threadIdx.x; v < n4; v += stride) {
		const float4 bv = reinterpret_cast<const float4*>(rhs)[v];
		reinterpret_cast<float4*>(x)[v] = zero4;
		reinterpret_cast<float4*>(r)[v] = bv;
		const float e[4] = {bv.x, bv.y, bv.z, bv.w};
		for (int u = 0; u < 4; u++) {
			zr += (double)(e[u] * e[u]);
			bb += (double)e[u] * (double)e[u];
		}
	}
	for (long long c = (n4 << 2) + blockIdx.x * (long long)kBlock + threadIdx.x; c < d.n; c += stride) {
		const float b = rhs[c];
		x[c] = 0.f;
		r[c] = b;
		zr += (double)(b * b);
		bb += (double)b * (double)b;
	}
	const double s0 = block_sum(zr, sh);
	const double s1 = block_sum(bb, sh);
	if (threadIdx.x == 0) {
		part[3 * blockIdx.x] = withDot ? s0 : 0.0;
		part[3 * blockIdx.x + 1] = s1;
		part[3 * blockIdx.x + 2] = 0.0;
	}
}

// CgState of a new solve (set on the device: no host copy to wait for)
__global__ void k_cg_state_init(CgState* st, float acc, int maxIter, int useL2, double relTol) {
	if (threadIdx.x != 0) return;
	CgState z;
	memset(&z, 0, sizeof(z));
	z.acc = acc;
	z.maxIter = maxIter;
	z.useL2 = useL2;
	z.relTol = relTol;
	z.fix = -1;
	z.fixFirst = ~0ull;
	*st = z;
}

// CG restart (cg_restart): iterate again from the current x; FIN_BETA0 then
// sets sigma from the replaced residual
__global__ void k_cg_restart(CgState* st) {
	if (threadIdx.x != 0) return;
	st->done = 0;
	st->converged = 0;
	st->first = 1;
	st->beta = 0.f;
}

// Zero-pressure cell selection (pressure.cpp:349-381) on the device, in two
// kernels: k_fix_scan notes whether any cell is empty (then nothing is
// pinned) and the first fluid cell of the bnd=1 interior in ijk order, i.e.
// the smallest such index; k_fix_select takes the first fluid cell among the
// top-centre positions (nx/2, ny-1-o, nz/2), o = 0, 1, 2, else that cell,
// and zeroes its right-hand side (fixPressure, pressure.cpp:226-245).
__global__ void __launch_bounds__(kBlock) k_fix_scan(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                     CgState* st) {
	const CellPos p = tile_cell(t);
	const bool in = p.i < d.nx && p.j < d.ny;
	const long long c = in ? d.idx(p.i, p.j, p.k) : 0;
	const uint8_t f = in ? f8[c] : (uint8_t)0;
	const bool inner = p.i >= 1 && p.i < d.nx - 1 && p.j >= 1 && p.j < d.ny - 1 &&
	                   (!d.is3d || (p.k >= 1 && p.k < d.nz - 1));
	unsigned long long first = (in && inner && (f & kFluid)) ? (unsigned long long)c : ~0ull;
	for (int o = 32; o > 0; o >>= 1) {
		const unsigned long long q = __shfl_xor(first, o, 64);
		first = q < first ? q : first;
	}
	const bool anyEmpty = __ballot((f & kEmpty) != 0) != 0;
	if ((threadIdx.x & 63) == 0) {
		if (first != ~0ull) atomicMin(&st->fixFirst, first);
		if (anyEmpty) atomicOr(&st->fixAnyEmpty, 1);
	}
}
__global__ void k_fix_select(Dims d, const uint8_t* __restrict__ f8, CgState* st, float* __restrict__ rhs) {
	if (threadIdx.x != 0) return;
	long long fix = -1;
	if (!st->fixAnyEmpty) {
		const int tcx = d.nx / 2, tcy = d.ny - 1, tcz = d.is3d ? d.nz / 2 : 0;
		for (int o = 0; o < 3 && fix < 0; o++) {
			const int y = tcy - o;
			if (y < 0) continue;
			const long long c = d.idx(tcx, y, tcz);
			if (f8[c] & kFluid) fix = c;
		}
		if (fix < 0 && st->fixFirst != ~0ull) fix = (long long)st->fixFirst;
	}
	st->fix = fix;
	if (fix >= 0) rhs[fix] = 0.f;
}

// enforceCompatibility (pressure.cpp:297-298): rhs += (Real)(-sum / (Real)cnt)
// over every cell; sum and cnt are the FP64 partials of k_make_rhs (pairs
// per block), added in a fixed order by one block
__global__ void __launch_bounds__(1024) k_compat_corr(const double* __restrict__ part, int nb,
                                                      float* __restrict__ corr) {
	__shared__ double sh[2][16];
	double sum = 0, cnt = 0;
	for (int b = threadIdx.x; b < nb; b += blockDim.x) {
		sum += part[2 * b];
		cnt += part[2 * b + 1];
	}
	sum = wave_sum(sum);
	cnt = wave_sum(cnt);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) {
		sh[0][w] = sum;
		sh[1][w] = cnt;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double a = 0, c = 0;
		for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
			a += sh[0][q];
			c += sh[1][q];
		}
		*corr = (float)(-a / (double)(float)c);
	}
}
__global__ void k_add_const_dev(float* a, long long n, const float* __restrict__ v) {
	const float c = *v;
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
		a[i] += c;
}

// Matrix code of cgSolveDiffusion (conjugategrad.cpp:355-373): the Laplace
// matrix of an all-fluid grid (every bnd=1 interior cell: 4 or 6 counted
// neighbours, off-diagonals toward every upper neighbour), obstacle rows of
// the real flags replaced by identity rows; the fluid bit (ApplyMatrix's
// identity test) is the real flags'.  Scaled by alpha in the matvec (sA0, sOff).
__global__ void __launch_bounds__(kBlock) k_build_code_diff(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                            uint8_t* __restrict__ code) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const bool inIJ = p.i >= 1 && p.i < d.nx - 1 && p.j >= 1 && p.j < d.ny - 1;
#pragma unroll
	for (int q = 0; q < kLightKZ; q++) {
		const int k = p.k + q;
		if (k >= d.nz) break;
		const long long c = d.idx(p.i, p.j, k);
		const uint8_t f = f8[c];
		const bool inner = inIJ && (!d.is3d || (k >= 1 && k < d.nz - 1));
		const bool rows = inner && !(f & kObstacle);
		const int cnt = rows ? (d.is3d ? 6 : 4) : 0;
		const int off = rows ? (8 | 16 | (d.is3d ? 32 : 0)) : 0;
		code[c] = (uint8_t)(cnt | off | ((f & kFluid) ? 64 : 0));
	}
}

// MakeRhsWE (waves.cpp:69-79) after rhs.clear(): bnd=1 cells get
// (Real)(2 ut - utm1) in double; with crankNic, rhs += s * (x-y 5-point
// stencil of ut) (double, stored as float); every other cell 0
__global__ void __launch_bounds__(kBlock) k_make_rhs_we(Dims d, Tiles t, const float* __restrict__ ut,
                                                        const float* __restrict__ utm1, float s, int crankNic,
                                                        float* __restrict__ rhs) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const bool inIJ = p.i >= 1 && p.i < d.nx - 1 && p.j >= 1 && p.j < d.ny - 1;
	const long long X = 1, Y = d.nx;
#pragma unroll
	for (int q = 0; q < kLightKZ; q++) {
		const int k = p.k + q;
		if (k >= d.nz) break;
		const long long c = d.idx(p.i, p.j, k);
		float v = 0.f;
		if (inIJ && (!d.is3d || (k >= 1 && k < d.nz - 1))) {
			v = (float)(2. * (double)ut[c] - (double)utm1[c]);
			if (crankNic)
				v = (float)((double)v + (double)s * (-4. * (double)ut[c] + 1. * (double)ut[c - X] +
				                                     1. * (double)ut[c + X] + 1. * (double)ut[c - Y] +
				                                     1. * (double)ut[c + Y]));
		}
		rhs[c] = v;
	}
}

// st->trueRR = sum of nb true-residual partials (fixed order; 1024 threads)
__device__ __forceinline__ void reduce_true_rr(CgState* st, const double* __restrict__ part, int nb) {
	__shared__ double sh[16];
	double a = 0;
	for (int b = threadIdx.x; b < nb; b += blockDim.x) a += part[3 * b];
	a = wave_sum(a);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
	__syncthreads();
	if (threadIdx.x == 0) {
		double s = 0;
		for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += sh[q];
		st->trueRR = s;
	}
}
__global__ void __launch_bounds__(1024) k_reduce_true_rr(CgState* st, const double* __restrict__ part, int nb) {
	reduce_true_rr(st, part, nb);
}
__device__ __forceinline__ void body_reduce_true_rr(const TrueRedP& P, long long, long long) {
	reduce_true_rr(P.st, P.part, P.nb);
}
// every problem's CgState into one array (one copy to the host for a batch)
__global__ void k_gather_states(int np, CgState* const* __restrict__ st, CgState* __restrict__ out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < np) out[i] = *st[i];
}

// sum fl(z*r) (GridDotProduct(mTmp, mResidual)) of 256 cells into part[3*slot]
__device__ __forceinline__ void body_dot(const DotP& P, long long blk, long long slot) {
	__shared__ double sh[8];
	if (P.st->done) return;
	const long long c = blk * kBlock + threadIdx.x;
	double a = 0;
	if (c < P.d.n) a = (double)(P.z[c] * P.r[c]);
	const double s = block_sum(a, sh);
	if (threadIdx.x == 0) P.part[3 * slot] = s;
}

// A: p_new = z + beta*p_old ; q = A p_new ; partial sums of fl(q*p_new).
// z-marching form: a 64 x TY tile of columns walks a chunk of planes.  The
// search direction of the current plane (tile + 1-cell ring) is staged in
// LDS, the planes below and above stay in registers, so every value is read
// from HBM once.  Row arithmetic and term order follow apply_row; neighbours
// outside the grid contribute 0 (their coefficient is 0 there).
//
// Loads run one plane ahead as raw values (z, p_old, code): p_new = z +
// beta p_old is formed only when a plane enters LDS, so the loads of plane
// k+2 are in flight during the whole step of plane k.
template <int TY, bool SC = false>
__device__ __forceinline__ void body_matvec(const MatvecP& P, long long blk, long long slot) {
	constexpr int TX = 64, RX = TX + 2, RY = TY + 2, NH = 2 * RX + 2 * (RY - 2);
	// plane buffers; the 64 slots past the ring take the writes of lanes
	// that have no ring cell
	__shared__ float ps[2][RX * RY + 64];
	__shared__ uint8_t cs[2][RX * RY + 64];
	__shared__ double tab[kZmMaxGroups][TY];
	const Dims& d = P.d;
	const int kchunk = P.kchunk;
	const CgState* __restrict__ st = P.st;
	const uint8_t* __restrict__ code = P.code;
	if (st->done) return;
	const float beta = st->beta;
	const bool first = st->first;
	const float* __restrict__ z = P.z;
	// in the first iteration p_old is not used: its loads read z's lines
	// (cache hits) instead of a second stream
	const float* __restrict__ pold = first ? P.z : P.pold;
	float* __restrict__ pnew = P.pnew;
	float* __restrict__ q = P.q;
	const unsigned ntx = (unsigned)(d.nx + TX - 1) / TX, nty = (unsigned)(d.ny + TY - 1) / TY;
	const unsigned b32 = (unsigned)blk, rest = b32 / ntx;
	const int ox = (int)(b32 - rest * ntx) * TX;
	const int oy = (int)(rest % nty) * TY;
	const int k0 = (int)(rest / nty) * kchunk, k1 = min(d.nz, k0 + kchunk);
	const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	const int gx = ox + tx, gy = oy + ty;
	const bool in = gx < d.nx && gy < d.ny;
	// one ring cell per thread (threads < NH)
	static_assert(NH <= 64 * TY, "one ring cell per thread");
	const int h = threadIdx.x;
	int hx = 0, hy = 0;
	if (h < RX) { hx = h; hy = 0; }
	else if (h < 2 * RX) { hx = h - RX; hy = RY - 1; }
	else if (h < 2 * RX + RY - 2) { hx = 0; hy = 1 + h - 2 * RX; }
	else { hx = RX - 1; hy = 1 + h - 2 * RX - (RY - 2); }
	const int hgx = ox - 1 + hx, hgy = oy - 1 + hy;
	const bool hin = h < NH && hgx >= 0 && hgy >= 0 && hgx < d.nx && hgy < d.ny;
	const long long gcol = in ? (long long)gx + (long long)d.nx * gy : 0;
	const long long hcol = hin ? (long long)hgx + (long long)d.nx * hgy : 0;
	// Raw values of one plane: own column (z, p_old, code) and ring cell
	// (same).  The six loads are issued without branches (a plane outside
	// the tile's planes k0-1..k1 or the grid reads a clamped plane, a lane
	// outside the grid reads column 0; in the first iteration the p_old
	// loads read z and are ignored), so the compiler can wait for one
	// plane's loads while the next plane's stay in flight; validity is
	// applied where the values are used.
	struct Raw {
		float z, po, hz, hpo;
		uint8_t c, hc;
		bool ok;
	};
	const int kLo = max(k0 - 1, 0), kHi = min(k1, d.nz - 1);
	auto load = [&](int kk, Raw& r) {
		const long long pl = d.sz * min(max(kk, kLo), kHi);
		r.z = z[gcol + pl];
		r.po = pold[gcol + pl];
		r.c = code[gcol + pl];
		r.hz = z[hcol + pl];
		r.hpo = pold[hcol + pl];
		r.hc = code[hcol + pl];
		r.ok = kk >= 0 && kk < d.nz;
	};
	auto pn = [&](float zv, float pv) -> float { return first ? zv : zv + beta * pv; };
	auto own = [&](const Raw& r) -> float { return in && r.ok ? pn(r.z, r.po) : 0.f; };
	auto ownc = [&](const Raw& r) -> uint8_t { return in && r.ok ? r.c : (uint8_t)0; };
	auto ring = [&](const Raw& r) -> float { return hin && r.ok ? pn(r.hz, r.hpo) : 0.f; };
	auto ringc = [&](const Raw& r) -> uint8_t { return hin && r.ok ? r.hc : (uint8_t)0; };
	const int L = (tx + 1) + RX * (ty + 1);
	const int HS = h < NH ? hx + RX * hy : RX * RY + tx;  // ring slot or spare
	float pPrev, pCur, pNext;
	uint8_t cPrev, cCur, cNext;
	Raw R;
	{
		load(k0 - 1, R);
		pPrev = own(R);
		cPrev = ownc(R);
		load(k0, R);
		pCur = own(R);
		cCur = ownc(R);
		ps[0][L] = pCur;
		cs[0][L] = cCur;
		ps[0][HS] = ring(R);
		cs[0][HS] = ringc(R);
	}
	load(k0 + 1, R);
	__syncthreads();
	double acc = 0.0;
	int b = 0;  // LDS buffer of plane k
	for (int k = k0; k < k1; k++, b ^= 1) {
		// R holds plane k+1: it goes to LDS buffer b^1 (whose plane k-1 was
		// last read before the previous barrier; after the last plane the
		// write is unused) without a branch, and only then is R reloaded
		// with plane k+2, so R keeps its registers across iterations
		pNext = own(R);
		cNext = ownc(R);
		ps[b ^ 1][L] = pNext;
		cs[b ^ 1][L] = cNext;
		ps[b ^ 1][HS] = ring(R);
		cs[b ^ 1][HS] = ringc(R);
		__builtin_amdgcn_sched_barrier(0);
		load(k + 2, R);
		if (in) {
			const float* __restrict__ pp = ps[b];
			const uint8_t* __restrict__ cp = cs[b];
			float qv;
			if (!(cCur & 64)) qv = pCur;
			else if (SC) {
				const float sc = P.scale;
				qv = pCur * sA0(cCur, sc) + pp[L - 1] * sOff(cp[L - 1], 8, sc) + pp[L + 1] * sOff(cCur, 8, sc) +
				     pp[L - RX] * sOff(cp[L - RX], 16, sc) + pp[L + RX] * sOff(cCur, 16, sc);
				if (d.is3d) qv = qv + pPrev * sOff(cPrev, 32, sc) + pNext * sOff(cCur, 32, sc);
			} else if (d.is3d)
				qv = pCur * cA0(cCur) + pp[L - 1] * cAi(cp[L - 1]) + pp[L + 1] * cAi(cCur) + pp[L - RX] * cAj(cp[L - RX]) +
				     pp[L + RX] * cAj(cCur) + pPrev * cAk(cPrev) + pNext * cAk(cCur);
			else
				qv = pCur * cA0(cCur) + pp[L - 1] * cAi(cp[L - 1]) + pp[L + 1] * cAi(cCur) + pp[L - RX] * cAj(cp[L - RX]) +
				     pp[L + RX] * cAj(cCur);
			const long long m = gcol + d.sz * k;
			pnew[m] = pCur;
			q[m] = qv;
			acc += (double)(qv * pCur);
		}
		if ((k + 1) % kDotGroup == 0 || k + 1 == k1) zm_dot_flush<TY>(acc, tab, k, k0);
		__syncthreads();
		pPrev = pCur;
		cPrev = cCur;
		pCur = pNext;
		cCur = cNext;
	}
	(void)slot;
	zm_dot_store<TY>(tab, d, zm_column(blk, d, TY), k0, k1, P.part);
}

// B: x += alpha p ; r += (-alpha) q ; min/max(r), sum fl(r*r), sum r^2.
// Tile blk of per (= tile count) tiles walks the float4 groups with a stride
// of per*256 (grid-stride order: concurrent tiles touch adjacent memory); the
// n % 4 trailing cells are spread the same way.
__device__ __forceinline__ void body_update(const UpdateP& P, long long blk, long long slot) {
	__shared__ double sh[8];
	__shared__ float shf[8];
	const Dims& d = P.d;
	const CgState* __restrict__ st = P.st;
	float* __restrict__ x = P.x;
	float* __restrict__ r = P.r;
	const float* __restrict__ p = P.p;
	const float* __restrict__ q = P.q;
	if (st->done) return;
	const float alpha = st->alpha, malpha = -alpha;
	float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
	double zr = 0, rr = 0;
	const long long n4 = d.n >> 2;
	const long long stride = P.per * kBlock;
	for (long long v = blk * kBlock + threadIdx.x; v < n4; v += stride) {
		float4 xv = reinterpret_cast<const float4*>(x)[v];
		float4 rv = reinterpret_cast<const float4*>(r)[v];
		const float4 pv = reinterpret_cast<const float4*>(p)[v];
		const float4 qv = reinterpret_cast<const float4*>(q)[v];
		xv.x = xv.x + alpha * pv.x;
		xv.y = xv.y + alpha * pv.y;
		xv.z = xv.z + alpha * pv.z;
		xv.w = xv.w + alpha * pv.w;
		rv.x = rv.x + malpha * qv.x;
		rv.y = rv.y + malpha * qv.y;
		rv.z = rv.z + malpha * qv.z;
		rv.w = rv.w + malpha * qv.w;
		reinterpret_cast<float4*>(x)[v] = xv;
		reinterpret_cast<float4*>(r)[v] = rv;
		const float e[4] = {rv.x, rv.y, rv.z, rv.w};
		for (int u = 0; u < 4; u++) {
			if (e[u] < mn) mn = e[u];
			if (e[u] > mx) mx = e[u];
			zr += (double)(e[u] * e[u]);
			rr += (double)e[u] * (double)e[u];
		}
	}
	for (long long c = (n4 << 2) + blk * kBlock + threadIdx.x; c < d.n; c += stride) {
		const float xn = x[c] + alpha * p[c];
		const float rn = r[c] + malpha * q[c];
		x[c] = xn;
		r[c] = rn;
		if (rn < mn) mn = rn;
		if (rn > mx) mx = rn;
		zr += (double)(rn * rn);
		rr += (double)rn * (double)rn;
	}
	const double s0 = block_sum(zr, sh);
	const double s1 = block_sum(rr, sh);
	mn = wave_min(mn);
	mx = wave_max(mx);
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	if (lane == 0) {
		shf[wv] = mn;
		shf[4 + wv] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		float a = shf[0], b = shf[4];
		for (int k = 1; k < 4; k++) {
			a = fminf(a, shf[k]);
			b = fmaxf(b, shf[4 + k]);
		}
		P.part[3 * slot] = P.withDot ? s0 : 0.0;
		P.part[3 * slot + 1] = s1;
		P.fpart[2 * slot] = a;
		P.fpart[2 * slot + 1] = b;
	}
}

// single-block deterministic reduction of nb partial triples
__device__ inline void reduce_parts(const double* __restrict__ part, int nb, int which, double* out, double* sh) {
	double a = 0;
	for (int b = threadIdx.x; b < nb; b += blockDim.x) a += part[3 * b + which];
	a = wave_sum(a);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
	__syncthreads();
	if (threadIdx.x == 0) {
		double s = 0;
		for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += sh[q];
		*out = s;
	}
	__syncthreads();
}

__device__ __forceinline__ void body_finalize(const FinP& P, long long, long long) {
	CgState* st = P.st;
	const double* __restrict__ part = P.part;
	const float* __restrict__ fpart = P.fpart;
	const int nb = P.nb, phase = P.phase, pcNone = P.pcNone;
	__shared__ double sh[16];
	__shared__ double r0, r1;
	__shared__ float smn[16], smx[16];
	if (phase != FIN_INIT && st->done) return;
	reduce_parts(part, nb, 0, &r0, sh);
	if (phase == FIN_INIT || phase == FIN_RES) reduce_parts(part, nb, 1, &r1, sh);
	float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
	if (phase == FIN_RES) {
		for (int b = threadIdx.x; b < nb; b += blockDim.x) {
			mn = fminf(mn, fpart[2 * b]);
			mx = fmaxf(mx, fpart[2 * b + 1]);
		}
		mn = wave_min(mn);
		mx = wave_max(mx);
		if ((threadIdx.x & 63) == 0) {
			smn[threadIdx.x >> 6] = mn;
			smx[threadIdx.x >> 6] = mx;
		}
		__syncthreads();
	}
	if (threadIdx.x != 0) return;
	if (phase == FIN_INIT) {
		st->sigma = (float)r0;
		st->bnorm2 = r1;
		st->iter = 0;
		st->done = st->maxIter <= 0 ? 1 : 0;
		st->converged = 0;
		st->diverged = 0;
		st->first = 1;
		st->beta = 0.f;
		st->resNorm = 1e20f;
		return;
	}
	if (phase == FIN_ALPHA) {
		const float dp = (float)r0;
		float alpha = 0.f;
		if (fabsf(dp) > 0.f) alpha = st->sigma / dp;
		st->alpha = alpha;
		st->first = 0;
		return;
	}
	if (phase == FIN_RES) {
		for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
			mn = fminf(mn, smn[q]);
			mx = fmaxf(mx, smx[q]);
		}
		st->iter += 1;
		st->rr = r1;
		float resNorm;
		bool conv;
		if (st->relTol > 0) {
			resNorm = (float)sqrt(r1 / (st->bnorm2 > 0 ? st->bnorm2 : 1.0));
			conv = sqrt(r1) <= st->relTol * sqrt(st->bnorm2);
		} else {
			if (st->useL2) resNorm = (float)r1;
			else resNorm = fmaxf(fabsf(mn), fabsf(mx));  // max(fabs(amin), fabs(amax))
			conv = resNorm < st->acc;
		}
		st->resNorm = resNorm;
		if (st->iter - 1 < 64) st->hist[st->iter - 1] = resNorm;
		if (conv) {
			st->sigma = resNorm;
			st->converged = 1;
			st->done = 1;
			return;
		}
		if (pcNone) {
			const float sigmaNew = (float)r0;
			st->beta = sigmaNew / st->sigma;
			st->sigma = sigmaNew;
			if (!(resNorm < 1e35f)) {
				st->diverged = 1;
				st->done = 1;
				return;
			}
			if (st->iter >= st->maxIter) st->done = 1;
		}
		return;
	}
	// FIN_BETA / FIN_BETA0 (preconditioned): sigmaNew = dot(z, r)
	const float sigmaNew = (float)r0;
	st->beta = sigmaNew / st->sigma;
	st->sigma = sigmaNew;
	if (phase == FIN_BETA0) st->first = 1;
	if (!(st->resNorm < 1e35f)) {
		st->diverged = 1;
		st->done = 1;
		return;
	}
	if (st->iter >= st->maxIter) st->done = 1;
}

// Three consecutive floats at a 4-byte aligned address, read by one load.
struct __attribute__((packed, aligned(4))) FTriple3 {
	float a, b, c;
};

// true residual ||b - A x||_2^2 with FP64 accumulation of the row sum;
// kLightKZ planes per thread, all loads issued first (linear neighbours
// outside [0, n) are replaced by the cell itself and never used).  The
// x-neighbours come from one 12-byte load at clamp(c - 1, 0, n - 3) (TRI:
// grids of at least 3 cells); the
// z-neighbours of inner planes are the thread's own adjacent planes, so only
// the first plane's lower and the last plane's upper neighbour are loaded.
// (a __device__ body: one launch per problem, or one for a batch; the
// partial goes to the logical tile's slot in both)
template <bool TRI, bool SC = false>
__device__ __forceinline__ void body_true_residual(const TrueResP& P, long long blk, long long slot) {
	const Dims& d = P.d;
	const uint8_t* __restrict__ code = P.code;
	const float* __restrict__ b = P.b;
	const float* __restrict__ x = P.x;
	double* __restrict__ part = P.part;
	const float scale = P.scale;
	__shared__ double sh[8];
	const CellPos p = tile_cell_of(P.t, (unsigned)blk);
	double rr = 0;
	if (p.i < d.nx && p.j < d.ny) {
		constexpr int KZ = kLightKZ;
		const long long X = 1, Y = d.nx, Z = d.sz, n = d.n;
		uint8_t cc[KZ], cxm[KZ], cym[KZ], czm[KZ];
		float bc[KZ], xc[KZ], xxm[KZ], xxp[KZ], xym[KZ], xyp[KZ], xzm[KZ], xzp[KZ];
		long long cl[KZ];
#pragma unroll
		for (int q = 0; q < KZ; q++) {
			const int k = min(p.k + q, d.nz - 1);
			const long long c = d.idx(p.i, p.j, k);
			cl[q] = c;
			const long long mx = c - X >= 0 ? c - X : c;
			const long long my = c - Y >= 0 ? c - Y : c, py = c + Y < n ? c + Y : c;
			if (TRI) {
				const long long s0 = min(max(c - 1, 0ll), n - 3);
				const FTriple3 tr = *reinterpret_cast<const FTriple3*>(x + s0);
				const long long o = c - s0;  // slot of c: 0 (c = 0), 1, or 2 (c = n - 1)
				xxm[q] = o == 2 ? tr.b : tr.a;
				xc[q] = o == 0 ? tr.a : (o == 1 ? tr.b : tr.c);
				xxp[q] = o == 0 ? tr.b : tr.c;
			} else {
				xxm[q] = x[mx];
				xc[q] = x[c];
				xxp[q] = x[c + X < n ? c + X : c];
			}
			cc[q] = code[c];
			cxm[q] = code[mx];
			cym[q] = code[my];
			bc[q] = b[c];
			xym[q] = x[my];
			xyp[q] = x[py];
		}
		{
			const long long mz = cl[0] - Z >= 0 ? cl[0] - Z : cl[0];
			const long long pz = cl[KZ - 1] + Z < n ? cl[KZ - 1] + Z : cl[KZ - 1];
			czm[0] = code[mz];
			xzm[0] = x[mz];
			xzp[KZ - 1] = x[pz];
		}
#pragma unroll
		for (int q = 1; q < KZ; q++) {
			czm[q] = cc[q - 1];
			xzm[q] = xc[q - 1];
			xzp[q - 1] = xc[q];
		}
#pragma unroll
		for (int q = 0; q < KZ; q++) {
			if (p.k + q >= d.nz) break;
			const long long c = d.idx(p.i, p.j, p.k + q);
			double ax;
			if (!(cc[q] & 64)) ax = xc[q];
			else if (SC) {
				const double o = -(double)scale;
				ax = (double)xc[q] * (double)sA0(cc[q], scale);
				if (c - X >= 0 && (cxm[q] & 8)) ax += o * xxm[q];
				if (c + X < n && (cc[q] & 8)) ax += o * xxp[q];
				if (c - Y >= 0 && (cym[q] & 16)) ax += o * xym[q];
				if (c + Y < n && (cc[q] & 16)) ax += o * xyp[q];
				if (d.is3d) {
					if (c - Z >= 0 && (czm[q] & 32)) ax += o * xzm[q];
					if (c + Z < n && (cc[q] & 32)) ax += o * xzp[q];
				}
			} else {
				ax = (double)xc[q] * (double)(cc[q] & 7);
				if (c - X >= 0 && (cxm[q] & 8)) ax -= xxm[q];
				if (c + X < n && (cc[q] & 8)) ax -= xxp[q];
				if (c - Y >= 0 && (cym[q] & 16)) ax -= xym[q];
				if (c + Y < n && (cc[q] & 16)) ax -= xyp[q];
				if (d.is3d) {
					if (c - Z >= 0 && (czm[q] & 32)) ax -= xzm[q];
					if (c + Z < n && (cc[q] & 32)) ax -= xzp[q];
				}
			}
			const double r = (double)bc[q] - ax;
			rr += r * r;
		}
	}
	const double s = block_sum(rr, sh);
	if (threadIdx.x == 0) part[3 * slot] = s;
}

// Residual replacement: r = b - A x with FP64 accumulation of the row sum,
// rounded once to float, in place of the recursively updated residual.  Emits
// the same partials as k_cg_update (sum fl(r*r) when z = r, sum r^2, min/max).
__global__ void __launch_bounds__(kBlock) k_cg_replace_residual(Dims d, Tiles t, const CgState* __restrict__ st,
                                                                const uint8_t* __restrict__ code,
                                                                const float* __restrict__ b,
                                                                const float* __restrict__ x, float* __restrict__ r,
                                                                int withDot, double* __restrict__ part,
                                                                float* __restrict__ fpart) {
	__shared__ double sh[8];
	__shared__ float shf[8];
	if (st->done) return;
	CellPos p = tile_cell(t);
	float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
	double zr = 0, rr = 0;
	if (p.i < d.nx && p.j < d.ny) {
		const long long c = d.idx(p.i, p.j, p.k);
		const uint8_t cc = code[c];
		double ax;
		if (!(cc & 64)) ax = x[c];
		else {
			const long long X = 1, Y = d.nx, Z = d.sz, n = d.n;
			ax = (double)x[c] * (double)(cc & 7);
			if (c - X >= 0 && (code[c - X] & 8)) ax -= x[c - X];
			if (c + X < n && (cc & 8)) ax -= x[c + X];
			if (c - Y >= 0 && (code[c - Y] & 16)) ax -= x[c - Y];
			if (c + Y < n && (cc & 16)) ax -= x[c + Y];
			if (d.is3d) {
				if (c - Z >= 0 && (code[c - Z] & 32)) ax -= x[c - Z];
				if (c + Z < n && (cc & 32)) ax -= x[c + Z];
			}
		}
		const float rn = (float)((double)b[c] - ax);
		r[c] = rn;
		mn = rn;
		mx = rn;
		zr = (double)(rn * rn);
		rr = (double)rn * (double)rn;
	}
	const double s0 = block_sum(zr, sh);
	const double s1 = block_sum(rr, sh);
	mn = wave_min(mn);
	mx = wave_max(mx);
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	if (lane == 0) {
		shf[wv] = mn;
		shf[4 + wv] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		float a = shf[0], bb = shf[4];
		for (int k = 1; k < 4; k++) {
			a = fminf(a, shf[k]);
			bb = fmaxf(bb, shf[4 + k]);
		}
		part[3 * blockIdx.x] = withDot ? s0 : 0.0;
		part[3 * blockIdx.x + 1] = s1;
		fpart[2 * blockIdx.x] = a;
		fpart[2 * blockIdx.x + 1] = bb;
	}
}

// ---------------------------------------------------------------------------
// launches of the CG loop kernels (single problem and work lists)
// ---------------------------------------------------------------------------
static std::mutex g_tune_mu;
static std::map<std::array<int, 3>, ZmTune> g_tune;
static std::set<std::array<int, 3>> g_tuning;  // shapes being tuned right now
ZmTune zm_tuned(const Dims& d) {
	std::lock_guard<std::mutex> g(g_tune_mu);
	auto it = g_tune.find({d.nx, d.ny, d.nz});
	return it == g_tune.end() ? ZmTune{} : it->second;
}
void zm_set_tuned(const Dims& d, ZmTune t) {
	std::lock_guard<std::mutex> g(g_tune_mu);
	g_tune[{d.nx, d.ny, d.nz}] = t;
	g_tuning.erase({d.nx, d.ny, d.nz});
}
bool zm_claim_tuning(const Dims& d) {
	std::lock_guard<std::mutex> g(g_tune_mu);
	const std::array<int, 3> key{d.nx, d.ny, d.nz};
	auto it = g_tune.find(key);
	if (it != g_tune.end() && it->second.mv > 0) return false;
	return g_tuning.insert(key).second;
}

long long matvec_tiles(const Dims& d, int* kchunk, bool tuned, int force) {
	// 64 x kMvTY column tiles; 3D: chunks of planes halved while fewer than
	// about 8192 wavefronts cover the grid, or the autotuned chunk
	const long long cols = (long long)((d.nx + 63) / 64) * ((d.ny + kMvTY - 1) / kMvTY);
	if (!d.is3d) {
		*kchunk = 1;
		return cols;
	}
	int kc = zm_chunk(d.nz);
	while (kc > 8 && cols * ((d.nz + kc - 1) / kc) < 8192 / kMvTY) kc = zm_chunk(kc / 2);
	if (tuned && zm_tuned(d).mv > 0) kc = zm_tuned(d).mv;
	if (force > 0) kc = force;
	*kchunk = kc;
	return cols * ((d.nz + kc - 1) / kc);
}
long long update_tiles(const Dims& d, long long* per) {
	// at most 2048 tiles; *per = tile count (the stride in 256-thread units)
	long long nt = std::min<long long>(2048, ((d.n >> 2) + kBlock - 1) / kBlock);
	if (nt < 1) nt = 1;
	*per = nt;
	return nt;
}
void run_matvec_one(hipStream_t s, unsigned nb, const MatvecP& p) {
	if (p.scaled)
		hipLaunchKernelGGL((k_one<MatvecP, body_matvec<kMvTY, true>, 64 * kMvTY>), dim3(nb), dim3(64 * kMvTY), 0, s,
		                   p);
	else
		hipLaunchKernelGGL((k_one<MatvecP, body_matvec<kMvTY>, 64 * kMvTY>), dim3(nb), dim3(64 * kMvTY), 0, s, p);
}
void run_matvec_many(hipStream_t s, unsigned nb, WorkList items, const MatvecP* p) {
	hipLaunchKernelGGL((k_many<MatvecP, body_matvec<kMvTY>, 64 * kMvTY>), dim3(nb), dim3(64 * kMvTY), 0, s, items, p);
}
void run_update_one(hipStream_t s, unsigned nb, const UpdateP& p) {
	hipLaunchKernelGGL((k_one<UpdateP, body_update, kBlock>), dim3(nb), dim3(kBlock), 0, s, p);
}
void run_update_many(hipStream_t s, unsigned nb, WorkList items, const UpdateP* p) {
	hipLaunchKernelGGL((k_many<UpdateP, body_update, kBlock>), dim3(nb), dim3(kBlock), 0, s, items, p);
}
void run_finalize_one(hipStream_t s, const FinP& p) {
	hipLaunchKernelGGL((k_one<FinP, body_finalize, 1024>), dim3(1), dim3(1024), 0, s, p);
}
void run_finalize_many(hipStream_t s, unsigned nprob, WorkList items, const FinP* p) {
	hipLaunchKernelGGL((k_many<FinP, body_finalize, 1024>), dim3(nprob), dim3(1024), 0, s, items, p);
}
void run_dot_one(hipStream_t s, unsigned nb, const DotP& p) {
	hipLaunchKernelGGL((k_one<DotP, body_dot, kBlock>), dim3(nb), dim3(kBlock), 0, s, p);
}
void run_dot_many(hipStream_t s, unsigned nb, WorkList items, const DotP* p) {
	hipLaunchKernelGGL((k_many<DotP, body_dot, kBlock>), dim3(nb), dim3(kBlock), 0, s, items, p);
}
static void run_true_residual_many(hipStream_t s, unsigned nb, WorkList items, const TrueResP* p) {
	launch_many<TrueResP, body_true_residual<true>, kBlock>(s, nb, items, p);
}
static void run_reduce_true_rr_many(hipStream_t s, unsigned nb, WorkList items, const TrueRedP* p) {
	launch_many<TrueRedP, body_reduce_true_rr, 1024>(s, nb, items, p);
}
static void run_correct_velocity_many(hipStream_t s, unsigned nb, WorkList items, const CorrVelP* p) {
	launch_many<CorrVelP, body_correct_velocity, kBlock>(s, nb, items, p);
}

// 64-bit hash of a matrix code array: the wrapping sum over cells of a
// mixed (index, code) word (splitmix64 finaliser), so it does not depend on
// the summation order.  Decides whether the hierarchy standing in for PcMIC
// still matches the matrix.
__global__ void k_code_hash(const uint8_t* __restrict__ code, long long n, unsigned long long* out) {
	unsigned long long h = 0;
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
		unsigned long long x = (unsigned long long)i * 0x9E3779B97F4A7C15ull + ((unsigned long long)code[i] + 1ull) * 0xC2B2AE3D27D4EB4Full;
		x ^= x >> 30;
		x *= 0xBF58476D1CE4E5B9ull;
		x ^= x >> 27;
		x *= 0x94D049BB133111EBull;
		x ^= x >> 31;
		h += x;
	}
	for (int off = 32; off > 0; off >>= 1) h += __shfl_down(h, off, 64);
	if ((threadIdx.x & 63) == 0) atomicAdd(out, h);
}

// ---------------------------------------------------------------------------
// Second-order obstacle and free-surface boundaries: MakeRhs with fractions,
// obstacle velocity, surface tension and per-cell correction
// (pressure.cpp:32-84), the fractions Laplacian with the ghost-fluid diagonal
// as four float arrays (conjugategrad.h:154-187, pressure.cpp:133-149), CG
// on that matrix, and the ghost-fluid velocity correction
// (pressure.cpp:151-219).  One thread per cell with a grid-stride loop; the
// arithmetic keeps the reference's float / double rounding, so RHS, matrix
// and corrections are bit-exact.
// ---------------------------------------------------------------------------
struct BndDev {
	const float* phi;   // fluid level set (ghost fluid), or null
	const float* pcc;   // per-cell divergence correction, or null
	const float* fx;    // fractions, SoA components, or null
	const float* fy;
	const float* fz;
	const float* ox;    // obstacle velocity (used with fractions), or null
	const float* oy;
	const float* oz;
	const float* curv;  // curvature (used with phi), or null
	float gfClamp, surfTens;
};

// thetaHelper (pressure.cpp:114-120)
__device__ inline float gf_theta(float inside, float outside) {
	const float denom = inside - outside;
	if ((double)denom > -1e-04) return 0.5f;
	float q = inside / denom;
	if (q > 1.f) q = 1.f;
	if (0.f > q) q = 0.f;
	return q;
}
// ghostFluidHelper (pressure.cpp:122-127): gfClamp itself below the clamp,
// else 1 - 1/alpha in double, stored as float
__device__ inline float gf_factor(float pin, float pout, float gfClamp) {
	const float alpha = gf_theta(pin, pout);
	if (alpha < gfClamp) return gfClamp;
	return (float)(1. - (1. / (double)alpha));
}
__device__ inline float gf_at(const BndDev& B, long long c, long long off) {
	return gf_factor(B.phi[c], B.phi[c + off], B.gfClamp);
}
// surfTensHelper (pressure.cpp:129-131)
__device__ inline float st_at(const BndDev& B, long long c, long long off) {
	return B.surfTens * (B.curv[c + off] - gf_at(B, c, off) * B.curv[c]);
}

__device__ inline void cell_of(const Dims& d, long long c, int& i, int& j, int& k) {
	const long long row = c / d.nx;
	i = (int)(c - row * d.nx);
	if (d.is3d) {
		j = (int)(row % d.ny);
		k = (int)(row / d.ny);
	} else {
		j = (int)row;
		k = 0;
	}
}

__global__ void __launch_bounds__(256) k_make_rhs_bnd(Dims d, const uint8_t* __restrict__ f8,
                                                      const float* __restrict__ u, const float* __restrict__ v,
                                                      const float* __restrict__ w, BndDev B,
                                                      float* __restrict__ rhs, double* __restrict__ part) {
	__shared__ double sh[8];
	double sum = 0, cnt = 0;
	const long long X = 1, Y = d.nx, Z = d.sz;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		int i, j, k;
		cell_of(d, c, i, j, k);
		float set = 0.f;
		if (interior_c(d, i, j, k) && (f8[c] & kFluid)) {
			if (!B.fx) {
				set = u[c] - u[c + X] + v[c] - v[c + Y];
				if (d.is3d) set += w[c] - w[c + Z];
			} else {
				set = B.fx[c] * u[c] - B.fx[c + X] * u[c + X] + B.fy[c] * v[c] - B.fy[c + Y] * v[c + Y];
				if (d.is3d) set += B.fz[c] * w[c] - B.fz[c + Z] * w[c + Z];
				if (B.ox) {
					set += (1 - B.fx[c]) * B.ox[c] - (1 - B.fx[c + X]) * B.ox[c + X] + (1 - B.fy[c]) * B.oy[c] -
					       (1 - B.fy[c + Y]) * B.oy[c + Y];
					if (d.is3d) set += (1 - B.fz[c]) * B.oz[c] - (1 - B.fz[c + Z]) * B.oz[c + Z];
				}
			}
			if (B.phi && B.curv) {
				if (f8[c - X] & kEmpty) set += st_at(B, c, -X);
				if (f8[c + X] & kEmpty) set += st_at(B, c, +X);
				if (f8[c - Y] & kEmpty) set += st_at(B, c, -Y);
				if (f8[c + Y] & kEmpty) set += st_at(B, c, +Y);
				if (d.is3d) {
					if (f8[c - Z] & kEmpty) set += st_at(B, c, -Z);
					if (f8[c + Z] & kEmpty) set += st_at(B, c, +Z);
				}
			}
			if (B.pcc) set += B.pcc[c];
			sum += (double)set;
			cnt += 1;
		}
		rhs[c] = set;
	}
	const double s = block_sum(sum, sh);
	const double nn = block_sum(cnt, sh);
	if (threadIdx.x == 0) {
		part[2 * blockIdx.x] = s;
		part[2 * blockIdx.x + 1] = nn;
	}
}

// MakeLaplaceMatrix (plain or with fractions) + ApplyGhostFluidDiagonal
__global__ void __launch_bounds__(256) k_make_laplace_bnd(Dims d, const uint8_t* __restrict__ f8, BndDev B,
                                                          float* __restrict__ A0, float* __restrict__ Ai,
                                                          float* __restrict__ Aj, float* __restrict__ Ak) {
	const long long X = 1, Y = d.nx, Z = d.sz;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		int i, j, k;
		cell_of(d, c, i, j, k);
		float a0 = 0.f, ai = 0.f, aj = 0.f, ak = 0.f;
		if (interior_c(d, i, j, k) && (f8[c] & kFluid)) {
			if (!B.fx) {
				if (!(f8[c - X] & kObstacle)) a0 += 1.f;
				if (!(f8[c + X] & kObstacle)) a0 += 1.f;
				if (!(f8[c - Y] & kObstacle)) a0 += 1.f;
				if (!(f8[c + Y] & kObstacle)) a0 += 1.f;
				if (d.is3d && !(f8[c - Z] & kObstacle)) a0 += 1.f;
				if (d.is3d && !(f8[c + Z] & kObstacle)) a0 += 1.f;
				if (f8[c + X] & kFluid) ai = -1.f;
				if (f8[c + Y] & kFluid) aj = -1.f;
				if (d.is3d && (f8[c + Z] & kFluid)) ak = -1.f;
			} else {
				a0 += B.fx[c];
				a0 += B.fx[c + X];
				a0 += B.fy[c];
				a0 += B.fy[c + Y];
				if (d.is3d) a0 += B.fz[c];
				if (d.is3d) a0 += B.fz[c + Z];
				if (f8[c + X] & kFluid) ai = -B.fx[c + X];
				if (f8[c + Y] & kFluid) aj = -B.fy[c + Y];
				if (d.is3d && (f8[c + Z] & kFluid)) ak = -B.fz[c + Z];
			}
			if (B.phi) {
				if (f8[c - X] & kEmpty) a0 -= gf_at(B, c, -X);
				if (f8[c + X] & kEmpty) a0 -= gf_at(B, c, +X);
				if (f8[c - Y] & kEmpty) a0 -= gf_at(B, c, -Y);
				if (f8[c + Y] & kEmpty) a0 -= gf_at(B, c, +Y);
				if (d.is3d) {
					if (f8[c - Z] & kEmpty) a0 -= gf_at(B, c, -Z);
					if (f8[c + Z] & kEmpty) a0 -= gf_at(B, c, +Z);
				}
			}
		}
		A0[c] = a0;
		Ai[c] = ai;
		Aj[c] = aj;
		Ak[c] = ak;
	}
}

// fixPressure (pressure.cpp:226-245) with value 0 on the float matrix, at
// the cell k_fix_select chose (which already set rhs there)
__global__ void k_fix_general(Dims d, const CgState* __restrict__ st, float* __restrict__ rhs,
                              float* __restrict__ A0, float* __restrict__ Ai, float* __restrict__ Aj,
                              float* __restrict__ Ak) {
	if (threadIdx.x != 0) return;
	const long long f = st->fix;
	if (f < 0) return;
	const long long X = 1, Y = d.nx, Z = d.sz;
	rhs[f + X] -= Ai[f] * 0.f;
	rhs[f + Y] -= Aj[f] * 0.f;
	rhs[f - X] -= Ai[f - X] * 0.f;
	rhs[f - Y] -= Aj[f - Y] * 0.f;
	if (d.is3d) {
		rhs[f + Z] -= Ak[f] * 0.f;
		rhs[f - Z] -= Ak[f - Z] * 0.f;
	}
	rhs[f] = 0.f;
	A0[f] = 1.f;
	Ai[f] = Aj[f] = Ak[f] = 0.f;
	Ai[f - X] = 0.f;
	Aj[f - Y] = 0.f;
	if (d.is3d) Ak[f - Z] = 0.f;
}

struct GenMat {
	const float *A0, *Ai, *Aj, *Ak;
};
// ApplyMatrix (conjugategrad.h:118-151) with src values from a functor;
// out-of-range neighbours contribute 0
template <typename SRC>
__device__ inline float gen_row(const Dims& d, const GenMat& M, long long c, SRC src) {
	const long long X = 1, Y = d.nx, Z = d.sz, n = d.n;
	auto S = [&](long long o) { return (c + o >= 0 && c + o < n) ? src(c + o) : 0.f; };
	auto Cf = [&](const float* A, long long o) { return (c + o >= 0 && c + o < n) ? A[c + o] : 0.f; };
	if (d.is3d)
		return src(c) * M.A0[c] + S(-X) * Cf(M.Ai, -X) + S(X) * M.Ai[c] + S(-Y) * Cf(M.Aj, -Y) + S(Y) * M.Aj[c] +
		       S(-Z) * Cf(M.Ak, -Z) + S(Z) * M.Ak[c];
	return src(c) * M.A0[c] + S(-X) * Cf(M.Ai, -X) + S(X) * M.Ai[c] + S(-Y) * Cf(M.Aj, -Y) + S(Y) * M.Aj[c];
}
// the same row sum in double (true residual, residual replacement)
__device__ inline double gen_row_d(const Dims& d, const GenMat& M, long long c, const float* __restrict__ x) {
	const long long X = 1, Y = d.nx, Z = d.sz, n = d.n;
	auto T = [&](const float* A, long long ao, long long xo) {
		return (c + xo >= 0 && c + xo < n && c + ao >= 0 && c + ao < n) ? (double)x[c + xo] * (double)A[c + ao] : 0.0;
	};
	double ax = (double)x[c] * (double)M.A0[c];
	ax += T(M.Ai, -X, -X) + T(M.Ai, 0, X) + T(M.Aj, -Y, -Y) + T(M.Aj, 0, Y);
	if (d.is3d) ax += T(M.Ak, -Z, -Z) + T(M.Ak, 0, Z);
	return ax;
}

// GridCg::iterate's search direction and matrix product on the float
// matrix: p = z (first iteration) or z + beta p_old, q = A p (identity rows
// outside the fluid), partial sums of fl(q p) per block
__global__ void __launch_bounds__(256) k_matvec_gen(Dims d, const CgState* __restrict__ st,
                                                    const uint8_t* __restrict__ f8, GenMat M,
                                                    const float* __restrict__ z, const float* __restrict__ pold,
                                                    float* __restrict__ pnew, float* __restrict__ q,
                                                    double* __restrict__ part) {
	__shared__ double sh[8];
	if (st->done) return;
	const int first = st->first;
	const float beta = st->beta;
	auto pn = [&](long long x) { return first ? z[x] : z[x] + beta * pold[x]; };
	double acc = 0;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		const float pc = pn(c);
		const float qc = (f8[c] & kFluid) ? gen_row(d, M, c, pn) : pc;
		pnew[c] = pc;
		q[c] = qc;
		acc += (double)(qc * pc);
	}
	const double s = block_sum(acc, sh);
	if (threadIdx.x == 0) part[3 * blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_true_residual_gen(Dims d, const uint8_t* __restrict__ f8, GenMat M,
                                                           const float* __restrict__ b, const float* __restrict__ x,
                                                           double* __restrict__ part) {
	__shared__ double sh[8];
	double rr = 0;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		const double ax = (f8[c] & kFluid) ? gen_row_d(d, M, c, x) : (double)x[c];
		const double r = (double)b[c] - ax;
		rr += r * r;
	}
	const double s = block_sum(rr, sh);
	if (threadIdx.x == 0) part[3 * blockIdx.x] = s;
}

// residual replacement on the float matrix (same partials as
// k_cg_replace_residual)
__global__ void __launch_bounds__(256) k_replace_residual_gen(Dims d, const CgState* __restrict__ st,
                                                              const uint8_t* __restrict__ f8, GenMat M,
                                                              const float* __restrict__ b,
                                                              const float* __restrict__ x, float* __restrict__ r,
                                                              int withDot, double* __restrict__ part,
                                                              float* __restrict__ fpart) {
	__shared__ double sh[8];
	__shared__ float shf[8];
	if (st->done) return;
	float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
	double zr = 0, rr = 0;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		const double ax = (f8[c] & kFluid) ? gen_row_d(d, M, c, x) : (double)x[c];
		const float rn = (float)((double)b[c] - ax);
		r[c] = rn;
		mn = fminf(mn, rn);
		mx = fmaxf(mx, rn);
		zr += (double)(rn * rn);
		rr += (double)rn * (double)rn;
	}
	const double s0 = block_sum(zr, sh);
	const double s1 = block_sum(rr, sh);
	mn = wave_min(mn);
	mx = wave_max(mx);
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	if (lane == 0) {
		shf[wv] = mn;
		shf[4 + wv] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		float a = shf[0], bb = shf[4];
		for (int k = 1; k < 4; k++) {
			a = fminf(a, shf[k]);
			bb = fmaxf(bb, shf[4 + k]);
		}
		part[3 * blockIdx.x] = withDot ? s0 : 0.0;
		part[3 * blockIdx.x + 1] = s1;
		fpart[2 * blockIdx.x] = a;
		fpart[2 * blockIdx.x + 1] = bb;
	}
}

// knCorrectVelocityGhostFluid (pressure.cpp:151-190), bnd=1, after
// knCorrectVelocity
__global__ void __launch_bounds__(256) k_correct_gf(Dims d, const uint8_t* __restrict__ f8,
                                                    const float* __restrict__ p, BndDev B, float* __restrict__ u,
                                                    float* __restrict__ v, float* __restrict__ w) {
	const long long X = 1, Y = d.nx, Z = d.sz;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		int i, j, k;
		cell_of(d, c, i, j, k);
		if (!interior_c(d, i, j, k)) continue;
		const uint8_t f = f8[c];
		float vu = u[c], vv = v[c], vw = w[c];
		if (f & kFluid) {
			if (f8[c - X] & kEmpty) vu += p[c] * gf_at(B, c, -X);
			if (f8[c - Y] & kEmpty) vv += p[c] * gf_at(B, c, -Y);
			if (d.is3d && (f8[c - Z] & kEmpty)) vw += p[c] * gf_at(B, c, -Z);
		} else if ((f & kEmpty) && !(f & kOutflow)) {
			if (f8[c - X] & kFluid) vu -= p[c - X] * gf_at(B, c - X, +X);
			else vu = 0.f;
			if (f8[c - Y] & kFluid) vv -= p[c - Y] * gf_at(B, c - Y, +Y);
			else vv = 0.f;
			if (d.is3d) {
				if (f8[c - Z] & kFluid) vw -= p[c - Z] * gf_at(B, c - Z, +Z);
				else vw = 0.f;
			}
		}
		if (B.curv) {
			if (f & kFluid) {
				if (f8[c - X] & kEmpty) vu += st_at(B, c, -X);
				if (f8[c - Y] & kEmpty) vv += st_at(B, c, -Y);
				if (d.is3d && (f8[c - Z] & kEmpty)) vw += st_at(B, c, -Z);
			} else if ((f & kEmpty) && !(f & kOutflow)) {
				vu -= (f8[c - X] & kFluid) ? st_at(B, c - X, +X) : 0.f;
				vv -= (f8[c - Y] & kFluid) ? st_at(B, c - Y, +Y) : 0.f;
				if (d.is3d) vw -= (f8[c - Z] & kFluid) ? st_at(B, c - Z, +Z) : 0.f;
			}
		}
		u[c] = vu;
		v[c] = vv;
		if (d.is3d) w[c] = vw;
	}
}

// ghostFluidWasClamped (pressure.cpp:194-199)
__device__ inline bool gf_clamped(const BndDev& B, long long c, long long off) {
	return gf_theta(B.phi[c], B.phi[c + off]) < B.gfClamp;
}
// knReplaceClampedGhostFluidVels (pressure.cpp:201-219), bnd=1: empty cells
// take a fluid neighbour's face where that face's fraction was clamped (the
// faces read belong to fluid cells, which this pass does not write)
__global__ void __launch_bounds__(256) k_replace_clamped(Dims d, const uint8_t* __restrict__ f8, BndDev B,
                                                         float* __restrict__ u, float* __restrict__ v,
                                                         float* __restrict__ w) {
	const long long X = 1, Y = d.nx, Z = d.sz;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		int i, j, k;
		cell_of(d, c, i, j, k);
		if (!interior_c(d, i, j, k) || !(f8[c] & kEmpty)) continue;
		float vu = u[c], vv = v[c], vw = w[c];
		if ((f8[c - X] & kFluid) && gf_clamped(B, c - X, +X)) vu = u[c - X];
		if ((f8[c - Y] & kFluid) && gf_clamped(B, c - Y, +Y)) vv = v[c - Y];
		if (d.is3d && (f8[c - Z] & kFluid) && gf_clamped(B, c - Z, +Z)) vw = w[c - Z];
		if ((f8[c + X] & kFluid) && gf_clamped(B, c + X, -X)) vu = u[c + X];
		if ((f8[c + Y] & kFluid) && gf_clamped(B, c + Y, -Y)) vv = v[c + Y];
		if (d.is3d && (f8[c + Z] & kFluid) && gf_clamped(B, c + Z, -Z)) vw = w[c + Z];
		u[c] = vu;
		v[c] = vv;
		if (d.is3d) w[c] = vw;
	}
}

// Multigrid on the float matrix: the V-cycle is built from the flags' plain
// Laplacian L; it is applied as S V S with S = sqrt(diag L / diag A) on fluid
// rows (1 elsewhere), a symmetric congruence that stays SPD and gives the
// preconditioned rows the system's diagonal (large ghost-fluid diagonals
// otherwise leave those rows under-weighted)
__global__ void __launch_bounds__(256) k_diag_scale(Dims d, const uint8_t* __restrict__ code,
                                                    const uint8_t* __restrict__ f8, const float* __restrict__ A0,
                                                    float* __restrict__ S) {
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		float sc = 1.f;
		if (f8[c] & kFluid) {
			const float dl = (float)(code[c] & 7), da = A0[c];
			if (dl > 0.f && da > 0.f) sc = sqrtf(dl / da);
		}
		S[c] = sc;
	}
}
__global__ void __launch_bounds__(256) k_scale_vec(long long n, const CgState* __restrict__ st,
                                                   const float* __restrict__ S, const float* __restrict__ in,
                                                   float* __restrict__ out) {
	if (st->done) return;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < n; c += (long long)gridDim.x * 256) out[c] = S[c] * in[c];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static unsigned flat_count(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

static uint64_t code_hash(Scene* s, const uint8_t* code) {
	if (!s->dhash) MFX_HIP(hipMalloc((void**)&s->dhash, sizeof(uint64_t)));
	if (!s->hhash) MFX_HIP(hipHostMalloc((void**)&s->hhash, sizeof(uint64_t), hipHostMallocDefault));
	MFX_HIP(hipMemsetAsync(s->dhash, 0, sizeof(uint64_t), s->stream));
	const unsigned nb = std::min<unsigned>(2048u, flat_blocks(s->d.n));
	hipLaunchKernelGGL(k_code_hash, dim3(nb), dim3(256), 0, s->stream, code, s->d.n,
	                   (unsigned long long*)s->dhash);
	MFX_HIP(hipMemcpyAsync(s->hhash, s->dhash, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
	sync(s);
	return *s->hhash;
}

// The hierarchy that stands in for PcMIC: kept between solves while the
// matrix stays the same (same flag grid generation and pinned cell, else
// the same code hash), rebuilt otherwise.  The reference's mIC is built
// from each solve's matrix (conjugategrad.cpp:66-97), so a stale hierarchy
// would change the preconditioner's meaning.
static MgHierarchy* mic_hierarchy(Scene* s, const Grid* flags, const uint8_t* code, bool fixing, int precision) {
	// the pinned cell is a function of the flags, so (flags, generation,
	// fixing on/off) determines the matrix code
	const long long fix = fixing ? 1 : 0;
	const bool sameKey = s->mgMic && s->mgMic->precision() == precision && s->micFlags == flags &&
	                     s->micGen == flags->gen && s->micFix == fix;
	if (!sameKey) {
		const uint64_t h = code_hash(s, code);
		if (!(s->mgMic && s->mgMic->precision() == precision && s->micHash == h)) {
			s->mgMic.reset(new MgHierarchy());
			s->mgMic->build(s, code, precision);
		}
		s->micHash = h;
		s->micFlags = flags;
		s->micGen = flags->gen;
		s->micFix = fix;
	}
	return s->mgMic.get();
}

// RHS (MakeRhs, pressure.cpp:32-84) into rhs; with enforceCompatibility the
// mean correction is reduced and added on the device.  part: FP64 partials
// (2 per block of the k_make_rhs launch, plus one float slot after them).
// code (not with the boundary inputs): the matrix code with no pinned cell
// is written by the same pass.
static void compute_rhs(Scene* s, Grid* flags, Grid* vel, float* rhs, int enforceCompatibility,
                        double* part = nullptr, const BndDev* bnd = nullptr, uint8_t* code = nullptr) {
	const Tiles t = tiles_for(s->d, kLightKZ);
	const bool extra = bnd && (bnd->fx || bnd->pcc || (bnd->phi && bnd->curv));
	const unsigned nb = extra ? flat_blocks(s->d.n) : (unsigned)t.count();
	if (!part) {
		s->ensurePartials(3 * (size_t)nb + 16, 2 * (size_t)nb + 16);
		part = s->dpart;
	}
	MFX_CHECK(!(extra && code), "the matrix code is built with the RHS only without boundary inputs");
	if (extra)
		hipLaunchKernelGGL(k_make_rhs_bnd, dim3(nb), dim3(256), 0, s->stream, s->d, (const uint8_t*)flags->f8,
		                   (const float*)vel->comp(0), (const float*)vel->comp(1), (const float*)vel->comp(2), *bnd,
		                   rhs, part);
	else if (code)
		hipLaunchKernelGGL(k_make_rhs<true>, dim3(nb), dim3(kBlock), 0, s->stream, s->d, t, flags->f8,
		                   (const float*)vel->comp(0), (const float*)vel->comp(1), (const float*)vel->comp(2), rhs,
		                   part, code);
	else
		hipLaunchKernelGGL(k_make_rhs<false>, dim3(nb), dim3(kBlock), 0, s->stream, s->d, t, flags->f8,
		                   (const float*)vel->comp(0), (const float*)vel->comp(1), (const float*)vel->comp(2), rhs,
		                   part, nullptr);
	if (enforceCompatibility) {
		float* corr = (float*)(part + 2 * (size_t)nb);
		hipLaunchKernelGGL(k_compat_corr, dim3(1), dim3(1024), 0, s->stream, (const double*)part, (int)nb, corr);
		hipLaunchKernelGGL(k_add_const_dev, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, rhs, s->d.n,
		                   (const float*)corr);
	}
}

// One pressure-system solve in three phases.  The batched solver runs the
// same set-up and finish per problem (each problem owns its scene, hence its
// partial-sum buffers and device CgState) and drives the iterations of all
// problems through work lists.
struct CgRun {
	Scene* s = nullptr;
	bool codeReady = false;  // the matrix code was written with the RHS (no pinned cell)
	Grid* flags = nullptr;
	float* rhs = nullptr;
	float* x = nullptr;
	mfx_pressure_opts o{};
	std::chrono::steady_clock::time_point t0;
	Tiles t{};
	unsigned nbT = 0, nbF = 0, nbM = 0, nbU = 0;
	int npM = 0;  // partials of fl(q p) left by the matvec
	int kchunk = 1;
	long long per = 0;
	bool fixing = false;  // zero-pressure fixing requested (the cell is chosen on the device)
	// pooled: buffers, matrix code and partial sums come from the scene (and
	// the z-marching chunk may be autotuned); otherwise the caller set code,
	// r, pa, pb, q, z, dpart and fpart before cg_setup
	bool pooled = true;
	double* dpart = nullptr;
	float* fpart = nullptr;
	uint8_t* code = nullptr;
	int pc = 0, maxIter = 0;
	bool pcNone = true;
	MgHierarchy* mg = nullptr;
	std::unique_ptr<MgHierarchy> mgDyn;
	float *r = nullptr, *pa = nullptr, *pb = nullptr, *q = nullptr, *z = nullptr;
	CgState* st = nullptr;  // device
	int launched = 0;
	int scaled = 0;     // scaled Laplace matrix (cgSolveDiffusion, cgSolveWE)
	float scale = 0.f;
	// float matrix of the second-order boundaries (fractions / ghost fluid),
	// pooled grids; null on the matrix-code path
	float *gA0 = nullptr, *gAi = nullptr, *gAj = nullptr, *gAk = nullptr;
	float *gS = nullptr, *gT = nullptr;  // V-cycle scaling and its scaled input (preconditioned only)
	unsigned nbG = 0;  // blocks of the float-matrix kernels
	GenMat gen() const { return GenMat{gA0, gAi, gAj, gAk}; }
	MatvecP mvp(bool odd) const {
		MatvecP p{s->d, kchunk, st, code, z, odd ? pb : pa, odd ? pa : pb, q, dpart, scaled, scale};
		return p;
	}
	UpdateP upd(bool odd) const {
		UpdateP p{s->d, per, st, x, r, odd ? pa : pb, q, pcNone ? 1 : 0, dpart, fpart};
		return p;
	}
	FinP fin(int nb, int phase) const {
		FinP p{st, dpart, fpart, nb, phase, pcNone ? 1 : 0};
		return p;
	}
	TrueResP trueResP() const { return TrueResP{s->d, tiles_for(s->d, kLightKZ), code, rhs, x, dpart, scale}; }
	TrueRedP trueRedP() const { return TrueRedP{st, dpart, (int)tiles_for(s->d, kLightKZ).count()}; }
};

// matrix, preconditioner, buffers, CgState and GridCg::doInit
// (conjugategrad.cpp:210-235, pressure.cpp:349-433).  With batchInit the
// initial V-cycle and sigma = dot(z, r) are left to the caller's batched
// launches (FIN_BETA0 afterwards).
// Autotuner of the z-marching plane chunk.  The first time a 3D grid shape
// is solved alone, the CG matvec and the level-0 multigrid down-sweep are
// timed (HIP events, three launches after one warm-up) for chunks of 4..64
// planes on the problem's own matrix code and pooled scratch grids under a
// scratch CgState; the fastest chunk of each is kept for the shape (the
// up-sweep shares the down-sweep's).  Single solves are latency-bound, so
// the winners are short chunks for small grids; batched solves pick their
// own chunk policy per batch layout (batch_policy below).  The dot
// partition of the z-marching kernels makes the result independent of the
// chunk, so single and batched solves still agree bit for bit.
static void autotune_zmarch(Scene* s, const uint8_t* code) {
	const Dims& d = s->d;
	if (!d.is3d) return;
	if (!zm_claim_tuning(d)) return;
	float* a = s->getReal();
	float* b = s->getReal();
	float* c = s->getReal();
	float* e = s->getReal();
	CgState* st = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	ZmTune best;
	auto cleanup = [&]() {
		if (st) (void)hipFree(st);
		if (e0) (void)hipEventDestroy(e0);
		if (e1) (void)hipEventDestroy(e1);
		s->putReal(a);
		s->putReal(b);
		s->putReal(c);
		s->putReal(e);
	};
	try {
		MFX_HIP(hipMalloc(&st, sizeof(CgState)));
		MFX_HIP(hipEventCreate(&e0));
		MFX_HIP(hipEventCreate(&e1));
		hipLaunchKernelGGL(k_cg_state_init, dim3(1), dim3(64), 0, s->stream, st, 1e-3f, 1, 0, 0.0);
		for (float* g : {a, b}) MFX_HIP(hipMemsetAsync(g, 0, sizeof(float) * (size_t)d.n, s->stream));
		const int maxc = zm_chunk(d.nz);
		auto timed = [&](auto launch) -> float {
			float ms = 0.f;
			launch();
			MFX_HIP(hipEventRecord(e0, s->stream));
			for (int r = 0; r < 3; r++) launch();
			MFX_HIP(hipEventRecord(e1, s->stream));
			MFX_HIP(hipEventSynchronize(e1));
			MFX_HIP(hipEventElapsedTime(&ms, e0, e1));
			return ms;
		};
		float tmv = 1e30f, tl0 = 1e30f;
		for (int kc = kDotGroup; kc <= maxc; kc *= 2) {
			// trial chunks stay local to this thread: other solves of the
			// shape use the heuristic chunk until the winners are published
			int k1 = 0, k2 = 0;
			const unsigned nbm = (unsigned)matvec_tiles(d, &k1, false, kc);
			const unsigned nbl = (unsigned)l0_tiles(d, &k2, false, kc);
			const MatvecP mp{d, k1, st, code, a, b, c, e, s->dpart};
			const L0DownP lp{d, k2, st, code, a, c, e};
			const float tm = timed([&]() { run_matvec_one(s->stream, nbm, mp); });
			if (tm < tmv) {
				tmv = tm;
				best.mv = kc;
			}
			const float tl = timed([&]() { run_l0_down_one(s->stream, true, nbl, lp); });
			if (tl < tl0) {
				tl0 = tl;
				best.l0 = kc;
			}
		}
		zm_set_tuned(d, best);
	} catch (...) {
		zm_set_tuned(d, ZmTune{});
		cleanup();
		throw;
	}
	cleanup();
}

static int precondition(CgRun& R);

static void cg_setup(CgRun& R, Scene* s, Grid* flags, float* rhs, float* x, const mfx_pressure_opts& o,
                     bool batchInit = false, const BndDev* bnd = nullptr) {
	R.t0 = std::chrono::steady_clock::now();
	R.s = s;
	R.flags = flags;
	R.rhs = rhs;
	R.x = x;
	R.o = o;
	const Dims& d = s->d;
	R.t = tiles_for(d);
	R.nbT = (unsigned)R.t.count();
	R.nbF = flat_count(d.n);
	R.nbU = (unsigned)update_tiles(d, &R.per);
	const unsigned nbMax = std::max(R.nbT, R.nbF);
	if (R.pooled) {
		s->ensurePartials(3 * (size_t)nbMax + 16, 2 * (size_t)nbMax + 16);
		R.dpart = s->dpart;
		R.fpart = s->fpart;
	}

	// preconditioner choice (pressure.cpp:404-433, conjugategrad.cpp:302-336)
	int pc = o.preconditioner;
	int mx = d.nx > d.ny ? d.nx : d.ny;
	if (d.nz > mx) mx = d.nz;
	if (pc == MFX_PC_NONE || pc == MFX_PC_MIC) {
		R.maxIter = (int)(o.cgMaxIterFac * (float)mx) * (d.is3d ? 1 : 4);
		// mIC only exists in 3D in the reference; the GPU runs the parallel
		// multigrid preconditioner in its place
		pc = (pc == MFX_PC_MIC && d.is3d) ? MFX_PC_MGSTATIC : MFX_PC_NONE;
	} else {
		R.maxIter = 100;
	}
	if (o.maxIterOverride > 0) R.maxIter = o.maxIterOverride;
	R.pc = pc;
	R.pcNone = pc == MFX_PC_NONE;

	R.st = (CgState*)s->dscal;
	hipLaunchKernelGGL(k_cg_state_init, dim3(1), dim3(64), 0, s->stream, R.st, o.cgAccuracy, R.maxIter, o.useL2Norm,
	                   o.relTol);
	// zero-pressure fixing (pressure.cpp:349-388), chosen and applied on the
	// device (no copy of the flags to the host)
	R.fixing = o.zeroPressureFixing || o.cgAccuracy < 1e-07f;
	if (R.fixing) {
		const Tiles t1 = tiles_for(d);
		hipLaunchKernelGGL(k_fix_scan, dim3((unsigned)t1.count()), dim3(kBlock), 0, s->stream, d, t1, flags->f8,
		                   R.st);
		hipLaunchKernelGGL(k_fix_select, dim3(1), dim3(64), 0, s->stream, d, flags->f8, R.st, rhs);
	}
	// the float matrix of fractions / ghost fluid (the matrix code below
	// then serves the multigrid preconditioner only)
	if (bnd && (bnd->fx || bnd->phi)) {
		R.nbG = flat_blocks(d.n);
		R.gA0 = s->getReal();
		R.gAi = s->getReal();
		R.gAj = s->getReal();
		R.gAk = s->getReal();
		hipLaunchKernelGGL(k_make_laplace_bnd, dim3(R.nbG), dim3(256), 0, s->stream, d, (const uint8_t*)flags->f8,
		                   *bnd, R.gA0, R.gAi, R.gAj, R.gAk);
		if (R.fixing)
			hipLaunchKernelGGL(k_fix_general, dim3(1), dim3(64), 0, s->stream, d, (const CgState*)R.st, rhs, R.gA0,
			                   R.gAi, R.gAj, R.gAk);
	}
	// matrix
	if (R.pooled) {
		if (!s->code) MFX_HIP(hipMalloc((void**)&s->code, (size_t)d.n));
		R.code = s->code;
	}
	if (R.codeReady) {
		MFX_CHECK(R.pooled && !R.fixing && R.code == s->code, "matrix code built with the RHS for another solve");
	} else {
		const Tiles tc = tiles_for(d, kLightKZ);
		hipLaunchKernelGGL(k_build_code, dim3((unsigned)tc.count()), dim3(kBlock), 0, s->stream, d, tc, flags->f8,
		                   R.code, (const CgState*)R.st);
	}
	if (R.pooled) autotune_zmarch(s, R.code);
	R.nbM = (unsigned)matvec_tiles(d, &R.kchunk, true);
	R.npM = (int)zm_dot_parts(d, kMvTY);

	if (!R.pcNone) {
		if (o.preconditioner == MFX_PC_MGDYNAMIC) {
			s->mgStatic.reset();
			R.mgDyn.reset(new MgHierarchy());
			R.mgDyn->build(s, R.code, o.precision);
			R.mg = R.mgDyn.get();
		} else if (o.preconditioner == MFX_PC_MIC) {
			R.mg = mic_hierarchy(s, flags, R.code, R.fixing, o.precision);
		} else {
			if (!s->mgStatic || s->mgStatic->precision() != o.precision) {
				s->mgStatic.reset(new MgHierarchy());
				s->mgStatic->build(s, R.code, o.precision);
			}
			R.mg = s->mgStatic.get();
		}
	}

	if (R.pooled) {
		R.r = s->getReal();
		R.pa = s->getReal();
		R.pb = s->getReal();
		R.q = s->getReal();
		R.z = R.pcNone ? R.r : s->getReal();
	}

	// doInit: x = 0, r = b, bnorm; preconditioned: z = M r, sigma = dot(z, r)
	hipLaunchKernelGGL(k_cg_init, dim3(R.nbU), dim3(kBlock), 0, s->stream, d, R.per, (const float*)rhs, x, R.r,
	                   R.pcNone ? 1 : 0, R.dpart);
	run_finalize_one(s->stream, FinP{R.st, R.dpart, R.fpart, (int)R.nbU, FIN_INIT, 0});
	if (R.gA0 && !R.pcNone) {
		R.gS = s->getReal();
		R.gT = s->getReal();
		hipLaunchKernelGGL(k_diag_scale, dim3(R.nbG), dim3(256), 0, s->stream, d, (const uint8_t*)R.code,
		                   (const uint8_t*)flags->f8, (const float*)R.gA0, R.gS);
	}
	if (!R.pcNone && !batchInit) {
		const int np = precondition(R);
		run_finalize_one(s->stream, FinP{R.st, R.dpart, R.fpart, np, FIN_BETA0, 0});
	}
	R.launched = 0;
}

// r = b - A x with FP64 row sums in place of the recursive residual (plus
// the update's partials); returns the number of partial-sum slots written
static int enqueue_replace_residual(CgRun& R) {
	Scene* s = R.s;
	if (R.gA0) {
		hipLaunchKernelGGL(k_replace_residual_gen, dim3(R.nbG), dim3(256), 0, s->stream, s->d, (const CgState*)R.st,
		                   (const uint8_t*)R.flags->f8, R.gen(), (const float*)R.rhs, (const float*)R.x, R.r,
		                   R.pcNone ? 1 : 0, R.dpart, R.fpart);
		return (int)R.nbG;
	}
	hipLaunchKernelGGL(k_cg_replace_residual, dim3(R.nbT), dim3(kBlock), 0, s->stream, s->d, R.t, (const CgState*)R.st,
	                   R.code, (const float*)R.rhs, (const float*)R.x, R.r, R.pcNone ? 1 : 0, R.dpart, R.fpart);
	return (int)R.nbT;
}

// z = M r (V-cycle, scaled on the float matrix) and the partials of
// fl(z r); returns their count
static int precondition(CgRun& R) {
	Scene* s = R.s;
	int np;
	if (R.gS) {
		const unsigned nb = flat_blocks(s->d.n);
		hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s->stream, s->d.n, (const CgState*)R.st,
		                   (const float*)R.gS, (const float*)R.r, R.gT);
		np = R.mg->vcycle(s, R.code, R.gT, R.z, R.st);
		hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s->stream, s->d.n, (const CgState*)R.st,
		                   (const float*)R.gS, (const float*)R.z, R.z);
	} else {
		np = R.mg->vcycle(s, R.code, R.r, R.z, R.st);
	}
	if (np == 0) {
		run_dot_one(s->stream, R.nbF, DotP{s->d, R.st, R.z, R.r, R.dpart});
		np = (int)R.nbF;
	}
	return np;
}

// GridCg::iterate (conjugategrad.cpp:238-299), queued on the scene stream
static void cg_iteration(CgRun& R) {
	Scene* s = R.s;
	const bool odd = (R.launched & 1) != 0;
	if (R.gA0) {
		const MatvecP m = R.mvp(odd);
		hipLaunchKernelGGL(k_matvec_gen, dim3(R.nbG), dim3(256), 0, s->stream, s->d, (const CgState*)R.st,
		                   (const uint8_t*)R.flags->f8, R.gen(), m.z, m.pold, m.pnew, m.q, R.dpart);
		run_finalize_one(s->stream, R.fin((int)R.nbG, FIN_ALPHA));
	} else {
		{
			KtScope kt("cg_matvec", s->stream);
			run_matvec_one(s->stream, R.nbM, R.mvp(odd));
		}
		run_finalize_one(s->stream, R.fin(R.npM, FIN_ALPHA));
	}
	{
		KtScope kt("cg_update", s->stream);
		run_update_one(s->stream, R.nbU, R.upd(odd));
	}
	int nbRes = (int)R.nbU;
	if (R.o.trueResidualEvery > 0 && (R.launched + 1) % R.o.trueResidualEvery == 0) {
		// FP64 residual replacement; its partials replace the update's
		nbRes = enqueue_replace_residual(R);
	}
	run_finalize_one(s->stream, R.fin(nbRes, FIN_RES));
	if (!R.pcNone) {
		const int np = precondition(R);
		run_finalize_one(s->stream, R.fin(np, FIN_BETA));
	}
	R.launched++;
}

// FP64 true residual reduced on the device, CgState to the scene's pinned
// mirror, buffers back to the pool (queued; nothing waits)
static void cg_true_residual_enqueue(CgRun& R) {
	Scene* s = R.s;
	if (R.gA0) {
		hipLaunchKernelGGL(k_true_residual_gen, dim3(R.nbG), dim3(256), 0, s->stream, s->d, (const uint8_t*)R.flags->f8,
		                   R.gen(), (const float*)R.rhs, (const float*)R.x, R.dpart);
		hipLaunchKernelGGL(k_reduce_true_rr, dim3(1), dim3(1024), 0, s->stream, R.st, (const double*)R.dpart,
		                   (int)R.nbG);
		MFX_HIP(hipMemcpyAsync(s->hscal, R.st, sizeof(CgState), hipMemcpyDeviceToHost, s->stream));
		return;
	}
	const Tiles tr = tiles_for(s->d, kLightKZ);
	const TrueResP tp = R.trueResP();
	if (R.scaled) launch_one<TrueResP, body_true_residual<false, true>, kBlock>(s->stream, tr.count(), tp);
	else if (s->d.n >= 3) launch_one<TrueResP, body_true_residual<true>, kBlock>(s->stream, tr.count(), tp);
	else launch_one<TrueResP, body_true_residual<false>, kBlock>(s->stream, tr.count(), tp);
	hipLaunchKernelGGL(k_reduce_true_rr, dim3(1), dim3(1024), 0, s->stream, R.st, (const double*)R.dpart,
	                   (int)tr.count());
	MFX_HIP(hipMemcpyAsync(s->hscal, R.st, sizeof(CgState), hipMemcpyDeviceToHost, s->stream));
}

static void cg_release(CgRun& R) {
	Scene* s = R.s;
	if (R.pooled) {
		s->putReal(R.r);
		s->putReal(R.pa);
		s->putReal(R.pb);
		s->putReal(R.q);
		if (!R.pcNone) s->putReal(R.z);
	}
	for (float** a : {&R.gA0, &R.gAi, &R.gAj, &R.gAk, &R.gS, &R.gT})
		if (*a) {
			s->putReal(*a);
			*a = nullptr;
		}
}

// CG restarted from the current x: r = b - A x (FP64 row sums), z = M r,
// p = z on the next iteration; the iteration count and ||b|| carry on
static void cg_restart(CgRun& R) {
	Scene* s = R.s;
	hipLaunchKernelGGL(k_cg_restart, dim3(1), dim3(64), 0, s->stream, R.st);
	int np = enqueue_replace_residual(R);
	if (!R.pcNone) np = precondition(R);
	run_finalize_one(s->stream, R.fin(np, FIN_BETA0));
}

// With relTol the stopping rule reads the recursively updated FP32 residual,
// which can drift from b - A x.  When it has met relTol but the FP64 true
// residual has not, CG restarts from the current x (at most kMaxRestarts
// times, within maxIter).  The scene's host mirror holds the state after the
// true residual.
constexpr int kMaxRestarts = 3;
static bool cg_needs_restart(const CgRun& R) {
	const CgState* hs = (const CgState*)R.s->hscal;
	if (!(R.o.relTol > 0) || R.scaled || !hs->converged || hs->diverged || hs->iter >= R.maxIter) return false;
	const double bb = hs->bnorm2;
	const double rel = bb > 0 ? std::sqrt(hs->trueRR / bb) : std::sqrt(hs->trueRR);
	return rel > R.o.relTol;
}
static void cg_iterate(CgRun& R, double target);
// Iterations, then the FP64 true residual and the state copy, until the
// copied state says done or maxIter iterations are queued.  The residual is
// queued as soon as the iterations in flight are predicted to converge; if
// they did not, it is superseded by the one after further iterations.
static void cg_iterate_finish(CgRun& R, double target) {
	while (true) {
		cg_iterate(R, target);
		cg_true_residual_enqueue(R);
		sync(R.s);
		if (((const CgState*)R.s->hscal)->done || R.launched >= R.maxIter) return;
	}
}

static void cg_verify(CgRun& R, double target) {
	for (int k = 0; k < kMaxRestarts && cg_needs_restart(R); k++) {
		// iterations queued past convergence did nothing; the budget follows
		// the device count (p = z after the restart, so the buffer parity is free)
		R.launched = ((const CgState*)R.s->hscal)->iter;
		cg_restart(R);
		cg_iterate_finish(R, target);
	}
}

// after the scene stream has drained: statistics and the divergence error
static void cg_finish_collect(CgRun& R, mfx_solve_stats* stats) {
	const CgState* hs = (const CgState*)R.s->hscal;
	const double bb = hs->bnorm2, rr = hs->trueRR;
	if (stats) {
		stats->iterations = hs->iter;
		stats->resNorm = hs->resNorm;
		stats->relResidual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
		stats->fixPidx = (int)hs->fix;
		stats->maxIter = R.maxIter;
		stats->preconditionerUsed = R.pc;
		stats->converged = hs->converged;
		stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - R.t0).count();
	}
	if (hs->diverged) {
		if (!R.pcNone)
			fprintf(stderr,
			        "GridCg::iterate: Warning - this diverging solve can be caused by the 'static' mode of the MG "
			        "preconditioner. If the static mode is active, try switching to dynamic.\n");
		throw Error("GridCg::iterate: The CG solver diverged, residual norm > 1e30, stopping.");
	}
}

// pressure system solve: rhs (modified by fixPressure), pressure out.
// Iterations are queued in batches whose size is predicted from the
// residual history; kernels stop on st->done.
static void cg_drive(CgRun& R, double target, mfx_solve_stats* stats);

static void solve_system(Scene* s, Grid* flags, float* rhs, float* x, const mfx_pressure_opts& o,
                         mfx_solve_stats* stats, const CgRun* bufs = nullptr, const BndDev* bnd = nullptr,
                         bool codeReady = false) {
	CgRun R;
	R.codeReady = codeReady;
	if (bufs) {
		// caller-provided buffers (allocation-free solve)
		R.pooled = false;
		R.code = bufs->code;
		R.r = bufs->r;
		R.pa = bufs->pa;
		R.pb = bufs->pb;
		R.q = bufs->q;
		R.z = bufs->r;
		R.dpart = bufs->dpart;
		R.fpart = bufs->fpart;
	}
	cg_setup(R, s, flags, rhs, x, o, false, bnd);
	cg_drive(R, o.relTol > 0 ? o.relTol : (double)o.cgAccuracy, stats);
}

// Queue iterations in batches sized from the residual history until the
// stopping rule holds or maxIter is reached, then finish (statistics,
// divergence error).  target: the stopping threshold the prediction aims at.
static void cg_drive(CgRun& R, double target, mfx_solve_stats* stats) {
	cg_iterate_finish(R, target);
	cg_verify(R, target);
	cg_release(R);
	cg_finish_collect(R, stats);
}

// Queues iterations in batches sized from the residual history.  The host
// reads the state of the next-to-last iteration of a batch while the GPU
// runs the last one, so the queue does not drain at these checks; it
// returns when the iterations queued are predicted to reach the target (or
// to be done, or maxIter is reached), and the caller queues the finish.
static void cg_iterate(CgRun& R, double target) {
	Scene* s = R.s;
	int batch = 4;
	CgState* hs = (CgState*)s->hscal;
	if (!s->evState) MFX_HIP(hipEventCreateWithFlags(&s->evState, hipEventDisableTiming));
	while (true) {
		const int nIt = std::min(batch, R.maxIter - R.launched);
		for (int it = 0; it < nIt; it++) {
			cg_iteration(R);
			if (it == nIt - 2) {
				MFX_HIP(hipMemcpyAsync(hs, R.st, sizeof(CgState), hipMemcpyDeviceToHost, s->stream));
				MFX_HIP(hipEventRecord(s->evState, s->stream));
			}
		}
		if (nIt < 2 || R.launched >= R.maxIter) return;
		MFX_HIP(hipEventSynchronize(s->evState));
		if (hs->done) return;  // the iteration in flight returns at once
		// iterations still needed after the one in flight (without a usable
		// rate: another batch of the same size)
		int more = batch;
		const int itNow = hs->iter;
		if (itNow >= 2 && itNow <= 64) {
			const double a = hs->hist[itNow - 2], b = hs->hist[itNow - 1];
			if (a > 0 && b > 0 && b < a && target > 0) {
				const double rate = b / a;
				double need = std::log(target / b) / std::log(rate);
				if (need < 1) need = 1;
				more = (int)std::ceil(need) - 1;
			}
		}
		if (more <= 0) return;
		batch = std::min(more, 64);
	}
}


// GridCg without preconditioner over a scaled Laplace matrix (the other
// GridCg clients): kind 0 builds cgSolveDiffusion's matrix code, kind 1 the
// pressure matrix of the flags (cgSolveWE); x is cleared by doInit and
// solved for rhs.  maxIter = (int)(cgMaxIterFac * max size) * (4 in 2D).
static void solve_scaled(Scene* s, Grid* flags, float* rhs, float* x, int kind, float scale, float cgAccuracy,
                         float cgMaxIterFac, mfx_solve_stats* stats) {
	CgRun R;
	R.t0 = std::chrono::steady_clock::now();
	R.s = s;
	R.flags = flags;
	R.rhs = rhs;
	R.x = x;
	mfx_pressure_opts_default(&R.o);
	R.o.cgAccuracy = cgAccuracy;
	R.o.preconditioner = MFX_PC_NONE;
	const Dims& d = s->d;
	R.t = tiles_for(d);
	R.nbT = (unsigned)R.t.count();
	R.nbF = flat_count(d.n);
	R.nbU = (unsigned)update_tiles(d, &R.per);
	const unsigned nbMax = std::max(R.nbT, R.nbF);
	s->ensurePartials(3 * (size_t)nbMax + 16, 2 * (size_t)nbMax + 16);
	R.dpart = s->dpart;
	R.fpart = s->fpart;
	int mx = d.nx > d.ny ? d.nx : d.ny;
	if (d.nz > mx) mx = d.nz;
	R.maxIter = (int)(cgMaxIterFac * (float)mx) * (d.is3d ? 1 : 4);
	R.pc = MFX_PC_NONE;
	R.pcNone = true;
	R.scaled = 1;
	R.scale = scale;
	R.st = (CgState*)s->dscal;
	hipLaunchKernelGGL(k_cg_state_init, dim3(1), dim3(64), 0, s->stream, R.st, cgAccuracy, R.maxIter, 0, 0.0);
	if (!s->code) MFX_HIP(hipMalloc((void**)&s->code, (size_t)d.n));
	R.code = s->code;
	const Tiles tc = tiles_for(d, kLightKZ);
	if (kind == 0)
		hipLaunchKernelGGL(k_build_code_diff, dim3((unsigned)tc.count()), dim3(kBlock), 0, s->stream, d, tc,
		                   flags->f8, R.code);
	else
		hipLaunchKernelGGL(k_build_code, dim3((unsigned)tc.count()), dim3(kBlock), 0, s->stream, d, tc, flags->f8,
		                   R.code, (const CgState*)R.st);
	R.nbM = (unsigned)matvec_tiles(d, &R.kchunk, true);
	R.npM = (int)zm_dot_parts(d, kMvTY);
	R.r = s->getReal();
	R.pa = s->getReal();
	R.pb = s->getReal();
	R.q = s->getReal();
	R.z = R.r;
	hipLaunchKernelGGL(k_cg_init, dim3(R.nbU), dim3(kBlock), 0, s->stream, d, R.per, (const float*)rhs, x, R.r, 1,
	                   R.dpart);
	run_finalize_one(s->stream, FinP{R.st, R.dpart, R.fpart, (int)R.nbU, FIN_INIT, 0});
	R.launched = 0;
	cg_drive(R, (double)cgAccuracy, stats);
}

static void correct_velocity(Scene* s, Grid* flags, Grid* vel, const float* pressure, const BndDev* bnd = nullptr) {
	const Tiles t = tiles_for(s->d, kLightKZ);
	launch_one<CorrVelP, body_correct_velocity, kBlock>(
	    s->stream, t.count(), CorrVelP{s->d, t, flags->f8, pressure, vel->comp(0), vel->comp(1), vel->comp(2)});
	if (bnd && bnd->phi) {
		// correctVelocity with phi (pressure.cpp:470-474)
		const unsigned nb = flat_blocks(s->d.n);
		hipLaunchKernelGGL(k_correct_gf, dim3(nb), dim3(256), 0, s->stream, s->d, (const uint8_t*)flags->f8, pressure,
		                   *bnd, vel->comp(0), vel->comp(1), vel->comp(2));
		hipLaunchKernelGGL(k_replace_clamped, dim3(nb), dim3(256), 0, s->stream, s->d, (const uint8_t*)flags->f8, *bnd,
		                   vel->comp(0), vel->comp(1), vel->comp(2));
	}
}

// ---------------------------------------------------------------------------
// Batched solve: many independent pressure systems (one per scene, any mix
// of sizes and 2D/3D).  Set-up and finish run per problem; every iteration
// is one launch per CG/MG phase over a work list that holds the tiles of all
// problems, so small problems share launches with large ones.  Iterations are
// queued a few ahead of pipelined convergence checks.  When a batch has the
// same launch layout as the previous batch led by the same scene, two
// iterations (both parities of the search-direction buffers) and the done-flag
// gather are captured once in a hipGraph and replayed from then on.  Each
// problem keeps its own device CgState, so converged problems drop out at
// once (their workgroups return on st->done).
// ---------------------------------------------------------------------------
namespace {

struct DevArena {
	std::vector<uint8_t> host;
	size_t add(const void* p, size_t bytes) {
		const size_t off = (host.size() + 255) & ~(size_t)255;
		host.resize(off + bytes);
		if (bytes) memcpy(host.data() + off, p, bytes);
		return off;
	}
};

// One batched launch: per-problem tile counts and parameters; on the device
// the tile starts (prefix sums) and the parameter array.
template <typename P>
struct Step {
	std::vector<int> tiles;
	std::vector<P> params;
	long long total = 0;
	size_t startOff = 0, paramOff = 0;
	WorkList dItems{nullptr, 0};
	const P* dParams = nullptr;
	explicit Step(size_t nprob = 0) : tiles(nprob, 0), params(nprob) {}
	void add_tiles(int prob, long long t) {
		tiles[(size_t)prob] += (int)t;
		total += t;
	}
	void place(DevArena& A) {
		std::vector<int> st(tiles.size() + 1, 0);
		for (size_t i = 0; i < tiles.size(); i++) st[i + 1] = st[i] + tiles[i];
		startOff = A.add(st.data(), sizeof(int) * st.size());
		paramOff = A.add(params.data(), sizeof(P) * params.size());
	}
	void bind(uint8_t* base) {
		dItems = WorkList{(const int*)(base + startOff), (int)tiles.size()};
		dParams = (const P*)(base + paramOff);
	}
	unsigned n() const { return (unsigned)total; }
};

__global__ void k_gather_done(int np, CgState* const* __restrict__ st, int* __restrict__ done) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < np) done[i] = st[i]->done;
}

struct BatchSolve {
	int P = 0;
	bool bf16 = false;
	// CG
	Step<MatvecP> mv[2];
	Step<UpdateP> up[2];
	Step<FinP> fa, fr, fb, fb0;
	// V-cycle
	Step<FillP> fill;
	Step<L0DownP> down[2];  // [is3d]
	std::vector<Step<RestrictP>> rs[2];
	std::vector<Step<SmoothP>> pre[2], post[2];
	std::vector<Step<ResidP>> res[2];
	Step<CoarsestP> crs;
	Step<L0UpP> l0up[2];
	Step<DotP> dot;
	// finish: true residual partials, their sums, the velocity correction
	Step<TrueResP> tres;
	Step<TrueRedP> tred;
	Step<CorrVelP> corr;
	bool anyMg = false;
	uint8_t* dmem = nullptr;
	CgState** dStates = nullptr;
	int* dDone = nullptr;

	void enqueue(hipStream_t s, int parity) {
		run_matvec_many(s, mv[parity].n(), mv[parity].dItems, mv[parity].dParams);
		run_finalize_many(s, fa.n(), fa.dItems, fa.dParams);
		run_update_many(s, up[parity].n(), up[parity].dItems, up[parity].dParams);
		run_finalize_many(s, fr.n(), fr.dItems, fr.dParams);
		if (!anyMg) return;
		enqueue_vcycle(s);
		run_finalize_many(s, fb.n(), fb.dItems, fb.dParams);
	}
	// z = M r for every preconditioned problem (+ dot(z, r) partials)
	void enqueue_vcycle(hipStream_t s) {
		run_fill_many(s, fill.n(), fill.dItems, fill.dParams);
		for (int d3 = 0; d3 < 2; d3++) run_l0_down_many(s, d3, down[d3].n(), down[d3].dItems, down[d3].dParams);
		const size_t nl = std::max(rs[0].size(), rs[1].size());
		for (size_t l = 0; l < nl; l++)
			for (int d3 = 0; d3 < 2; d3++) {
				if (l < rs[d3].size()) run_restrict_many(s, d3, rs[d3][l].n(), rs[d3][l].dItems, rs[d3][l].dParams);
				if (l < pre[d3].size()) {
					run_smooth_many(s, d3, bf16, false, pre[d3][l].n(), pre[d3][l].dItems, pre[d3][l].dParams);
					run_resid_many(s, d3, bf16, res[d3][l].n(), res[d3][l].dItems, res[d3][l].dParams);
				}
			}
		run_coarsest_many(s, crs.n(), crs.dItems, crs.dParams);
		for (size_t l = nl; l-- > 0;)
			for (int d3 = 0; d3 < 2; d3++)
				if (l < post[d3].size())
					run_smooth_many(s, d3, bf16, true, post[d3][l].n(), post[d3][l].dItems, post[d3][l].dParams);
		for (int d3 = 0; d3 < 2; d3++) run_l0_up_many(s, d3, l0up[d3].n(), l0up[d3].dItems, l0up[d3].dParams);
		run_dot_many(s, dot.n(), dot.dItems, dot.dParams);
	}
};

}  // namespace

// Online autotuner of the batches' plane chunks.  A batch layout (precision
// and the problems' shapes, in order) is solved with one chunk policy per
// solve - the throughput heuristic, the single-solve autotuner's per-shape
// winners, and the heuristic's chunks doubled and quadrupled (capped at the
// grid depth) - twice each, interleaved, timing the iteration phase on the
// host; from then on the layout uses the fastest.  The dot partition of the z-marching kernels
// makes every policy give the same bits.  A layout seen once (a serving tick)
// gets the heuristic, the first policy.
constexpr int kBatchPolicies = 4, kBatchTrials = 2, kBatchTuneMax = 256;
struct BatchTune {
	double best[kBatchPolicies] = {1e30, 1e30, 1e30, 1e30};
	int runs[kBatchPolicies] = {0, 0, 0, 0};
	int chosen = -1;
	unsigned long long used = 0;
};
static std::mutex g_btune_mu;
static std::map<std::vector<int>, BatchTune> g_btune;
static unsigned long long g_btune_clock = 0;

static int batch_policy(const std::vector<int>& key) {
	std::lock_guard<std::mutex> g(g_btune_mu);
	auto it = g_btune.find(key);
	if (it == g_btune.end()) {
		if ((int)g_btune.size() >= kBatchTuneMax) {
			// forget the least recently used layout
			auto old = g_btune.begin();
			for (auto j = g_btune.begin(); j != g_btune.end(); ++j)
				if (j->second.used < old->second.used) old = j;
			g_btune.erase(old);
		}
		it = g_btune.emplace(key, BatchTune{}).first;
	}
	BatchTune& t = it->second;
	t.used = ++g_btune_clock;
	if (t.chosen >= 0) return t.chosen;
	int total = 0;
	for (int p = 0; p < kBatchPolicies; p++) total += t.runs[p];
	return total % kBatchPolicies;
}
static void batch_policy_done(const std::vector<int>& key, int policy, double seconds) {
	std::lock_guard<std::mutex> g(g_btune_mu);
	auto it = g_btune.find(key);
	if (it == g_btune.end() || it->second.chosen >= 0) return;
	BatchTune& t = it->second;
	t.runs[policy] += 1;
	t.best[policy] = std::min(t.best[policy], seconds);
	for (int p = 0; p < kBatchPolicies; p++)
		if (t.runs[p] < kBatchTrials) return;
	int b = 0;
	for (int p = 1; p < kBatchPolicies; p++)
		if (t.best[p] < t.best[b]) b = p;
	t.chosen = b;
}
// plane chunks of a 3D problem under a batch policy: tuned (per-shape
// winners) and the forced chunk for the matvec and the level-0 kernels
static void batch_chunks(const Dims& d, int policy, bool* tuned, int* forceMv, int* forceL0) {
	*tuned = policy == 1;
	*forceMv = *forceL0 = 0;
	if (policy < 2 || !d.is3d) return;
	int kmv = 0, kl0 = 0;
	matvec_tiles(d, &kmv, false);
	l0_tiles(d, &kl0, false);
	const int cap = zm_chunk(d.nz), f = policy == 2 ? 2 : 4;
	*forceMv = std::min(f * kmv, cap);
	*forceL0 = std::min(f * kl0, cap);
}

// f(i) for i in [0, n) on up to eight host threads (the calling one
// included), each problem's calls on its own scene and stream; the first
// exception (by index) is rethrown after all have finished.  The per-problem
// set-up and finishing of a batch are host-bound (launches, an event wait
// per problem), so spreading them over threads shortens the batch.
template <typename F>
static void parallel_for(int n, int device, F&& f) {
	const int nt = std::min(n, 8);
	std::vector<std::exception_ptr> err((size_t)n);
	std::atomic<int> next{0};
	auto work = [&]() {
		for (;;) {
			const int i = next.fetch_add(1);
			if (i >= n) return;
			try {
				f(i);
			} catch (...) {
				err[(size_t)i] = std::current_exception();
			}
		}
	};
	std::vector<std::thread> th;
	for (int t = 1; t < nt; t++)
		th.emplace_back([&, device]() {
			(void)hipSetDevice(device);
			work();
		});
	work();
	for (auto& t : th) t.join();
	for (auto& e : err)
		if (e) std::rethrow_exception(e);
}

static void solve_batch(const std::vector<Grid*>& vel, const std::vector<Grid*>& pres,
                        const std::vector<Grid*>& flg, const mfx_pressure_opts& o, mfx_solve_stats* stats) {
	const int P = (int)vel.size();
	std::vector<float*> rhs((size_t)P, nullptr);
	std::vector<CgRun> R((size_t)P);
	auto release = [&]() {
		for (int i = 0; i < P; i++)
			if (rhs[(size_t)i]) vel[(size_t)i]->s->putReal(rhs[(size_t)i]);
	};
	const int device = vel[0]->s->device;
	try {
		parallel_for(P, device, [&](int i) {
			Scene* s = vel[(size_t)i]->s;
			rhs[(size_t)i] = s->getReal();
			// without a pinned cell the matrix code is written by the RHS pass
			const bool fuseCode = !(o.zeroPressureFixing || o.cgAccuracy < 1e-07f);
			if (fuseCode && !s->code) MFX_HIP(hipMalloc((void**)&s->code, (size_t)s->d.n));
			compute_rhs(s, flg[(size_t)i], vel[(size_t)i], rhs[(size_t)i], o.enforceCompatibility, nullptr, nullptr,
			            fuseCode ? s->code : nullptr);
			R[(size_t)i].codeReady = fuseCode;
			cg_setup(R[(size_t)i], s, flg[(size_t)i], rhs[(size_t)i], pres[(size_t)i]->real(), o, true);
			// marks the end of this problem's set-up for the batch stream
			if (!s->ev) MFX_HIP(hipEventCreateWithFlags(&s->ev, hipEventDisableTiming));
			MFX_HIP(hipEventRecord(s->ev, s->stream));
		});

		BatchSolve B;
		B.P = P;
		B.bf16 = o.precision == 1;
		for (int par = 0; par < 2; par++) {
			B.mv[par] = Step<MatvecP>((size_t)P);
			B.up[par] = Step<UpdateP>((size_t)P);
		}
		B.fa = Step<FinP>((size_t)P);
		B.fr = Step<FinP>((size_t)P);
		B.fb = Step<FinP>((size_t)P);
		B.fb0 = Step<FinP>((size_t)P);
		B.fill = Step<FillP>((size_t)P);
		B.crs = Step<CoarsestP>((size_t)P);
		B.dot = Step<DotP>((size_t)P);
		B.tres = Step<TrueResP>((size_t)P);
		B.tred = Step<TrueRedP>((size_t)P);
		B.corr = Step<CorrVelP>((size_t)P);
		for (int d3 = 0; d3 < 2; d3++) {
			B.down[d3] = Step<L0DownP>((size_t)P);
			B.l0up[d3] = Step<L0UpP>((size_t)P);
		}
		std::vector<int> tkey = {o.precision, P};
		for (int i = 0; i < P; i++) {
			const Dims& d = R[(size_t)i].s->d;
			tkey.push_back(d.nx);
			tkey.push_back(d.ny);
			tkey.push_back(d.nz);
		}
		const int policy = batch_policy(tkey);
		int maxIt = 0;
		for (int i = 0; i < P; i++) {
			CgRun& r = R[(size_t)i];
			const Dims& d = r.s->d;
			maxIt = std::max(maxIt, r.maxIter);
			bool tuned = false;
			int fMv = 0, fL0 = 0;
			batch_chunks(d, policy, &tuned, &fMv, &fL0);
			int kb = 1;
			const long long nbMv = matvec_tiles(d, &kb, tuned, fMv);
			for (int par = 0; par < 2; par++) {
				B.mv[par].params[(size_t)i] = r.mvp(par == 1);
				B.mv[par].params[(size_t)i].kchunk = kb;
				B.mv[par].add_tiles(i, nbMv);
				B.up[par].params[(size_t)i] = r.upd(par == 1);
				B.up[par].add_tiles(i, r.nbU);
			}
			MFX_CHECK(!r.scaled && !r.gA0 && d.n >= 3, "batched solve: unexpected system kind");
			const long long nLight = tiles_for(d, kLightKZ).count();
			B.tres.params[(size_t)i] = r.trueResP();
			B.tres.add_tiles(i, nLight);
			B.tred.params[(size_t)i] = r.trueRedP();
			B.tred.add_tiles(i, 1);
			B.corr.params[(size_t)i] = CorrVelP{d, tiles_for(d, kLightKZ), flg[(size_t)i]->f8, pres[(size_t)i]->real(),
			                                    vel[(size_t)i]->comp(0), vel[(size_t)i]->comp(1), vel[(size_t)i]->comp(2)};
			B.corr.add_tiles(i, nLight);
			B.fa.params[(size_t)i] = r.fin(r.npM, FIN_ALPHA);
			B.fa.add_tiles(i, 1);
			B.fr.params[(size_t)i] = r.fin((int)r.nbU, FIN_RES);
			B.fr.add_tiles(i, 1);
			if (r.pcNone) continue;
			B.anyMg = true;
			const MgHierarchy::Steps S = r.mg->steps(d, r.code, r.r, r.z, r.st, r.s->dpart, tuned, fL0);
			const int d3 = d.is3d ? 1 : 0;
			int np;
			if (S.nl == 1) {
				B.fill.params[(size_t)i] = S.fill;
				B.fill.add_tiles(i, (d.n + 255) / 256);
				B.crs.params[(size_t)i] = S.coarsest;
				B.crs.add_tiles(i, coarsest_tiles(S.coarsest.nc));
				B.dot.params[(size_t)i] = DotP{d, r.st, r.z, r.r, r.s->dpart};
				B.dot.add_tiles(i, dot_tiles(d));
				np = (int)dot_tiles(d);
			} else {
				B.down[d3].params[(size_t)i] = S.down;
				B.down[d3].add_tiles(i, S.l0Tiles);
				const int nlv = (int)S.lv.size();
				while ((int)B.rs[d3].size() < nlv) B.rs[d3].emplace_back((size_t)P);
				while ((int)B.pre[d3].size() < nlv - 1) {
					B.pre[d3].emplace_back((size_t)P);
					B.res[d3].emplace_back((size_t)P);
					B.post[d3].emplace_back((size_t)P);
				}
				for (int l = 0; l < nlv; l++) {
					B.rs[d3][(size_t)l].params[(size_t)i] = S.lv[(size_t)l].restrict_;
					B.rs[d3][(size_t)l].add_tiles(i, S.lv[(size_t)l].tiles);
					if (l == nlv - 1) break;
					B.pre[d3][(size_t)l].params[(size_t)i] = S.lv[(size_t)l].pre;
					B.pre[d3][(size_t)l].add_tiles(i, S.lv[(size_t)l].blkTiles);
					B.res[d3][(size_t)l].params[(size_t)i] = S.lv[(size_t)l].resid;
					B.res[d3][(size_t)l].add_tiles(i, S.lv[(size_t)l].tiles);
					B.post[d3][(size_t)l].params[(size_t)i] = S.lv[(size_t)l].post;
					B.post[d3][(size_t)l].add_tiles(i, S.lv[(size_t)l].blkTiles);
				}
				B.crs.params[(size_t)i] = S.coarsest;
				B.crs.add_tiles(i, coarsest_tiles(S.coarsest.nc));
				B.l0up[d3].params[(size_t)i] = S.up;
				B.l0up[d3].add_tiles(i, S.l0Tiles);
				np = (int)S.nPart;
			}
			B.fb.params[(size_t)i] = r.fin(np, FIN_BETA);
			B.fb.add_tiles(i, 1);
			B.fb0.params[(size_t)i] = r.fin(np, FIN_BETA0);
			B.fb0.add_tiles(i, 1);
		}
		// one device block for every list and parameter array
		DevArena A;
		auto placeAll = [&](auto&& f) {
			for (int par = 0; par < 2; par++) {
				f(B.mv[par]);
				f(B.up[par]);
			}
			f(B.fa);
			f(B.fr);
			f(B.fb);
			f(B.fb0);
			f(B.fill);
			f(B.crs);
			f(B.dot);
			f(B.tres);
			f(B.tred);
			f(B.corr);
			for (int d3 = 0; d3 < 2; d3++) {
				f(B.down[d3]);
				f(B.l0up[d3]);
				for (auto& x : B.rs[d3]) f(x);
				for (auto& x : B.pre[d3]) f(x);
				for (auto& x : B.res[d3]) f(x);
				for (auto& x : B.post[d3]) f(x);
			}
		};
		placeAll([&](auto& st) { st.place(A); });
		std::vector<CgState*> states((size_t)P);
		for (int i = 0; i < P; i++) states[(size_t)i] = R[(size_t)i].st;
		const size_t stOff = A.add(states.data(), sizeof(CgState*) * (size_t)P);
		std::vector<int> zeros((size_t)P, 0);
		const size_t doneOff = A.add(zeros.data(), sizeof(int) * (size_t)P);
		std::vector<CgState> gatherInit((size_t)P);
		const size_t gatherOff = A.add(gatherInit.data(), sizeof(CgState) * (size_t)P);
		Scene* s0 = vel[0]->s;
		hipStream_t q = s0->stream;
		if (s0->arenaCap < A.host.size()) {
			// grow (hipFree waits for the device; this happens once per size class)
			if (s0->arena) MFX_HIP(hipFree(s0->arena));
			s0->arena = nullptr;
			s0->arenaCap = 0;
			MFX_HIP(hipMalloc(&s0->arena, A.host.size() * 2));
			s0->arenaCap = A.host.size() * 2;
		}
		B.dmem = (uint8_t*)s0->arena;
		// staged through pinned memory, so the upload does not hold the host
		// (the previous batch led by this scene has finished: its solve synced)
		if (s0->hstageCap < A.host.size()) {
			if (s0->hstage) MFX_HIP(hipHostFree(s0->hstage));
			s0->hstage = nullptr;
			s0->hstageCap = 0;
			MFX_HIP(hipHostMalloc(&s0->hstage, A.host.size() * 2, hipHostMallocDefault));
			s0->hstageCap = A.host.size() * 2;
		}
		memcpy(s0->hstage, A.host.data(), A.host.size());
		MFX_HIP(hipMemcpyAsync(B.dmem, s0->hstage, A.host.size(), hipMemcpyHostToDevice, q));
		placeAll([&](auto& st) { st.bind(B.dmem); });
		B.dStates = (CgState**)(B.dmem + stOff);
		B.dDone = (int*)(B.dmem + doneOff);
		// the batch stream waits for every problem's set-up
		for (int i = 1; i < P; i++) MFX_HIP(hipStreamWaitEvent(q, vel[(size_t)i]->s->ev, 0));
		// initial V-cycle of every preconditioned problem, sigma = dot(z, r)
		if (B.anyMg) {
			B.enqueue_vcycle(q);
			run_finalize_many(q, B.fb0.n(), B.fb0.dItems, B.fb0.dParams);
		}

		// launch layout of one iteration; equal keys give identical launches
		std::vector<long long> key = {P, B.bf16 ? 1 : 0, B.anyMg ? 1 : 0, (long long)(uintptr_t)s0->arena};
		placeAll([&](auto& st) {
			key.push_back(st.total);
			key.push_back((long long)st.startOff);
			key.push_back((long long)st.paramOff);
		});
		key.push_back((long long)stOff);
		key.push_back((long long)doneOff);
		hipGraphExec_t gexec = nullptr;
		if (s0->bexec && s0->bkey == key) {
			gexec = s0->bexec;
		} else if (s0->blastKey == key) {
			// second batch in a row with this layout: capture it once
			hipGraph_t graph = nullptr;
			MFX_HIP(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
			B.enqueue(q, 0);
			B.enqueue(q, 1);
			hipLaunchKernelGGL(k_gather_done, dim3((P + 255) / 256), dim3(256), 0, q, P, (CgState* const*)B.dStates,
			                   B.dDone);
			MFX_HIP(hipStreamEndCapture(q, &graph));
			hipGraphExec_t e = nullptr;
			const hipError_t ie = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
			(void)hipGraphDestroy(graph);
			MFX_HIP(ie);
			if (s0->bexec) (void)hipGraphExecDestroy(s0->bexec);
			s0->bexec = e;
			s0->bkey = key;
			gexec = e;
		}
		s0->blastKey = key;
		// Iterations are queued directly on the batch stream, a few ahead of
		// the host: after every second iteration the done flags of all
		// problems are copied to a pinned slot and an event recorded; the host
		// waits only for the oldest outstanding check while later iterations
		// are already queued.  Iterations queued after every problem has
		// converged return at once (their kernels test the done flag).
		constexpr int kAhead = 4;  // iterations queued beyond the last completed check
		int* hdone = (int*)s0->hscal;  // pinned, 4 KiB: one slot of P flags per outstanding check
		const int nslot = std::max(1, (int)(4096 / (sizeof(int) * (size_t)P)));
		const bool pinned = nslot >= 2;
		std::vector<int> done((size_t)P, 0);
		struct Check {
			int its;
			int slot;
			hipEvent_t ev;
		};
		std::vector<hipEvent_t> evPool;
		std::deque<Check> pending;
		int its = 0, checked = 0, nextSlot = 0;
		auto take_event = [&]() {
			hipEvent_t e;
			if (!evPool.empty()) {
				e = evPool.back();
				evPool.pop_back();
			} else {
				MFX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
			}
			return e;
		};
		bool converged = false;
		const auto tIter = std::chrono::steady_clock::now();
		try {
			while (!converged) {
				while (its < maxIt && its - checked < kAhead && (pinned ? (int)pending.size() < nslot : pending.empty())) {
					if (gexec && maxIt - its >= 2) {
						// (its is even here: the graph starts with parity 0)
						MFX_HIP(hipGraphLaunch(gexec, q));
						its += 2;
					} else {
						B.enqueue(q, its & 1);
						its++;
					}
					if ((its & 1) == 0 || its == maxIt) {
						const int slot = pinned ? nextSlot : 0;
						nextSlot = (nextSlot + 1) % nslot;
						if (!gexec || (its & 1))
							hipLaunchKernelGGL(k_gather_done, dim3((P + 255) / 256), dim3(256), 0, q, P,
							                   (CgState* const*)B.dStates, B.dDone);
						MFX_HIP(hipMemcpyAsync(pinned ? hdone + (size_t)slot * P : done.data(), B.dDone,
						                       sizeof(int) * (size_t)P, hipMemcpyDeviceToHost, q));
						hipEvent_t e = take_event();
						MFX_HIP(hipEventRecord(e, q));
						pending.push_back(Check{its, slot, e});
					}
				}
				if (pending.empty()) break;
				Check c = pending.front();
				pending.pop_front();
				MFX_HIP(hipEventSynchronize(c.ev));
				evPool.push_back(c.ev);
				checked = c.its;
				const int* dn = pinned ? hdone + (size_t)c.slot * P : done.data();
				bool all = true;
				for (int i = 0; i < P && all; i++) all = dn[i] != 0;
				converged = all || (its >= maxIt && pending.empty());
			}
		} catch (...) {
			for (auto& c : pending) evPool.push_back(c.ev);
			for (hipEvent_t e : evPool) (void)hipEventDestroy(e);
			throw;
		}
		for (auto& c : pending) evPool.push_back(c.ev);
		for (hipEvent_t e : evPool) (void)hipEventDestroy(e);
		batch_policy_done(tkey, policy,
		                  std::chrono::duration<double>(std::chrono::steady_clock::now() - tIter).count());
		for (int i = 0; i < P; i++) R[(size_t)i].launched = its;
		// finish, batched: every problem's true residual and its sum, one
		// copy of all CgStates to the host, each problem's state into its
		// scene's host mirror (as a single solve leaves it)
		run_true_residual_many(q, B.tres.n(), B.tres.dItems, B.tres.dParams);
		run_reduce_true_rr_many(q, B.tred.n(), B.tred.dItems, B.tred.dParams);
		CgState* dGather = (CgState*)(B.dmem + gatherOff);
		hipLaunchKernelGGL(k_gather_states, dim3((P + 255) / 256), dim3(256), 0, q, P, (CgState* const*)B.dStates,
		                   dGather);
		MFX_HIP(hipMemcpyAsync(s0->hstage, dGather, sizeof(CgState) * (size_t)P, hipMemcpyDeviceToHost, q));
		sync(s0);
		for (int i = 0; i < P; i++)
			std::memcpy(vel[(size_t)i]->s->hscal, (const CgState*)s0->hstage + i, sizeof(CgState));
		// the relTol check on the true residual, as in a single solve; a
		// problem that needs a restart continues alone on its own stream
		// (and has synchronized it when cg_verify returns)
		const double target = o.relTol > 0 ? o.relTol : (double)o.cgAccuracy;
		for (int i = 0; i < P; i++) {
			cg_verify(R[(size_t)i], target);
			cg_release(R[(size_t)i]);
		}
		run_correct_velocity_many(q, B.corr.n(), B.corr.dItems, B.corr.dParams);
		sync(s0);
		std::vector<std::string> errs((size_t)P);
		for (int i = 0; i < P; i++) {
			try {
				cg_finish_collect(R[(size_t)i], stats ? &stats[i] : nullptr);
			} catch (const Error& e) {
				errs[(size_t)i] = e.what();
			}
		}
		std::string err;
		for (const auto& e : errs)
			if (err.empty() && !e.empty()) err = e;
		release();
		if (!err.empty()) throw Error(err);
	} catch (...) {
		release();
		throw;
	}
}

}  // namespace mfx

using namespace mfx;

// Not part of the public interface (include/mfx.h): the plane chunks the
// autotuner picked for a grid shape (0 = not tuned).  Used by tests and the
// bench report.
extern "C" void mfx_testing_zmarch_chunks(int nx, int ny, int nz, int* mv, int* l0) {
	Dims d{};
	d.nx = nx;
	d.ny = ny;
	d.nz = nz;
	const ZmTune t = zm_tuned(d);
	if (mv) *mv = t.mv;
	if (l0) *l0 = t.l0;
}
// Testing hook: set the plane chunks that single solves of a shape use
// (multiples of 4 up to 64; 0,0 = autotune again).  Returns 0, or -1 for a
// chunk the z-marching kernels do not accept.
extern "C" int mfx_testing_set_zmarch_chunks(int nx, int ny, int nz, int mv, int l0) {
	auto ok = [](int c) { return c > 0 && c <= kZmMaxChunk && c % kDotGroup == 0; };
	if (!(mv == 0 && l0 == 0) && !(ok(mv) && ok(l0))) return -1;
	Dims d{};
	d.nx = nx;
	d.ny = ny;
	d.nz = nz;
	zm_set_tuned(d, ZmTune{mv, l0});
	return 0;
}

struct mfx_grid : public Grid {};
struct mfx_scene : public Scene {};
static inline Grid* G(const mfx_grid* g) { return const_cast<Grid*>(static_cast<const Grid*>(g)); }

#define API_BEGIN try {
#define API_END                                 \
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

static mfx_pressure_opts opts_or_default(const mfx_pressure_opts* o) {
	mfx_pressure_opts r;
	if (o) r = *o;
	else mfx_pressure_opts_default(&r);
	return r;
}

static void check_pressure_args(Grid* vel, Grid* p, Grid* flags) {
	MFX_CHECK(vel && p && flags, "solvePressure: vel, pressure and flags are required");
	MFX_CHECK(vel->isMAC(), "solvePressure: vel must be a MACGrid");
	MFX_CHECK(p->isReal(), "solvePressure: pressure must be a RealGrid");
	MFX_CHECK(flags->isFlags(), "solvePressure: flags must be a FlagGrid");
	MFX_CHECK(vel->s == p->s && vel->s == flags->s, "grids belong to different solvers");
}

extern "C" {

void mfx_pressure_opts_default(mfx_pressure_opts* o) {
	o->cgAccuracy = 1e-3f;
	o->cgMaxIterFac = 1.5f;
	o->preconditioner = MFX_PC_MIC;
	o->enforceCompatibility = 0;
	o->useL2Norm = 0;
	o->zeroPressureFixing = 0;
	o->relTol = 0;
	o->maxIterOverride = 0;
	o->precision = 0;
	o->trueResidualEvery = 0;
	o->deferCorrection = 0;
}

void mfx_pressure_bnd_default(mfx_pressure_bnd* b) {
	if (!b) return;
	memset(b, 0, sizeof(*b));
	b->gfClamp = 1e-04f;
	b->surfTens = 0.f;
}

}  // extern "C"

// the optional grids of the second-order boundaries, checked and resolved to
// device pointers (MAC / Vec3 grids as their three SoA components)
static BndDev bnd_of(const mfx_pressure_bnd* b, const Grid* flags) {
	BndDev B{};
	B.gfClamp = b ? b->gfClamp : 1e-04f;
	B.surfTens = b ? b->surfTens : 0.f;
	if (!b) return B;
	auto real = [&](const mfx_grid* h, const char* name) -> const float* {
		if (!h) return nullptr;
		const Grid* g = G(h);
		MFX_CHECK(g->isReal(), std::string("solvePressure: ") + name + " must be a Real or Levelset grid");
		MFX_CHECK(g->s == flags->s, "grids belong to different solvers");
		return g->real();
	};
	auto vec = [&](const mfx_grid* h, const char* name, const float** c) {
		c[0] = c[1] = c[2] = nullptr;
		if (!h) return;
		const Grid* g = G(h);
		MFX_CHECK(g->isVec(), std::string("solvePressure: ") + name + " must be a MACGrid");
		MFX_CHECK(g->s == flags->s, "grids belong to different solvers");
		for (int q = 0; q < 3; q++) c[q] = g->comp(q);
	};
	B.phi = real(b->phi, "phi");
	B.pcc = real(b->perCellCorr, "perCellCorr");
	B.curv = real(b->curv, "curv");
	const float* fr[3];
	vec(b->fractions, "fractions", fr);
	B.fx = fr[0];
	B.fy = fr[1];
	B.fz = fr[2];
	const float* ob[3];
	vec(b->obvel, "obvel", ob);
	B.ox = ob[0];
	B.oy = ob[1];
	B.oz = ob[2];
	return B;
}

extern "C" {

int mfx_solve_pressure_bnd(mfx_grid* vel, mfx_grid* pressure, const mfx_grid* flags, const mfx_pressure_opts* opts,
                           const mfx_pressure_bnd* bnd, mfx_solve_stats* stats, mfx_grid* retRhs) {
	API_BEGIN
	Grid *v = G(vel), *p = G(pressure), *f = G(flags);
	check_pressure_args(v, p, f);
	MFX_CHECK(!retRhs || G(retRhs)->isReal(), "solvePressure: retRhs must be a RealGrid");
	const mfx_pressure_opts o = opts_or_default(opts);
	MFX_CHECK(!o.deferCorrection, "deferCorrection is honoured by mfx_solve_pressure only");
	const BndDev B = bnd_of(bnd, f);
	Scene* s = v->s;
	float* rhs = s->getReal();
	try {
		compute_rhs(s, f, v, rhs, o.enforceCompatibility, nullptr, &B);
		solve_system(s, f, rhs, p->real(), o, stats, nullptr, &B);
		correct_velocity(s, f, v, p->real(), &B);
		if (retRhs)
			MFX_HIP(hipMemcpyAsync(G(retRhs)->real(), rhs, sizeof(float) * (size_t)s->d.n, hipMemcpyDeviceToDevice,
			                       s->stream));
	} catch (...) {
		s->putReal(rhs);
		throw;
	}
	s->putReal(rhs);
	API_END
}

int mfx_compute_pressure_rhs_bnd(mfx_grid* rhs, const mfx_grid* vel, const mfx_grid* flags,
                                 const mfx_pressure_opts* opts, const mfx_pressure_bnd* bnd) {
	API_BEGIN
	Grid *r = G(rhs), *v = G(vel), *f = G(flags);
	check_pressure_args(v, r, f);
	const mfx_pressure_opts o = opts_or_default(opts);
	const BndDev B = bnd_of(bnd, f);
	compute_rhs(v->s, f, v, r->real(), o.enforceCompatibility, nullptr, &B);
	API_END
}

int mfx_solve_pressure_system_bnd(mfx_grid* rhs, mfx_grid* pressure, const mfx_grid* flags,
                                  const mfx_pressure_opts* opts, const mfx_pressure_bnd* bnd, mfx_solve_stats* stats) {
	API_BEGIN
	Grid *r = G(rhs), *p = G(pressure), *f = G(flags);
	MFX_CHECK(r && p && f && r->isReal() && p->isReal() && f->isFlags(),
	          "solvePressureSystem: needs rhs, pressure (RealGrid) and flags (FlagGrid)");
	MFX_CHECK(r->s == p->s && r->s == f->s, "grids belong to different solvers");
	const mfx_pressure_opts o = opts_or_default(opts);
	const BndDev B = bnd_of(bnd, f);
	solve_system(r->s, f, r->real(), p->real(), o, stats, nullptr, &B);
	API_END
}

int mfx_correct_velocity_bnd(mfx_grid* vel, const mfx_grid* pressure, const mfx_grid* flags,
                             const mfx_pressure_bnd* bnd) {
	API_BEGIN
	Grid *v = G(vel), *p = G(pressure), *f = G(flags);
	check_pressure_args(v, p, f);
	const BndDev B = bnd_of(bnd, f);
	correct_velocity(v->s, f, v, p->real(), &B);
	API_END
}

int mfx_solve_pressure(mfx_grid* vel, mfx_grid* pressure, const mfx_grid* flags, const mfx_pressure_opts* opts,
                       mfx_solve_stats* stats, mfx_grid* retRhs) {
	API_BEGIN
	Grid *v = G(vel), *p = G(pressure), *f = G(flags);
	check_pressure_args(v, p, f);
	MFX_CHECK(!retRhs || G(retRhs)->isReal(), "solvePressure: retRhs must be a RealGrid");
	const mfx_pressure_opts o = opts_or_default(opts);
	Scene* s = v->s;
	// without a pinned cell the matrix code is written by the RHS pass
	const bool fuseCode = !(o.zeroPressureFixing || o.cgAccuracy < 1e-07f);
	if (fuseCode && !s->code) MFX_HIP(hipMalloc((void**)&s->code, (size_t)s->d.n));
	float* rhs = s->getReal();
	try {
		compute_rhs(s, f, v, rhs, o.enforceCompatibility, nullptr, nullptr, fuseCode ? s->code : nullptr);
		solve_system(s, f, rhs, p->real(), o, stats, nullptr, nullptr, fuseCode);
		if (!o.deferCorrection) correct_velocity(s, f, v, p->real());
		if (retRhs)
			MFX_HIP(hipMemcpyAsync(G(retRhs)->real(), rhs, sizeof(float) * (size_t)s->d.n, hipMemcpyDeviceToDevice,
			                       s->stream));
	} catch (...) {
		s->putReal(rhs);
		throw;
	}
	s->putReal(rhs);
	API_END
}

// ---------------------------------------------------------------------------
// Allocation-free plain-CG solve: every buffer beyond the caller's grids
// (right-hand side, CG vectors, matrix code, FP64 partial sums) lives in one
// file-scope device array.  Each GPU holds its own copy of the array, whose
// address hipGetSymbolAddress gives once per device; host code uses only that
// address, never the array's name.
// ---------------------------------------------------------------------------
namespace {
constexpr long long kNoAllocMaxCells = 1ll << 20;
constexpr size_t kNoAllocMaxPartBlocks = 65536;  // partial-sum slots (blocks) of one launch
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaVec = al256((size_t)kNoAllocMaxCells * sizeof(float));
constexpr size_t kNaCode = al256((size_t)kNoAllocMaxCells);
constexpr size_t kNaDPart = al256((3 * kNoAllocMaxPartBlocks + 16) * sizeof(double));
constexpr size_t kNaFPart = al256((2 * kNoAllocMaxPartBlocks + 16) * sizeof(float));
constexpr size_t kNoAllocBytes = 5 * kNaVec + kNaCode + kNaDPart + kNaFPart;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_dev[64];

// device address of `scratch` on device dev (looked up once per device, with
// that device current)
static unsigned char* scratch_on(int dev) {
	MFX_CHECK(dev >= 0 && dev < 64, "solvePressureNoAlloc: device index out of range");
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_dev[dev]) {
		int cur = 0;
		MFX_HIP(hipGetDevice(&cur));
		if (cur != dev) MFX_HIP(hipSetDevice(dev));
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (cur != dev) MFX_HIP(hipSetDevice(cur));
		MFX_CHECK(e == hipSuccess && p != nullptr,
		          std::string("solvePressureNoAlloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
		g_scratch_dev[dev] = p;
	}
	return (unsigned char*)g_scratch_dev[dev];
}

int mfx_solve_pressure_batch(int count, mfx_grid* const* vel, mfx_grid* const* pressure,
                             const mfx_grid* const* flags, const mfx_pressure_opts* opts, mfx_solve_stats* stats) {
	API_BEGIN
	MFX_CHECK(count >= 1 && vel && pressure && flags, "solvePressureBatch: needs count >= 1 and the grid arrays");
	std::vector<Grid*> v((size_t)count), p((size_t)count), f((size_t)count);
	for (int i = 0; i < count; i++) {
		v[(size_t)i] = G(vel[i]);
		p[(size_t)i] = G(pressure[i]);
		f[(size_t)i] = G(flags[i]);
		check_pressure_args(v[(size_t)i], p[(size_t)i], f[(size_t)i]);
		for (int j = 0; j < i; j++)
			MFX_CHECK(v[(size_t)j]->s != v[(size_t)i]->s, "solvePressureBatch: each problem needs its own solver");
	}
	MFX_CHECK(!opts || !opts->deferCorrection, "deferCorrection is honoured by mfx_solve_pressure only");
	solve_batch(v, p, f, opts_or_default(opts), stats);
	API_END
}

int mfx_solve_pressure_noalloc(mfx_grid* vel, mfx_grid* pressure, const mfx_grid* flags,
                               const mfx_pressure_opts* opts, mfx_solve_stats* stats) {
	API_BEGIN
	Grid *v = G(vel), *p = G(pressure), *f = G(flags);
	check_pressure_args(v, p, f);
	const mfx_pressure_opts o = opts_or_default(opts);
	MFX_CHECK(!o.deferCorrection, "deferCorrection is honoured by mfx_solve_pressure only");
	Scene* s = v->s;
	const Dims& d = s->d;
	MFX_CHECK(o.preconditioner == MFX_PC_NONE || (o.preconditioner == MFX_PC_MIC && !d.is3d),
	          "solvePressureNoAlloc: only unpreconditioned CG (PcNone, or PcMIC in 2D) runs without allocation");
	MFX_CHECK(d.n <= kNoAllocMaxCells, "solvePressureNoAlloc: grid has " + std::to_string(d.n) +
	                                       " cells, more than the allocation-free solve's maximum of " +
	                                       std::to_string(kNoAllocMaxCells));
	// partial-sum slots of the launches of this shape (RHS, CG init/update,
	// matvec, true residual) must fit the scratch partials
	{
		long long per = 0;
		const long long nbT = tiles_for(d).count();
		const long long nbF = flat_count(d.n);
		const long long nbL = tiles_for(d, kLightKZ).count();
		const long long nbU = update_tiles(d, &per);
		int kc = 1;
		matvec_tiles(d, &kc, true);
		const long long nbMv = zm_dot_parts(d, kMvTY);
		const long long need = std::max(std::max(nbT, nbF), std::max(std::max(nbL, nbU), nbMv));
		MFX_CHECK(need <= (long long)kNoAllocMaxPartBlocks,
		          "solvePressureNoAlloc: grid shape needs " + std::to_string(need) +
		              " partial sums per launch, more than the allocation-free solve's maximum of " +
		              std::to_string(kNoAllocMaxPartBlocks));
	}
	int dev = 0;
	MFX_HIP(hipGetDevice(&dev));
	MFX_CHECK(dev == s->device, "solvePressureNoAlloc: the solver's device is not the current device");
	unsigned char* base = scratch_on(dev);
	float* rhs = (float*)base;
	CgRun B;
	B.r = (float*)(base + kNaVec);
	B.pa = (float*)(base + 2 * kNaVec);
	B.pb = (float*)(base + 3 * kNaVec);
	B.q = (float*)(base + 4 * kNaVec);
	B.code = base + 5 * kNaVec;
	B.dpart = (double*)(base + 5 * kNaVec + kNaCode);
	B.fpart = (float*)(base + 5 * kNaVec + kNaCode + kNaDPart);
	compute_rhs(s, f, v, rhs, o.enforceCompatibility, B.dpart);
	solve_system(s, f, rhs, p->real(), o, stats, &B);
	correct_velocity(s, f, v, p->real());
	API_END
}

int mfx_cg_solve_diffusion(const mfx_grid* flags, mfx_grid* grid, float alpha, float cgMaxIterFac, float cgAccuracy,
                           mfx_solve_stats* stats) {
	API_BEGIN
	MFX_CHECK(flags && grid, "cgSolveDiffusion: flags and grid are required");
	Grid *f = G(flags), *g = G(grid);
	MFX_CHECK(f->isFlags(), "cgSolveDiffusion: flags must be a FlagGrid");
	MFX_CHECK(f->s == g->s, "grids belong to different solvers");
	MFX_CHECK(g->isReal() || g->isVec(),
	          "cgSolveDiffusion: Grid Type is not supported (only Real, Vec3, MAC, or Levelset)");
	Scene* s = f->s;
	const size_t bytes = sizeof(float) * (size_t)s->d.n;
	float* rhs = s->getReal();
	try {
		// Vec3 / MAC grids: every component (two in 2D) with the same matrix;
		// a component of the SoA grid is itself a Real grid, so getComponent
		// and setComponent are the solve's own input and output
		const int nc = g->isVec() ? (s->d.is3d ? 3 : 2) : 1;
		for (int c = 0; c < nc; c++) {
			float* u = g->comp(c);
			MFX_HIP(hipMemcpyAsync(rhs, u, bytes, hipMemcpyDeviceToDevice, s->stream));
			solve_scaled(s, f, rhs, u, 0, alpha, cgAccuracy, cgMaxIterFac, stats);
		}
	} catch (...) {
		s->putReal(rhs);
		throw;
	}
	s->putReal(rhs);
	API_END
}

int mfx_cg_solve_we(const mfx_grid* flags, mfx_grid* ut, mfx_grid* utm1, mfx_grid* out, int crankNic, float cSqr,
                    float cgMaxIterFac, float cgAccuracy, mfx_solve_stats* stats) {
	API_BEGIN
	MFX_CHECK(flags && ut && utm1 && out, "cgSolveWE: flags, ut, utm1 and out are required");
	Grid *f = G(flags), *a = G(ut), *b = G(utm1), *o = G(out);
	MFX_CHECK(f->isFlags() && a->isReal() && b->isReal() && o->isReal(),
	          "cgSolveWE: needs a FlagGrid and three RealGrids");
	MFX_CHECK(f->s == a->s && f->s == b->s && f->s == o->s, "grids belong to different solvers");
	MFX_CHECK(a != b && a != o && b != o, "cgSolveWE: ut, utm1 and out must be different grids");
	Scene* s = f->s;
	// Real s = dt*dt*cSqr * 0.5 (waves.cpp:107)
	const float dt = s->dt;
	volatile float dd = dt * dt;
	volatile float dc = dd * cSqr;
	const float sc = (float)((double)dc * 0.5);
	float* rhs = s->getReal();
	try {
		const Tiles t = tiles_for(s->d, kLightKZ);
		hipLaunchKernelGGL(k_make_rhs_we, dim3((unsigned)t.count()), dim3(kBlock), 0, s->stream, s->d, t,
		                   (const float*)a->real(), (const float*)b->real(), sc, crankNic ? 1 : 0, rhs);
		solve_scaled(s, f, rhs, o->real(), 1, sc, cgAccuracy, cgMaxIterFac, stats);
		// utm1.swap(ut); ut.copyFrom(out)
		std::swap(a->data, b->data);
		MFX_HIP(hipMemcpyAsync(a->real(), o->real(), sizeof(float) * (size_t)s->d.n, hipMemcpyDeviceToDevice,
		                       s->stream));
	} catch (...) {
		s->putReal(rhs);
		throw;
	}
	s->putReal(rhs);
	API_END
}

int mfx_compute_pressure_rhs(mfx_grid* rhs, const mfx_grid* vel, const mfx_grid* flags, const mfx_pressure_opts* opts) {
	API_BEGIN
	Grid *r = G(rhs), *v = G(vel), *f = G(flags);
	check_pressure_args(v, r, f);
	const mfx_pressure_opts o = opts_or_default(opts);
	compute_rhs(v->s, f, v, r->real(), o.enforceCompatibility);
	API_END
}

int mfx_solve_pressure_system(mfx_grid* rhs, mfx_grid* pressure, const mfx_grid* flags, const mfx_pressure_opts* opts,
                              mfx_solve_stats* stats) {
	API_BEGIN
	Grid *r = G(rhs), *p = G(pressure), *f = G(flags);
	MFX_CHECK(r && p && f && r->isReal() && p->isReal() && f->isFlags(),
	          "solvePressureSystem: needs rhs, pressure (RealGrid) and flags (FlagGrid)");
	MFX_CHECK(r->s == p->s && r->s == f->s, "grids belong to different solvers");
	const mfx_pressure_opts o = opts_or_default(opts);
	solve_system(r->s, f, r->real(), p->real(), o, stats);
	API_END
}

int mfx_correct_velocity(mfx_grid* vel, const mfx_grid* pressure, const mfx_grid* flags) {
	API_BEGIN
	Grid *v = G(vel), *p = G(pressure), *f = G(flags);
	check_pressure_args(v, p, f);
	correct_velocity(v->s, f, v, p->real());
	API_END
}

int mfx_correct_velocity_wall_bcs(mfx_grid* vel, const mfx_grid* pressure, const mfx_grid* flags) {
	API_BEGIN
	Grid *v = G(vel), *p = G(pressure), *f = G(flags);
	check_pressure_args(v, p, f);
	Scene* s = v->s;
	const Tiles t = tiles_for(s->d, kLightKZ);
	launch_one<CorrVelP, body_correct_velocity_wall, kBlock>(
	    s->stream, t.count(), CorrVelP{s->d, t, f->f8, p->real(), v->comp(0), v->comp(1), v->comp(2)});
	API_END
}

int mfx_release_mg(mfx_scene* s) {
	API_BEGIN
	// without a scene: every live scene's hierarchies (pressure.cpp:254-258)
	std::vector<Scene*> all = s ? std::vector<Scene*>{static_cast<Scene*>(s)} : live_scenes();
	for (Scene* sc : all) {
		sync(sc);
		sc->mgStatic.reset();
		sc->mgMic.reset();
	}
	API_END
}

}  // extern "C"
