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
#include <chrono>
#include <cmath>
#include <cstring>

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

__global__ void __launch_bounds__(kBlock) k_build_code(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                       uint8_t* __restrict__ code, long long fix) {
	CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const int i = p.i, j = p.j, k = p.k;
	const long long c = d.idx(i, j, k);
	const bool fl = f8[c] & kFluid;
	int A0 = 0, bx = 0, by = 0, bz = 0;
	if (fl && interior_c(d, i, j, k)) {
		if (c == fix) {
			A0 = 1;
		} else {
			A0 += !(f8[c - 1] & kObstacle);
			A0 += !(f8[c + 1] & kObstacle);
			A0 += !(f8[c - d.nx] & kObstacle);
			A0 += !(f8[c + d.nx] & kObstacle);
			if (d.is3d) {
				A0 += !(f8[c - d.sz] & kObstacle);
				A0 += !(f8[c + d.sz] & kObstacle);
			}
			bx = off_coef(d, f8, c, i, j, k, c + 1, fix);
			by = off_coef(d, f8, c, i, j, k, c + d.nx, fix);
			bz = d.is3d ? off_coef(d, f8, c, i, j, k, c + d.sz, fix) : 0;
		}
	}
	int lo = 0;  // any lower off-diagonal into c
	if (i > 0 && off_coef(d, f8, c - 1, i - 1, j, k, c, fix)) lo = 1;
	if (j > 0 && off_coef(d, f8, c - d.nx, i, j - 1, k, c, fix)) lo = 1;
	if (d.is3d && k > 0 && off_coef(d, f8, c - d.sz, i, j, k - 1, c, fix)) lo = 1;
	const int trivial = (A0 == 1 && !bx && !by && !bz && !lo) ? 1 : 0;
	code[c] = (uint8_t)(A0 | (bx << 3) | (by << 4) | (bz << 5) | ((fl ? 1 : 0) << 6) | (trivial << 7));
}

__device__ inline float cA0(uint8_t c) { return (float)(c & 7); }
__device__ inline float cAi(uint8_t c) { return (c & 8) ? -1.0f : 0.0f; }
__device__ inline float cAj(uint8_t c) { return (c & 16) ? -1.0f : 0.0f; }
__device__ inline float cAk(uint8_t c) { return (c & 32) ? -1.0f : 0.0f; }

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
__global__ void __launch_bounds__(kBlock) k_make_rhs(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                     const float* __restrict__ u, const float* __restrict__ v,
                                                     const float* __restrict__ w, float* __restrict__ rhs,
                                                     double* __restrict__ part) {
	__shared__ double sh[8];
	CellPos p = tile_cell(t);
	double sum = 0, cnt = 0;
	if (p.i < d.nx && p.j < d.ny) {
		const long long c = d.idx(p.i, p.j, p.k);
		float set = 0.f;
		if (interior_c(d, p.i, p.j, p.k) && (f8[c] & kFluid)) {
			set = u[c] - u[c + 1] + v[c] - v[c + d.nx];
			if (d.is3d) set += w[c] - w[c + d.sz];
			sum = (double)set;
			cnt = 1;
		}
		rhs[c] = set;
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

// knCorrectVelocity (pressure.cpp:87-109), bnd=1
__global__ void __launch_bounds__(kBlock) k_correct_velocity(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                             const float* __restrict__ pr, float* __restrict__ u,
                                                             float* __restrict__ v, float* __restrict__ w) {
	CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	if (!interior_c(d, p.i, p.j, p.k)) return;
	const long long c = d.idx(p.i, p.j, p.k);
	const long long xm = c - 1, ym = c - d.nx, zm = c - d.sz;
	const uint8_t f = f8[c];
	if (f & kFluid) {
		const float pc = pr[c];
		if (f8[xm] & kFluid) u[c] -= (pc - pr[xm]);
		if (f8[ym] & kFluid) v[c] -= (pc - pr[ym]);
		if (d.is3d && (f8[zm] & kFluid)) w[c] -= (pc - pr[zm]);
		if (f8[xm] & kEmpty) u[c] -= pc;
		if (f8[ym] & kEmpty) v[c] -= pc;
		if (d.is3d && (f8[zm] & kEmpty)) w[c] -= pc;
	} else if ((f & kEmpty) && !(f & kOutflow)) {
		if (f8[xm] & kFluid) u[c] += pr[xm];
		else u[c] = 0.f;
		if (f8[ym] & kFluid) v[c] += pr[ym];
		else v[c] = 0.f;
		if (d.is3d) {
			if (f8[zm] & kFluid) w[c] += pr[zm];
			else w[c] = 0.f;
		}
	}
}

// ---------------------------------------------------------------------------
// CG kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_cg_init(Dims d, const float* __restrict__ rhs, float* __restrict__ x,
                                                    float* __restrict__ r, int withDot, double* __restrict__ part) {
	__shared__ double sh[8];
	const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
	double zr = 0, bb = 0;
	if (c < d.n) {
		const float b = rhs[c];
		x[c] = 0.f;
		r[c] = b;
		zr = (double)(b * b);  // GridDotProduct(tmp=r, r): float product
		bb = (double)b * (double)b;
	}
	const double s0 = block_sum(zr, sh);
	const double s1 = block_sum(bb, sh);
	if (threadIdx.x == 0) {
		part[3 * blockIdx.x] = withDot ? s0 : 0.0;
		part[3 * blockIdx.x + 1] = s1;
		part[3 * blockIdx.x + 2] = 0.0;
	}
}

// sum fl(z*r) (GridDotProduct(mTmp, mResidual)) into part[3*b]
__global__ void __launch_bounds__(kBlock) k_dot_zr(Dims d, const CgState* __restrict__ st, const float* __restrict__ z,
                                                   const float* __restrict__ r, double* __restrict__ part) {
	__shared__ double sh[8];
	if (st->done) return;
	const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
	double a = 0;
	if (c < d.n) a = (double)(z[c] * r[c]);
	const double s = block_sum(a, sh);
	if (threadIdx.x == 0) part[3 * blockIdx.x] = s;
}

// A: p_new = z + beta*p_old ; q = A p_new ; partial sums of fl(q*p_new).
// z-marching form: a 64 x TY tile of columns walks a chunk of planes.  The
// search direction of the current plane (tile + 1-cell ring) is staged in
// LDS, the planes below and above stay in registers, so every value is read
// from HBM once.  Row arithmetic and term order follow apply_row; neighbours
// outside the grid contribute 0 (their coefficient is 0 there).
constexpr int kMvTY = 8;
template <int TY>
__global__ void __launch_bounds__(64 * TY) k_cg_matvec25(Dims d, int kchunk, const CgState* __restrict__ st,
                                                         const uint8_t* __restrict__ code, const float* __restrict__ z,
                                                         const float* __restrict__ pold, float* __restrict__ pnew,
                                                         float* __restrict__ q, double* __restrict__ part) {
	constexpr int TX = 64, RX = TX + 2, RY = TY + 2, NH = 2 * RX + 2 * (RY - 2);
	__shared__ float ps[RX * RY];
	__shared__ uint8_t cs[RX * RY];
	__shared__ double sh[16];
	if (st->done) return;
	const float beta = st->beta;
	const bool first = st->first;
	const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
	const long long blk = xcd_tile(blockIdx.x, gridDim.x);
	const int ox = (int)(blk % ntx) * TX;
	const long long rest = blk / ntx;
	const int oy = (int)(rest % nty) * TY;
	const int k0 = (int)(rest / nty) * kchunk, k1 = min(d.nz, k0 + kchunk);
	const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	const int gx = ox + tx, gy = oy + ty;
	const bool in = gx < d.nx && gy < d.ny;
	auto pn = [&](long long m) -> float { return first ? z[m] : z[m] + beta * pold[m]; };
	// own column: planes k-1, k, k+1 in registers, plane k+2 prefetched
	const long long gcol = in ? (long long)gx + (long long)d.nx * gy : 0;
	auto fcol = [&](int kk, float& v, uint8_t& c) {
		v = 0.f;
		c = 0;
		if (in && kk >= 0 && kk < d.nz) {
			const long long m = gcol + d.sz * kk;
			v = pn(m);
			c = code[m];
		}
	};
	// one ring cell per thread (threads < NH), prefetched one plane ahead
	static_assert(NH <= 64 * TY, "one ring cell per thread");
	const int h = threadIdx.x;
	int hx = 0, hy = 0;
	if (h < RX) { hx = h; hy = 0; }
	else if (h < 2 * RX) { hx = h - RX; hy = RY - 1; }
	else if (h < 2 * RX + RY - 2) { hx = 0; hy = 1 + h - 2 * RX; }
	else { hx = RX - 1; hy = 1 + h - 2 * RX - (RY - 2); }
	const int hgx = ox - 1 + hx, hgy = oy - 1 + hy;
	const bool hin = h < NH && hgx >= 0 && hgy >= 0 && hgx < d.nx && hgy < d.ny;
	const long long hcol = hin ? (long long)hgx + (long long)d.nx * hgy : 0;
	auto fhalo = [&](int kk, float& v, uint8_t& c) {
		v = 0.f;
		c = 0;
		if (hin && kk < d.nz) {
			const long long m = hcol + d.sz * kk;
			v = pn(m);
			c = code[m];
		}
	};
	float pPrev, pCur, pNext, pN2, hv;
	uint8_t cPrev, cCur, cNext, cN2, hc;
	fcol(k0 - 1, pPrev, cPrev);
	fcol(k0, pCur, cCur);
	fcol(k0 + 1, pNext, cNext);
	fhalo(k0, hv, hc);
	const int L = (tx + 1) + RX * (ty + 1);
	double acc = 0.0;
	for (int k = k0; k < k1; k++) {
		ps[L] = pCur;
		cs[L] = cCur;
		if (h < NH) {
			ps[hx + RX * hy] = hv;
			cs[hx + RX * hy] = hc;
		}
		fcol(k + 2, pN2, cN2);
		fhalo(k + 1, hv, hc);
		__syncthreads();
		if (in) {
			float qv;
			if (!(cCur & 64)) qv = pCur;
			else if (d.is3d)
				qv = pCur * cA0(cCur) + ps[L - 1] * cAi(cs[L - 1]) + ps[L + 1] * cAi(cCur) + ps[L - RX] * cAj(cs[L - RX]) +
				     ps[L + RX] * cAj(cCur) + pPrev * cAk(cPrev) + pNext * cAk(cCur);
			else
				qv = pCur * cA0(cCur) + ps[L - 1] * cAi(cs[L - 1]) + ps[L + 1] * cAi(cCur) + ps[L - RX] * cAj(cs[L - RX]) +
				     ps[L + RX] * cAj(cCur);
			const long long m = d.idx(gx, gy, k);
			pnew[m] = pCur;
			q[m] = qv;
			acc += (double)(qv * pCur);
		}
		__syncthreads();
		pPrev = pCur;
		cPrev = cCur;
		pCur = pNext;
		cCur = cNext;
		pNext = pN2;
		cNext = cN2;
	}
	const double s = block_sum(acc, sh);
	if (threadIdx.x == 0) part[3 * blockIdx.x] = s;
}

// B: x += alpha p ; r += (-alpha) q ; min/max(r), sum fl(r*r), sum r^2.
// A bounded grid walks the arrays four elements per thread and step.
__global__ void __launch_bounds__(kBlock) k_cg_update(Dims d, const CgState* __restrict__ st, float* __restrict__ x,
                                                      float* __restrict__ r, const float* __restrict__ p,
                                                      const float* __restrict__ q, int withDot,
                                                      double* __restrict__ part, float* __restrict__ fpart) {
	__shared__ double sh[8];
	__shared__ float shf[8];
	if (st->done) return;
	const float alpha = st->alpha, malpha = -alpha;
	float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
	double zr = 0, rr = 0;
	const long long n4 = d.n >> 2;
	const long long stride = (long long)gridDim.x * blockDim.x;
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n4; v += stride) {
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
	for (long long c = (n4 << 2) + blockIdx.x * (long long)blockDim.x + threadIdx.x; c < d.n; c += stride) {
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
		part[3 * blockIdx.x] = withDot ? s0 : 0.0;
		part[3 * blockIdx.x + 1] = s1;
		fpart[2 * blockIdx.x] = a;
		fpart[2 * blockIdx.x + 1] = b;
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

enum { FIN_INIT = 0, FIN_ALPHA = 1, FIN_RES = 2, FIN_BETA = 3 };

__global__ void __launch_bounds__(1024) k_cg_finalize(CgState* st, const double* __restrict__ part,
                                                      const float* __restrict__ fpart, int nb, int phase, int pcNone) {
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
	// FIN_BETA (preconditioned): sigmaNew = dot(z, r)
	const float sigmaNew = (float)r0;
	st->beta = sigmaNew / st->sigma;
	st->sigma = sigmaNew;
	if (!(st->resNorm < 1e35f)) {
		st->diverged = 1;
		st->done = 1;
		return;
	}
	if (st->iter >= st->maxIter) st->done = 1;
}

// true residual ||b - A x||_2^2 with FP64 accumulation of the row sum
__global__ void __launch_bounds__(kBlock) k_true_residual(Dims d, Tiles t, const uint8_t* __restrict__ code,
                                                          const float* __restrict__ b, const float* __restrict__ x,
                                                          double* __restrict__ part) {
	__shared__ double sh[8];
	CellPos p = tile_cell(t);
	double rr = 0;
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
		const double r = (double)b[c] - ax;
		rr = r * r;
	}
	const double s = block_sum(rr, sh);
	if (threadIdx.x == 0) part[3 * blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static unsigned flat_count(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// zero-pressure cell selection (pressure.cpp:349-381), on the host from the
// packed flags (setup-time work)
static long long find_fix_cell(Scene* s, Grid* flags) {
	const Dims& d = s->d;
	std::vector<uint8_t> f((size_t)d.n);
	MFX_HIP(hipMemcpyAsync(f.data(), flags->f8, (size_t)d.n, hipMemcpyDeviceToHost, s->stream));
	sync(s);
	long long numEmpty = 0;
	for (long long i = 0; i < d.n; i++)
		if (f[i] & kEmpty) numEmpty++;
	if (numEmpty != 0) return -1;
	const int tcx = d.nx / 2, tcy = d.ny - 1, tcz = d.is3d ? d.nz / 2 : 0;
	for (int o = 0; o < 3; o++) {
		const int y = tcy - o;
		if (y < 0) continue;
		const long long ix = d.idx(tcx, y, tcz);
		if (f[ix] & kFluid) return ix;
	}
	const int k0 = d.is3d ? 1 : 0, k1 = d.is3d ? d.nz - 1 : 1;
	for (int k = k0; k < k1; k++)
		for (int j = 1; j < d.ny - 1; j++)
			for (int i = 1; i < d.nx - 1; i++)
				if (f[d.idx(i, j, k)] & kFluid) return d.idx(i, j, k);
	return -1;
}

static void compute_rhs(Scene* s, Grid* flags, Grid* vel, float* rhs, int enforceCompatibility) {
	const Tiles t = tiles_for(s->d);
	const unsigned nb = (unsigned)t.count();
	s->ensurePartials(3 * (size_t)nb + 16, 2 * (size_t)nb + 16);
	hipLaunchKernelGGL(k_make_rhs, dim3(nb), dim3(kBlock), 0, s->stream, s->d, t, flags->f8,
	                   (const float*)vel->comp(0), (const float*)vel->comp(1), (const float*)vel->comp(2), rhs,
	                   s->dpart);
	if (enforceCompatibility) {
		std::vector<double> h(2 * (size_t)nb);
		MFX_HIP(hipMemcpyAsync(h.data(), s->dpart, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s->stream));
		sync(s);
		double sum = 0, cnt = 0;
		for (unsigned b = 0; b < nb; b++) {
			sum += h[2 * b];
			cnt += h[2 * b + 1];
		}
		// rhs += (Real)(-sum / (Real)cnt), over all cells (pressure.cpp:297-298)
		const float corr = (float)(-sum / (double)(float)cnt);
		hipLaunchKernelGGL(k_add_const, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, rhs, s->d.n, corr);
	}
}

struct SolveWork {
	float *r = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr, *z = nullptr;
	uint8_t* code = nullptr;
};

// pressure system solve: rhs (modified by fixPressure), pressure out
static void solve_system(Scene* s, Grid* flags, float* rhs, float* x, const mfx_pressure_opts& o,
                         mfx_solve_stats* stats) {
	auto t0 = std::chrono::steady_clock::now();
	const Dims& d = s->d;
	const Tiles t = tiles_for(d);
	const unsigned nbT = (unsigned)t.count();
	const unsigned nbF = flat_count(d.n);
	const unsigned nbMax = nbT > nbF ? nbT : nbF;
	s->ensurePartials(3 * (size_t)nbMax + 16, 2 * (size_t)nbMax + 16);
	// z-marching matvec: 64 x kMvTY column tiles, chunks of planes sized so
	// that about 1024 workgroups cover the grid
	const long long cols = (long long)((d.nx + 63) / 64) * ((d.ny + kMvTY - 1) / kMvTY);
	int kchunk = d.nz;
	while (kchunk > 8 && cols * ((d.nz + kchunk - 1) / kchunk) < 1024) kchunk = (kchunk + 1) / 2;
	const unsigned nbM = (unsigned)(cols * ((d.nz + kchunk - 1) / kchunk));
	// update: grid-stride over float4 groups, bounded partial count
	const unsigned nbU = (unsigned)std::max(1LL, std::min<long long>(2048, (d.n / 4 + kBlock - 1) / kBlock));

	// zero-pressure fixing (pressure.cpp:349-388)
	long long fix = -1;
	if (o.zeroPressureFixing || o.cgAccuracy < 1e-07f) {
		fix = find_fix_cell(s, flags);
		if (fix >= 0) {
			const float zero = 0.f;
			MFX_HIP(hipMemcpyAsync(rhs + fix, &zero, sizeof(float), hipMemcpyHostToDevice, s->stream));
		}
	}
	// matrix
	if (!s->code) MFX_HIP(hipMalloc((void**)&s->code, (size_t)d.n));
	uint8_t* code = s->code;
	hipLaunchKernelGGL(k_build_code, dim3(nbT), dim3(kBlock), 0, s->stream, d, t, flags->f8, code, fix);

	// preconditioner choice (pressure.cpp:404-433, conjugategrad.cpp:302-336)
	int pc = o.preconditioner;
	int mx = d.nx > d.ny ? d.nx : d.ny;
	if (d.nz > mx) mx = d.nz;
	int maxIter;
	if (pc == MFX_PC_NONE || pc == MFX_PC_MIC) {
		maxIter = (int)(o.cgMaxIterFac * (float)mx) * (d.is3d ? 1 : 4);
		// mIC only exists in 3D in the reference; the GPU runs the parallel
		// multigrid preconditioner in its place
		pc = (pc == MFX_PC_MIC && d.is3d) ? MFX_PC_MGSTATIC : MFX_PC_NONE;
	} else {
		maxIter = 100;
	}
	if (o.maxIterOverride > 0) maxIter = o.maxIterOverride;
	const bool pcNone = pc == MFX_PC_NONE;

	MgHierarchy* mg = nullptr;
	std::unique_ptr<MgHierarchy> mgDyn;
	if (!pcNone) {
		if (o.preconditioner == MFX_PC_MGDYNAMIC) {
			s->mgStatic.reset();
			mgDyn.reset(new MgHierarchy());
			mgDyn->build(s, code, o.precision);
			mg = mgDyn.get();
		} else {
			if (!s->mgStatic) {
				s->mgStatic.reset(new MgHierarchy());
				s->mgStatic->build(s, code, o.precision);
			}
			mg = s->mgStatic.get();
		}
		mg->setTrivial(code);
	}

	float* r = s->getReal();
	float* pa = s->getReal();
	float* pb = s->getReal();
	float* q = s->getReal();
	float* z = pcNone ? r : s->getReal();

	CgState hst;
	memset(&hst, 0, sizeof(hst));
	hst.acc = o.cgAccuracy;
	hst.maxIter = maxIter;
	hst.useL2 = o.useL2Norm;
	hst.relTol = o.relTol;
	CgState* st = (CgState*)s->dscal;
	MFX_HIP(hipMemcpyAsync(st, &hst, sizeof(CgState), hipMemcpyHostToDevice, s->stream));

	// doInit (conjugategrad.cpp:210-235)
	hipLaunchKernelGGL(k_cg_init, dim3(nbF), dim3(kBlock), 0, s->stream, d, (const float*)rhs, x, r, pcNone ? 1 : 0,
	                   s->dpart);
	if (!pcNone) {
		// bnorm from the init partials, sigma from dot(z, r) after one V-cycle
		hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(1024), 0, s->stream, st, (const double*)s->dpart,
		                   (const float*)s->fpart, (int)nbF, (int)FIN_INIT, 0);
		int np = mg->vcycle(s, code, r, z, st);
		if (np == 0) {
			hipLaunchKernelGGL(k_dot_zr, dim3(nbF), dim3(kBlock), 0, s->stream, d, (const CgState*)st,
			                   (const float*)z, (const float*)r, s->dpart);
			np = (int)nbF;
		}
		// sigma = dot(z, r); keep bnorm2
		hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(1024), 0, s->stream, st, (const double*)s->dpart,
		                   (const float*)s->fpart, np, (int)FIN_BETA, 0);
	} else {
		hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(1024), 0, s->stream, st, (const double*)s->dpart,
		                   (const float*)s->fpart, (int)nbF, (int)FIN_INIT, 0);
	}
	if (!pcNone) {
		// FIN_BETA above set sigma = dot(z,r); restart the search direction
		static const int one = 1;
		MFX_HIP(hipMemcpyAsync(&st->first, &one, sizeof(int), hipMemcpyHostToDevice, s->stream));
	}

	// iterations in batches; the kernels stop on st->done
	int launched = 0;
	int batch = 4;
	float* pcur = pa;
	float* pnext = pb;
	CgState* hs = (CgState*)s->hscal;
	while (true) {
		const int nIt = std::min(batch, maxIter - launched);
		for (int it = 0; it < nIt; it++) {
			{
				KtScope kt("cg_matvec", s->stream);
				hipLaunchKernelGGL(k_cg_matvec25<kMvTY>, dim3(nbM), dim3(64 * kMvTY), 0, s->stream, d, kchunk,
				                   (const CgState*)st, code, (const float*)z, (const float*)pcur, pnext, q, s->dpart);
			}
			hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(1024), 0, s->stream, st, (const double*)s->dpart,
			                   (const float*)s->fpart, (int)nbM, (int)FIN_ALPHA, pcNone ? 1 : 0);
			{
				KtScope kt("cg_update", s->stream);
				hipLaunchKernelGGL(k_cg_update, dim3(nbU), dim3(kBlock), 0, s->stream, d, (const CgState*)st, x, r,
				                   (const float*)pnext, (const float*)q, pcNone ? 1 : 0, s->dpart, s->fpart);
			}
			hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(1024), 0, s->stream, st, (const double*)s->dpart,
			                   (const float*)s->fpart, (int)nbU, (int)FIN_RES, pcNone ? 1 : 0);
			if (!pcNone) {
				int np = mg->vcycle(s, code, r, z, st);
				if (np == 0) {
					hipLaunchKernelGGL(k_dot_zr, dim3(nbF), dim3(kBlock), 0, s->stream, d, (const CgState*)st,
					                   (const float*)z, (const float*)r, s->dpart);
					np = (int)nbF;
				}
				hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(1024), 0, s->stream, st, (const double*)s->dpart,
				                   (const float*)s->fpart, np, (int)FIN_BETA, 0);
			}
			std::swap(pcur, pnext);
		}
		launched += nIt;
		MFX_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s->stream));
		sync(s);
		if (hs->done || launched >= maxIter) break;
		// predict the remaining iterations from the residual history
		int itNow = hs->iter;
		int predict = batch;
		if (itNow >= 2 && itNow <= 64) {
			const double a = hs->hist[itNow - 2], b = hs->hist[itNow - 1];
			const double target = o.relTol > 0 ? o.relTol : (double)o.cgAccuracy;
			if (a > 0 && b > 0 && b < a && target > 0) {
				const double rate = b / a;
				double need = std::log(target / b) / std::log(rate);
				if (need < 1) need = 1;
				predict = (int)std::ceil(need);
			}
		}
		batch = std::max(1, std::min(predict, 64));
	}
	if (!hs->done) {
		// cap reached without the kernels flagging it (maxIter launched)
	}
	// true relative residual (FP64)
	hipLaunchKernelGGL(k_true_residual, dim3(nbT), dim3(kBlock), 0, s->stream, d, t, code, (const float*)rhs,
	                   (const float*)x, s->dpart);
	std::vector<double> h(3 * (size_t)nbT);
	MFX_HIP(hipMemcpyAsync(h.data(), s->dpart, sizeof(double) * 3 * nbT, hipMemcpyDeviceToHost, s->stream));
	MFX_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s->stream));
	sync(s);
	double rr = 0;
	for (unsigned b = 0; b < nbT; b++) rr += h[3 * b];
	const double bb = hs->bnorm2;

	s->putReal(r);
	s->putReal(pa);
	s->putReal(pb);
	s->putReal(q);
	if (!pcNone) s->putReal(z);

	if (stats) {
		stats->iterations = hs->iter;
		stats->resNorm = hs->resNorm;
		stats->relResidual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
		stats->fixPidx = (int)fix;
		stats->maxIter = maxIter;
		stats->preconditionerUsed = pc;
		stats->converged = hs->converged;
		stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	}
	if (hs->diverged) {
		if (!pcNone)
			fprintf(stderr,
			        "GridCg::iterate: Warning - this diverging solve can be caused by the 'static' mode of the MG "
			        "preconditioner. If the static mode is active, try switching to dynamic.\n");
		throw Error("GridCg::iterate: The CG solver diverged, residual norm > 1e30, stopping.");
	}
}

static void correct_velocity(Scene* s, Grid* flags, Grid* vel, const float* pressure) {
	const Tiles t = tiles_for(s->d);
	hipLaunchKernelGGL(k_correct_velocity, dim3((unsigned)t.count()), dim3(kBlock), 0, s->stream, s->d, t, flags->f8,
	                   pressure, vel->comp(0), vel->comp(1), vel->comp(2));
}

}  // namespace mfx

using namespace mfx;
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
}

int mfx_solve_pressure(mfx_grid* vel, mfx_grid* pressure, const mfx_grid* flags, const mfx_pressure_opts* opts,
                       mfx_solve_stats* stats, mfx_grid* retRhs) {
	API_BEGIN
	Grid *v = G(vel), *p = G(pressure), *f = G(flags);
	check_pressure_args(v, p, f);
	MFX_CHECK(!retRhs || G(retRhs)->isReal(), "solvePressure: retRhs must be a RealGrid");
	const mfx_pressure_opts o = opts_or_default(opts);
	Scene* s = v->s;
	float* rhs = s->getReal();
	try {
		compute_rhs(s, f, v, rhs, o.enforceCompatibility);
		solve_system(s, f, rhs, p->real(), o, stats);
		correct_velocity(s, f, v, p->real());
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

int mfx_release_mg(mfx_scene* s) {
	API_BEGIN
	if (s) static_cast<Scene*>(s)->mgStatic.reset();
	API_END
}

}  // extern "C"
