// Multigrid preconditioner for the pressure CG.
//
// Structure after the reference's GridMg (multigrid.cpp, Dick et al. 2015):
// vertex-centred levels of size (n+2)/2 until every axis is <= 5 or a level
// has <= 1000 vertices (:220-262), full-weighting restriction and
// (tri)linear interpolation with weights 1/2^(#odd coordinates) (:904-959),
// Galerkin coarse operators A_l = R A_{l-1} I stored as a symmetric 27-point
// (9-point in 2D) stencil (:580-666), trivial rows scaled by 1e-6 (:350-415),
// one multi-colour Gauss-Seidel pre- and post-sweep with reversed colour
// order on the way up so that the V-cycle is symmetric (:668-737).
//
// Differences, chosen for the GPU:
//  * coarse vertices are selected by a parallel rule: a coarse vertex is
//    active if any active fine vertex lies in its restriction stencil, and it
//    is dropped again when its Galerkin diagonal vanishes.  The reference's
//    sequential bucket-heap selection (:520-578) keeps the interpolation full
//    rank; here rank-deficient directions are harmless because
//  * the coarsest level is solved exactly by a dense generalised inverse in
//    FP64 (Cholesky of the active operator with zero pivots dropped, built once
//    per hierarchy on the host, applied as a GEMV on the GPU) instead of a
//    sequential FP64 Jacobi-PCG (:796-902).  This also makes the V-cycle well
//    defined for pure-Neumann problems without pinning a pressure value.
//  * level 0 is matrix-free: its stencil comes from the per-cell matrix code.
#include <cmath>
#include <cstring>
#include <type_traits>

#include "mfx_batch.h"
#include "mfx_mg.h"
#include "mfx_prof.h"

namespace mfx {

__device__ __host__ inline int oddc(int i, int j, int k) { return (i & 1) + (j & 1) + (k & 1); }
// 1/2^(#odd coordinates): exact powers of two (same values as the division)
__device__ inline float wgt(int i, int j, int k) {
	const int o = oddc(i, j, k);
	return o == 0 ? 1.0f : (o == 1 ? 0.5f : (o == 2 ? 0.25f : 0.125f));
}

static inline LvDev lvdev(const MgLevel& l) {
	return LvDev{l.nx, l.ny, l.nz, l.sy, l.sz, l.n, fastdiv_make((unsigned)l.nx), fastdiv_make((unsigned)l.ny)};
}

__device__ inline bool l0_active(uint8_t c) { return (c & 7) != 0; }
__device__ inline float l0_diag(uint8_t c) { return (c & 128) ? (float)(c & 7) * 1e-6f : (float)(c & 7); }
__device__ inline float l0_scale(uint8_t c) { return (c & 128) ? 1e-6f : 1.0f; }

// ---------------------------------------------------------------------------
// hierarchy construction
// ---------------------------------------------------------------------------
// Coarse vertices, a parallel restatement of genCoarseGrid
// (multigrid.cpp:520-578).  There, a fine vertex with one undecided parent
// left activates it, and a fine vertex popped with several free parents
// activates the first and removes the others, so each fine vertex adds at
// most one parent beyond the "standard" ones.  Here, in two passes:
//  1. standard: V is active iff its injection point, fine vertex 2V, is active;
//  2. every active fine vertex activates its first (loop-order) parent that
//     is not standard.
// Parents of walls and obstacles (fine 2V inactive) are thereby active where
// a fine vertex needs them, but a fine vertex never turns on several weak,
// nearly dependent parents - those make the Galerkin operator ill-conditioned.
__device__ inline bool fine_active(const uint8_t* code, const uint8_t* ftype, long long u) {
	return code ? l0_active(code[u]) : (ftype[u] != 0);
}
__global__ void k_mg_activate(LvDev c, LvDev f, int is3d, const uint8_t* __restrict__ code,
                              const uint8_t* __restrict__ ftype, uint8_t* __restrict__ ctype) {
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x) {
		const int i = (int)(v % c.nx), j = (int)((v / c.nx) % c.ny), k = (int)(v / c.sz);
		const int ux = 2 * i, uy = 2 * j, uz = is3d ? 2 * k : 0;
		bool act = false;
		if (ux < f.nx && uy < f.ny && uz < f.nz) act = fine_active(code, ftype, ux + f.sy * uy + f.sz * uz);
		ctype[v] = act ? 1 : 0;
	}
}
__global__ void k_mg_cover(LvDev c, LvDev f, int is3d, const uint8_t* __restrict__ code,
                           const uint8_t* __restrict__ ftype, uint8_t* __restrict__ ctype) {
	for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < f.n;
	     u += (long long)gridDim.x * blockDim.x) {
		if (!fine_active(code, ftype, u)) continue;
		const int ux = (int)(u % f.nx), uy = (int)((u / f.nx) % f.ny), uz = (int)(u / f.sz);
		long long pick = -1;
		for (int iz = uz / 2; iz <= (uz + 1) / 2 && pick < 0; iz++)
			for (int iy = uy / 2; iy <= (uy + 1) / 2 && pick < 0; iy++)
				for (int ix = ux / 2; ix <= (ux + 1) / 2; ix++) {
					if (ix >= c.nx || iy >= c.ny || iz >= c.nz) continue;
					const long long pi = ix + c.sy * iy + c.sz * iz;
					// pass-1 marks (1) decide "standard"; pass-2 marks (3) written by
					// other fine vertices are not standard either, so the choice
					// does not depend on the order of the writes
					if (ctype[pi] != 1) {
						pick = pi;
						break;
					}
				}
		if (pick >= 0) ctype[pick] = 3;
	}
}
__global__ void k_mg_cover_done(LvDev c, uint8_t* __restrict__ ctype) {
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x)
		if (ctype[v] == 3) ctype[v] = 1;
}

// fine-level matrix entry A(U, W) for W = U + off, off in [-1,1]^3;
// level 0 from the code (7-point), coarse levels from the symmetric storage
struct FineOp {
	const uint8_t* code;  // level 0 if non-null
	const float* A;       // coarse storage otherwise
	const uint8_t* type;
	LvDev f;
	int S, C;  // stencil size and centre index (14/13 in 3D, 5/4 in 2D)
	int is3d;
};
__device__ inline float fine_entry(const FineOp& F, long long u, long long w, int dx, int dy, int dz) {
	if (F.code) {
		const int nz = (dx != 0) + (dy != 0) + (dz != 0);
		if (nz == 0) return l0_diag(F.code[u]);
		if (nz > 1) return 0.f;
		if (dx == 1) return (F.code[u] & 8) ? -1.f : 0.f;
		if (dx == -1) return (F.code[w] & 8) ? -1.f : 0.f;
		if (dy == 1) return (F.code[u] & 16) ? -1.f : 0.f;
		if (dy == -1) return (F.code[w] & 16) ? -1.f : 0.f;
		if (dz == 1) return (F.code[u] & 32) ? -1.f : 0.f;
		return (F.code[w] & 32) ? -1.f : 0.f;
	}
	const int s = F.is3d ? (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1) : (dx + 1) + 3 * (dy + 1);
	if (s >= F.C) return F.A[(long long)(s - F.C) * F.f.n + u];
	return F.A[(long long)(F.C - s) * F.f.n + w];
}
__device__ inline bool fine_active(const FineOp& F, long long u) {
	return F.code ? l0_active(F.code[u]) : F.type[u] != 0;
}

// Galerkin operator A_l = R A_{l-1} I at coarse vertex V (upper half), plus a
// diagonal scale proxy used to drop vertices whose diagonal cancels
__global__ void k_mg_galerkin(FineOp F, LvDev c, const uint8_t* __restrict__ ctype, float* __restrict__ Ac,
                              float* __restrict__ proxyOut) {
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x) {
		const int S = F.S, C = F.C;
		if (!ctype[v]) {
			for (int q = 0; q < S; q++) Ac[(long long)q * c.n + v] = 0.f;
			proxyOut[v] = 0.f;
			continue;
		}
		const int vi = (int)(v % c.nx), vj = (int)((v / c.nx) % c.ny), vk = (int)(v / c.sz);
		double acc[14];
		for (int q = 0; q < 14; q++) acc[q] = 0.0;
		double proxy = 0.0;
		const LvDev f = F.f;
		const int uk0 = F.is3d ? max(0, 2 * vk - 1) : 0, uk1 = F.is3d ? min(f.nz - 1, 2 * vk + 1) : 0;
		for (int uz = uk0; uz <= uk1; uz++)
			for (int uy = max(0, 2 * vj - 1); uy <= min(f.ny - 1, 2 * vj + 1); uy++)
				for (int ux = max(0, 2 * vi - 1); ux <= min(f.nx - 1, 2 * vi + 1); ux++) {
					const long long u = ux + f.sy * uy + f.sz * uz;
					if (!fine_active(F, u)) continue;
					const double rw = 1.0 / (double)(1 << oddc(ux, uy, uz));
					const int dz0 = F.is3d ? -1 : 0, dz1 = F.is3d ? 1 : 0;
					for (int dz = dz0; dz <= dz1; dz++)
						for (int dy = -1; dy <= 1; dy++)
							for (int dx = -1; dx <= 1; dx++) {
								if (F.code && ((dx != 0) + (dy != 0) + (dz != 0)) > 1) continue;
								const int wx = ux + dx, wy = uy + dy, wz = uz + dz;
								if (wx < 0 || wy < 0 || wz < 0 || wx >= f.nx || wy >= f.ny || wz >= f.nz) continue;
								const long long w = wx + f.sy * wy + f.sz * wz;
								if (!fine_active(F, w)) continue;
								const float a = fine_entry(F, u, w, dx, dy, dz);
								if (a == 0.f) continue;
								const double iw = 1.0 / (double)(1 << oddc(wx, wy, wz));
								if (dx == 0 && dy == 0 && dz == 0) proxy += rw * fabs((double)a) * iw;
								// parents N of W: [W/2, (W+1)/2]
								for (int nz = wz / 2; nz <= (wz + 1) / 2; nz++)
									for (int ny = wy / 2; ny <= (wy + 1) / 2; ny++)
										for (int nx = wx / 2; nx <= (wx + 1) / 2; nx++) {
											const int ox = nx - vi, oy = ny - vj, oz = nz - vk;
											if (ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1) continue;
											const int s = F.is3d ? (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)
											                     : (ox + 1) + 3 * (oy + 1);
											if (s < C) continue;
											if (nx >= c.nx || ny >= c.ny || nz >= c.nz) continue;
											const long long nn = nx + c.sy * ny + c.sz * nz;
											if (!ctype[nn]) continue;
											acc[s - C] += rw * (double)a * iw;
										}
							}
				}
		for (int q = 0; q < S; q++) Ac[(long long)q * c.n + v] = (float)acc[q];
		proxyOut[v] = (float)proxy;
	}
}

__global__ void k_mg_drop(LvDev c, uint8_t* __restrict__ ctype, const float* __restrict__ Ac,
                          const float* __restrict__ proxy) {
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x) {
		if (ctype[v] && !(Ac[v] > 1e-6f * proxy[v])) ctype[v] = 0;
	}
}

// Full 27-point (9-point in 2D) stencil of vertex v from the symmetric
// upper-half storage, entry C = S-1 being the diagonal.  Returns false when
// a neighbour lies outside the level or is inactive.
__device__ inline bool mg_full_stencil(const LvDev& c, int is3d, int S, const float* __restrict__ A,
                                       const uint8_t* __restrict__ type, long long v, float* out) {
	const int C = S - 1;
	const int i = (int)(v % c.nx), j = (int)((v / c.nx) % c.ny), k = (int)(v / c.sz);
	bool ok = true;
	int s = 0;
	for (int dz = is3d ? -1 : 0; dz <= (is3d ? 1 : 0); dz++)
		for (int dy = -1; dy <= 1; dy++)
			for (int dx = -1; dx <= 1; dx++, s++) {
				const int ni = i + dx, nj = j + dy, nk = k + dz;
				if (ni < 0 || nj < 0 || nk < 0 || ni >= c.nx || nj >= c.ny || nk >= c.nz) {
					ok = false;
					out[s] = 0.f;
					continue;
				}
				const long long nn = ni + c.sy * nj + c.sz * nk;
				if (!type[nn]) ok = false;
				out[s] = s >= C ? A[(long long)(s - C) * c.n + v] : A[(long long)(C - s) * c.n + nn];
			}
	return ok;
}

// full stencils of a few sample vertices (for picking the canonical one)
__global__ void k_mg_sample(LvDev c, int is3d, int S, const float* __restrict__ A, const uint8_t* __restrict__ type,
                            const long long* __restrict__ samples, int ns, float* __restrict__ out,
                            int* __restrict__ ok) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= ns) return;
	const long long v = samples[t];
	float st[27];
	for (int s = 0; s < 27; s++) st[s] = 0.f;
	const bool good = type[v] && mg_full_stencil(c, is3d, S, A, type, v, st);
	for (int s = 0; s < 27; s++) out[27 * t + s] = st[s];
	ok[t] = good ? 1 : 0;
}

// type 2 = active, all neighbours active, stencil bitwise equal to cn.  Only
// nonzero type values are ever written, so concurrent readers testing
// type != 0 are unaffected.
__global__ void k_mg_classify(LvDev c, int is3d, int S, const float* __restrict__ A, uint8_t* __restrict__ type,
                              Canon cn) {
	const int ns = is3d ? 27 : 9;
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x) {
		if (!type[v]) continue;
		float st[27];
		if (!mg_full_stencil(c, is3d, S, A, type, v, st)) continue;
		bool same = true;
		for (int s = 0; s < ns; s++) same = same && (__float_as_uint(st[s]) == __float_as_uint(cn.a[s]));
		if (same) type[v] = 2;
	}
}

// ---------------------------------------------------------------------------
// V-cycle kernels
// ---------------------------------------------------------------------------
// Coarse-level pointwise kernels run on runs of 256 consecutive vertices (a
// thread per vertex; 32-bit indices: coarse levels hold fewer than 2^31
// vertices).
long long lc_tiles(const LvDev& c) { return (c.n + 255) / 256; }
struct VtxPos {
	int i, j, k, v;
	bool in;
};
__device__ inline VtxPos lc_vertex(const LvDev& c, long long blk) {
	VtxPos p;
	p.v = (int)blk * 256 + (int)threadIdx.x;
	p.in = p.v < c.n;
	const unsigned v = p.in ? (unsigned)p.v : 0u, row = fastdiv(v, c.fdx), k = fastdiv(row, c.fdy);
	p.i = (int)(v - row * (unsigned)c.nx);
	p.j = (int)(row - k * (unsigned)c.ny);
	p.k = (int)k;
	return p;
}

// Coarse-level stencil coefficients are FP32, or BF16 (upper 16 bits of the
// FP32 value, round-to-nearest-even) in the mixed-precision mode; the
// arithmetic is FP32 either way.
__device__ inline float ldA(const float* __restrict__ A, long long i) { return A[i]; }
__device__ inline float ldA(const uint16_t* __restrict__ A, long long i) {
	return __uint_as_float((unsigned)A[i] << 16);
}
__host__ __device__ inline uint16_t to_bf16(float f) {
	unsigned u;
	memcpy(&u, &f, 4);
	u += 0x7FFFu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
	const unsigned u = (unsigned)h << 16;
	float f;
	memcpy(&f, &u, 4);
	return f;
}
__global__ void k_to_bf16(const float* __restrict__ in, uint16_t* __restrict__ out, long long n) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
		out[i] = to_bf16(in[i]);
}

// Three consecutive floats at a 4-byte aligned address, read by one load.
struct __attribute__((packed, aligned(4))) FTriple {
	float a, b, c;
};

// b_c[V] = sum_{U in [2V-1,2V+1]} w(U) r_f[U]   (r_f is 0 on inactive vertices).
// Fixed 3x3x3 (3x3) trip counts with predicates: the loops unroll and all
// fine loads are in flight together; skipped terms are exact zeros.  The
// three fine values of a row (x = 2i-1, 2i, 2i+1) come from one 12-byte load
// at x0 = clamp(2i-1, 0, nx_f-3); only the first and last coarse vertex of a
// row have 2i-1 != x0 and pick their values at a shifted slot.  Fine levels
// narrower than 3 vertices in x take one load per value.
template <bool IS3D>
__device__ __forceinline__ void body_restrict(const RestrictP& P, long long blk, long long) {
	const CgState* __restrict__ st = P.st;
	const LvDev& c = P.c;
	const LvDev& f = P.f;
	const uint8_t* __restrict__ ctype = P.ctype;
	const float* __restrict__ rf = P.rf;
	float* __restrict__ bc = P.bc;
	if (st->done) return;
	const VtxPos p = lc_vertex(c, blk);
	if (!p.in) return;
	const int i = p.i, j = p.j, k = p.k, v = p.v;
	float sum = 0.f;
	if (f.nx >= 3) {
		constexpr int NZ = IS3D ? 3 : 1;
		const int x0 = min(max(2 * i - 1, 0), f.nx - 3);
		const int d = 2 * i - 1 - x0;  // -1 (i = 0), 0, 1 or 2 (last vertex of a row)
		long long ay[3], az[3];
		bool kx[3], ky[3], kz[3];
#pragma unroll
		for (int e = 0; e < 3; e++) {
			const int ux = 2 * i + e - 1, uy = 2 * j + e - 1, uz = 2 * k + (IS3D ? e - 1 : 0);
			kx[e] = ux >= 0 && ux < f.nx;
			ky[e] = uy >= 0 && uy < f.ny;
			kz[e] = uz >= 0 && uz < f.nz;
			ay[e] = ky[e] ? f.sy * uy : 0;
			az[e] = kz[e] ? f.sz * uz : 0;
		}
		const uint8_t ct = ctype[v];
		FTriple row[NZ * 3];
#pragma unroll
		for (int r = 0; r < NZ * 3; r++) row[r] = *reinterpret_cast<const FTriple*>(rf + x0 + ay[r % 3] + az[IS3D ? r / 3 : 1]);
#pragma unroll
		for (int e = 0; e < NZ * 9; e++) {
			const int ex = e % 3, ey = (e / 3) % 3, ez = IS3D ? e / 9 : 1;
			const FTriple& t = row[ey + 3 * (IS3D ? ez : 0)];
			// slot of x = 2i - 1 + ex in the triple: d + ex (0..2 whenever kx[ex])
			const int s = d + ex;
			const float val = s <= 0 ? t.a : (s == 1 ? t.b : t.c);
			const int ux = 2 * i + ex - 1, uy = 2 * j + ey - 1, uz = 2 * k + (IS3D ? ez - 1 : 0);
			if (kx[ex] && ky[ey] && kz[ez]) sum += wgt(ux, uy, uz) * val;
		}
		bc[v] = ct ? sum : 0.f;
		return;
	}
	if (ctype[v]) {
		// branch-free: per-axis offsets of the fine coordinates 2V-1..2V+1
		// (0 where outside the level, so every address is valid), every load
		// issued before the sum; terms outside the level add exactly +0
		constexpr int NZ = IS3D ? 3 : 1;
		long long ax[3], ay[3], az[3];
		bool kx[3], ky[3], kz[3];
#pragma unroll
		for (int d = 0; d < 3; d++) {
			const int ux = 2 * i + d - 1, uy = 2 * j + d - 1, uz = 2 * k + (IS3D ? d - 1 : 0);
			kx[d] = ux >= 0 && ux < f.nx;
			ky[d] = uy >= 0 && uy < f.ny;
			kz[d] = uz >= 0 && uz < f.nz;
			ax[d] = kx[d] ? ux : 0;
			ay[d] = ky[d] ? f.sy * uy : 0;
			az[d] = kz[d] ? f.sz * uz : 0;
		}
		float rv[NZ * 9];
		bool ok[NZ * 9];
#pragma unroll
		for (int e = 0; e < NZ * 9; e++) {
			const int ex = e % 3, ey = (e / 3) % 3, ez = IS3D ? e / 9 : 1;
			ok[e] = kx[ex] && ky[ey] && kz[ez];
			rv[e] = rf[ax[ex] + ay[ey] + az[ez]];
		}
#pragma unroll
		for (int e = 0; e < NZ * 9; e++) {
			const int dx = e % 3 - 1, dy = (e / 3) % 3 - 1, dz = IS3D ? e / 9 - 1 : 0;
			const int ux = 2 * i + dx, uy = 2 * j + dy, uz = 2 * k + dz;
			if (ok[e]) sum += wgt(ux, uy, uz) * rv[e];
		}
	}
	bc[v] = sum;
}

// r = b - A x on a coarse level (0 on inactive vertices)
template <bool IS3D, typename AT>
__device__ __forceinline__ void body_resid(const ResidP& P, long long blk, long long) {
	const CgState* __restrict__ st = P.st;
	const LvDev& c = P.c;
	const AT* __restrict__ A = (const AT*)P.A;
	const uint8_t* __restrict__ type = P.type;
	const float* __restrict__ b = P.b;
	const float* __restrict__ x = P.x;
	float* __restrict__ r = P.r;
	const Canon& cn = P.cn;
	constexpr int C = IS3D ? 13 : 4;
	if (st->done) return;
	const VtxPos p = lc_vertex(c, blk);
	if (!p.in) return;
	const int i = p.i, j = p.j, k = p.k, v = p.v;
	const int sy = (int)c.sy, sz = (int)c.sz;
	const uint8_t tv = type[v];
	if (!tv) {
		r[v] = 0.f;
		return;
	}
	float sum = b[v];
	if (tv == 2) {
		// canonical stencil, every neighbour inside and active: one fused
		// multiply-add chain per z layer; the three x-neighbours of a row
		// (v-1, v, v+1 shifted by the row offset) come from one 12-byte load
		const float* __restrict__ xv = x + v - 1;
		FTriple xr[IS3D ? 9 : 3];
#pragma unroll
		for (int r = 0; r < (IS3D ? 9 : 3); r++)
			xr[r] = *reinterpret_cast<const FTriple*>(xv + sy * (r % 3 - 1) + sz * (IS3D ? r / 3 - 1 : 0));
		float acc[IS3D ? 3 : 1];
#pragma unroll
		for (int lz = 0; lz < (IS3D ? 3 : 1); lz++) {
			float sacc = 0.f;
#pragma unroll
			for (int dy = -1; dy <= 1; dy++) {
				const FTriple& t = xr[(dy + 1) + 3 * lz];
				sacc = fmaf(cn.a[3 * (dy + 1) + 9 * lz], t.a, sacc);
				sacc = fmaf(cn.a[1 + 3 * (dy + 1) + 9 * lz], t.b, sacc);
				sacc = fmaf(cn.a[2 + 3 * (dy + 1) + 9 * lz], t.c, sacc);
			}
			acc[lz] = sacc;
		}
		r[v] = sum - (IS3D ? (acc[0] + acc[1]) + acc[IS3D ? 2 : 0] : acc[0]);
		return;
	}
	// general stencil, branch-free: neighbours outside the level or inactive
	// contribute an exact +0 (every load is issued before the sum)
	constexpr int NS = IS3D ? 27 : 9;
	float av[NS], xv[NS];
	bool ok[NS];
#pragma unroll
	for (int s = 0; s < NS; s++) {
		const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = IS3D ? s / 9 - 1 : 0;
		const int ni = i + dx, nj = j + dy, nk = k + dz;
		const bool in = ni >= 0 && nj >= 0 && nk >= 0 && ni < c.nx && nj < c.ny && nk < c.nz;
		const int nn = in ? ni + sy * nj + sz * nk : v;
		const uint8_t tn = type[nn];  // unconditional: nn is always a valid vertex
		ok[s] = in & (tn != 0);
		av[s] = s >= C ? ldA(A, (long long)(s - C) * c.n + v) : ldA(A, (long long)(C - s) * c.n + nn);
		xv[s] = x[nn];
	}
#pragma unroll
	for (int s = 0; s < NS; s++)
		if (ok[s]) sum -= av[s] * xv[s];
	r[v] = sum;
}

// ---------------------------------------------------------------------------
// LDS-brick kernels.  A workgroup owns a brick of cells/vertices and stages
// the brick plus a halo in LDS, so that a whole smoothing sweep (all
// colours) is one launch.
// ---------------------------------------------------------------------------
template <int BX, int BY, int BZ>
struct Brick {
	static constexpr int H = 2;
	static constexpr int RX = BX + 2 * H, RY = BY + 2 * H, RZ = BZ > 1 ? BZ + 2 * H : 1;
	static constexpr int RN = RX * RY * RZ;
	static constexpr int HZ = BZ > 1 ? H : 0;
};

__device__ inline void brick_origin(long long blk, int nx, int ny, int BX, int BY, int BZ, int& ox, int& oy,
                                    int& oz) {
	const unsigned nbx = (unsigned)(nx + BX - 1) / BX, nby = (unsigned)(ny + BY - 1) / BY;
	const unsigned b = (unsigned)blk, rest = b / nbx, kz = rest / nby;
	ox = (int)(b - rest * nbx) * BX;
	oy = (int)(rest - kz * nby) * BY;
	oz = (int)kz * BZ;
}

// sum of x over the fluid neighbours coupled to cell q (coefficient -1 each),
// for a cell in the LDS region with code cs and region strides 1/RX/RXY
template <bool IS3D>
__device__ inline float l0_nb_sum(const uint8_t* cs, const float* xs, int q, int RX, int RXY) {
	const uint8_t cc = cs[q];
	float sum = 0.f;
	if (cs[q - 1] & 8) sum += xs[q - 1];
	if (cc & 8) sum += xs[q + 1];
	if (cs[q - RX] & 16) sum += xs[q - RX];
	if (cc & 16) sum += xs[q + RX];
	if (IS3D) {
		if (cs[q - RXY] & 32) sum += xs[q - RXY];
		if (cc & 32) sum += xs[q + RXY];
	}
	return sum;
}

// Level-0 down leg: x = 0, red sweep, black sweep (exact red-black
// Gauss-Seidel: the black halo ring is recomputed from a 2-cell halo of red
// values), then the residual r0 = b - A x of the owned cells.
// Reads b (4 B) and the matrix code (1 B), writes x and r0: 13 B/cell.
template <int BX, int BY, int BZ, bool IS3D>
__device__ __forceinline__ void body_l0_down_brick(const L0DownP& P, long long blk, long long) {
	const Dims& d = P.d;
	const CgState* __restrict__ st = P.st;
	const uint8_t* __restrict__ code = P.code;
	const float* __restrict__ rin = P.rin;
	float* __restrict__ z = P.xout;
	float* __restrict__ r0 = P.r0;
	using B = Brick<BX, BY, BZ>;
	__shared__ float xs[B::RN];
	__shared__ float bs[B::RN];
	__shared__ uint8_t cs[B::RN];
	if (st->done) return;
	int ox, oy, oz;
	brick_origin(blk, d.nx, d.ny, BX, BY, BZ, ox, oy, oz);
	ox -= B::H;
	oy -= B::H;
	oz -= B::HZ;
	constexpr int RXY = B::RX * B::RY;
	for (int q = threadIdx.x; q < B::RN; q += blockDim.x) {
		const int gx = ox + q % B::RX, gy = oy + (q / B::RX) % B::RY, gz = oz + q / RXY;
		uint8_t cc = 0;
		float b = 0.f;
		if (gx >= 0 && gy >= 0 && gz >= 0 && gx < d.nx && gy < d.ny && gz < d.nz) {
			const long long g = d.idx(gx, gy, gz);
			cc = code[g];
			b = rin[g] * l0_scale(cc);
		}
		cs[q] = cc;
		bs[q] = b;
		xs[q] = (l0_active(cc) && ((gx + gy + gz) & 1) == 0) ? b / l0_diag(cc) : 0.f;
	}
	__syncthreads();
	// black cells of the region without its outermost ring
	constexpr int IX = B::RX - 2, IY = B::RY - 2, IZ = IS3D ? B::RZ - 2 : 1;
	for (int t = threadIdx.x; t < IX * IY * IZ; t += blockDim.x) {
		const int lx = 1 + t % IX, ly = 1 + (t / IX) % IY, lz = (IS3D ? 1 : 0) + t / (IX * IY);
		const int q = lx + B::RX * ly + RXY * lz;
		const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
		const uint8_t cc = cs[q];
		if (((gx + gy + gz) & 1) == 1 && l0_active(cc))
			xs[q] = (bs[q] + l0_nb_sum<IS3D>(cs, xs, q, B::RX, RXY)) / l0_diag(cc);
	}
	__syncthreads();
	for (int t = threadIdx.x; t < BX * BY * BZ; t += blockDim.x) {
		const int lx = B::H + t % BX, ly = B::H + (t / BX) % BY, lz = B::HZ + t / (BX * BY);
		const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
		if (gx >= d.nx || gy >= d.ny || gz >= d.nz) continue;
		const int q = lx + B::RX * ly + RXY * lz;
		const long long g = d.idx(gx, gy, gz);
		const uint8_t cc = cs[q];
		const float x = xs[q];
		z[g] = x;
		r0[g] = l0_active(cc) ? (bs[q] + l0_nb_sum<IS3D>(cs, xs, q, B::RX, RXY)) - l0_diag(cc) * x : 0.f;
	}
}

// Level-0 up leg: x = xpre + I x_1 (prolongation on load), black sweep, red
// sweep (reversed colour order), write z, and the partial sums of fl(z*r) for
// the CG scalar sigma = dot(z, r).  Reads xpre, b, code (9 B), writes z (4 B).
// Input and output are different arrays: the halo of one brick is owned by
// its neighbours, which write their result concurrently.
template <int BX, int BY, int BZ, bool IS3D>
__device__ __forceinline__ void body_l0_up_brick(const L0UpP& P, long long blk, long long slot) {
	const Dims& d = P.d;
	const CgState* __restrict__ st = P.st;
	const uint8_t* __restrict__ code = P.code;
	const float* __restrict__ rin = P.rin;
	const float* __restrict__ xpre = P.xpre;
	float* __restrict__ z = P.z;
	const LvDev& c1 = P.c1;
	const float* __restrict__ xc = P.xc;
	double* __restrict__ part = P.part;
	using B = Brick<BX, BY, BZ>;
	// coarse values covering the region's parents (inactive coarse vertices
	// always hold 0, so no type lookups are needed)
	constexpr int CRX = B::RX / 2 + 2, CRY = B::RY / 2 + 2, CRZ = IS3D ? B::RZ / 2 + 2 : 1;
	__shared__ float xs[B::RN];
	__shared__ float bs[B::RN];
	__shared__ uint8_t cs[B::RN];
	__shared__ float xcs[CRX * CRY * CRZ];
	__shared__ double sh[16];
	if (st->done) return;
	int ox, oy, oz;
	brick_origin(blk, d.nx, d.ny, BX, BY, BZ, ox, oy, oz);
	ox -= B::H;
	oy -= B::H;
	oz -= B::HZ;
	const int cx0 = ox >> 1, cy0 = oy >> 1, cz0 = oz >> 1;  // floor division
	for (int q = threadIdx.x; q < CRX * CRY * CRZ; q += blockDim.x) {
		const int px = cx0 + q % CRX, py = cy0 + (q / CRX) % CRY, pz = cz0 + q / (CRX * CRY);
		float v = 0.f;
		if (px >= 0 && py >= 0 && pz >= 0 && px < c1.nx && py < c1.ny && pz < c1.nz) v = xc[px + c1.sy * py + c1.sz * pz];
		xcs[q] = v;
	}
	__syncthreads();
	constexpr int RXY = B::RX * B::RY;
	for (int q = threadIdx.x; q < B::RN; q += blockDim.x) {
		const int gx = ox + q % B::RX, gy = oy + (q / B::RX) % B::RY, gz = oz + q / RXY;
		uint8_t cc = 0;
		float b = 0.f, x = 0.f;
		if (gx >= 0 && gy >= 0 && gz >= 0 && gx < d.nx && gy < d.ny && gz < d.nz) {
			const long long g = d.idx(gx, gy, gz);
			cc = code[g];
			b = rin[g] * l0_scale(cc);
			x = xpre[g];
			if (l0_active(cc)) {
				float sum = 0.f;
				for (int pz = gz >> 1; pz <= (gz + 1) >> 1; pz++)
					for (int py = gy >> 1; py <= (gy + 1) >> 1; py++)
						for (int px = gx >> 1; px <= (gx + 1) >> 1; px++)
							sum += xcs[(px - cx0) + CRX * (py - cy0) + CRX * CRY * (pz - cz0)];
				x += wgt(gx, gy, gz) * sum;
			}
		}
		cs[q] = cc;
		bs[q] = b;
		xs[q] = x;
	}
	__syncthreads();
	constexpr int IX = B::RX - 2, IY = B::RY - 2, IZ = IS3D ? B::RZ - 2 : 1;
	for (int t = threadIdx.x; t < IX * IY * IZ; t += blockDim.x) {
		const int lx = 1 + t % IX, ly = 1 + (t / IX) % IY, lz = (IS3D ? 1 : 0) + t / (IX * IY);
		const int q = lx + B::RX * ly + RXY * lz;
		const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
		const uint8_t cc = cs[q];
		if (((gx + gy + gz) & 1) == 1 && l0_active(cc))
			xs[q] = (bs[q] + l0_nb_sum<IS3D>(cs, xs, q, B::RX, RXY)) / l0_diag(cc);
	}
	__syncthreads();
	double acc = 0.0;
	for (int t = threadIdx.x; t < BX * BY * BZ; t += blockDim.x) {
		const int lx = B::H + t % BX, ly = B::H + (t / BX) % BY, lz = B::HZ + t / (BX * BY);
		const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
		if (gx >= d.nx || gy >= d.ny || gz >= d.nz) continue;
		const int q = lx + B::RX * ly + RXY * lz;
		const long long g = d.idx(gx, gy, gz);
		const uint8_t cc = cs[q];
		float x = xs[q];
		if (((gx + gy + gz) & 1) == 0 && l0_active(cc)) x = (bs[q] + l0_nb_sum<IS3D>(cs, xs, q, B::RX, RXY)) / l0_diag(cc);
		z[g] = x;
		acc += (double)(x * rin[g]);
	}
	const double s = block_sum(acc, sh);
	if (threadIdx.x == 0) part[3 * slot] = s;
}

// ---------------------------------------------------------------------------
// z-marching level-0 kernels (3D).  A workgroup owns a 64 x TY column tile
// and a chunk of planes [k0, k1).  It walks the planes once, staging each
// plane of the tile plus a 2-cell ring in a 6-slot LDS ring buffer, and runs
// the three dependent stages of red-black Gauss-Seidel as a pipeline whose
// stages are two planes apart:
//   A(p)   load plane p (code, rhs, x), first colour pointwise
//   B(p-2) second colour on plane p-2 (tile + 1-cell ring): needs the first
//          colour of planes p-3 .. p-1
//   C(p-4) final output on the tile's cells of plane p-4: needs both colours
//          of planes p-5 .. p-3
// Each stage reads only what the previous iterations finished, and the three
// touch different ring slots, so one barrier per plane suffices.  Every input
// element is read from HBM once plus 4 halo planes per chunk.  Arithmetic
// and operation order equal the brick kernels above.
// ---------------------------------------------------------------------------
constexpr int ZRING = 6;  // LDS ring slots: planes p-5 .. p
template <int TY>
struct ZTile {
	static constexpr int TX = 64, H = 2, RX = TX + 2 * H, RY = TY + 2 * H, RP = RX * RY, NS = ZRING, NT = TX * TY;
};

__device__ inline void zm_origin(long long blk, const Dims& d, int TY, int kchunk, int& ox, int& oy, int& k0,
                                 int& k1) {
	const unsigned ntx = (unsigned)(d.nx + 63) / 64, nty = (unsigned)(d.ny + TY - 1) / TY;
	const unsigned b = (unsigned)blk, rest = b / ntx, kz = rest / nty;
	ox = (int)(b - rest * ntx) * 64;
	oy = (int)(rest - kz * nty) * TY;
	k0 = (int)kz * kchunk;
	k1 = min(d.nz, k0 + kchunk);
}

// x if bit B of c is set, else +0.0: the bit pattern AND the bit's sign
// extension (v_bfe_i32 + v_and_b32 instead of a compare and a select; the
// extract is written as an instruction because the compiler turns a
// sign-extended bit of a byte back into a compare and a select)
template <int B>
__device__ inline float zm_keep(uint8_t c, float x) {
	int m;
	asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"((int)c), "i"(B));
	return __int_as_float(__float_as_int(x) & m);
}
// l0_nb_sum with the plane below (cM, xM) and above (xU) in separate slots;
// cc = cP[q].  Every neighbour is read and added, masked to +0.0 where its
// coupling flag is clear.  The running sum starts at +0.0 and a round-to-
// nearest sum is -0.0 only when both addends are, so it never is; adding
// +0.0 to it is therefore exact and this equals the conditional adds of
// l0_nb_sum, in the same order, without a branch per neighbour.
__device__ inline float zm_nb_sum(uint8_t cc, const uint8_t* cP, const float* xP, const uint8_t* cM, const float* xM,
                                  const float* xU, int q, int RX) {
	const uint8_t cw = cP[q - 1], cs = cP[q - RX], cm = cM[q];
	const float w = xP[q - 1], e = xP[q + 1], so = xP[q - RX], no = xP[q + RX], dn = xM[q], up = xU[q];
	float sum = 0.f;
	sum += zm_keep<3>(cw, w);
	sum += zm_keep<3>(cc, e);
	sum += zm_keep<4>(cs, so);
	sum += zm_keep<4>(cc, no);
	sum += zm_keep<5>(cm, dn);
	sum += zm_keep<5>(cc, up);
	return sum;
}

// (diag, 1/diag, scale) of the level-0 row codes in LDS, index (c & 7) + 8
// for the 1e-6-scaled trivial rows (bit 7); filled by threads 0..15 before
// the kernel's first barrier (scale = l0_scale, diag = l0_diag)
__device__ inline int l0_tab_index(uint8_t c) { return (c & 7) | ((c >> 4) & 8); }
__device__ inline void l0_tab_fill(float4* t) {
	if (threadIdx.x < 16) {
		const int a = threadIdx.x & 7;
		const bool tr = (threadIdx.x & 8) != 0;
		const float dg = tr ? (float)a * 1e-6f : (float)a;
		t[threadIdx.x] = make_float4(dg, a ? 1.0f / dg : 0.f, tr ? 1e-6f : 1.0f, 0.f);
	}
}
__device__ inline float4 l0_row(uint8_t c, const float4* t) { return t[l0_tab_index(c)]; }
// v / diag as q0 = v (1/d) plus one residual correction (Markstein):
// checked exhaustively over all binary32 v, it equals the IEEE quotient for
// every finite v when d = 1..5 and for every normal quotient when d = 6; on
// the 1e-6-scaled trivial rows it can differ from it in the last bit
__device__ inline float l0_div_r(float v, const float4& e) {
	const float q0 = v * e.y;
	return fmaf(fmaf(-q0, e.x, v), e.y, q0);
}


// The (at most two) region cells of one thread: q = tid and tid + NT.
template <typename T>
struct ZmCells {
	static_assert(T::RP <= 2 * T::NT, "two region cells per thread");
	unsigned g[2];   // in-plane offset i + nx*j (nx*ny < 2^31 is checked on the host)
	bool ok[2];      // inside the region and the grid
	bool par[2];     // (i + j) & 1
	__device__ ZmCells(const Dims& d, int rx0, int ry0) {
#pragma unroll
		for (int u = 0; u < 2; u++) {
			const int q = threadIdx.x + u * T::NT;
			const int gx = rx0 + q % T::RX, gy = ry0 + q / T::RX;
			ok[u] = q < T::RP && gx >= 0 && gy >= 0 && gx < d.nx && gy < d.ny;
			g[u] = ok[u] ? (unsigned)gx + (unsigned)d.nx * (unsigned)gy : 0u;
			par[u] = ((gx + gy) & 1) != 0;
		}
	}
};

// x / diag of a level-0 row (A0 = 1..6, or A0 * 1e-6 on trivial rows)
__device__ inline float l0_div(float v, uint8_t c) { return v / l0_diag(c); }

// Per-thread positions of the z-marching kernels that do not change from
// plane to plane (the plane loop only rotates ring slots).
template <typename T>
struct ZmPos {
	// load stage: at most two region cells per thread (q = tid, tid + NT)
	int qa[2];
	// second-colour stage: region without its outer ring (at most two cells)
	int qb[2];
	bool parb[2];
	bool inb[2];
	// output stage: the thread's tile cell
	int qc;
	unsigned gc;
	bool parc;
	bool inc;
	__device__ ZmPos(const Dims& d, int ox, int oy) {
		const int rx0 = ox - T::H, ry0 = oy - T::H;
		constexpr int IX = T::RX - 2, IN = IX * (T::RY - 2);
		static_assert(IN <= 2 * T::NT, "two second-colour cells per thread");
#pragma unroll
		for (int u = 0; u < 2; u++) {
			qa[u] = threadIdx.x + u * T::NT;
			const int t = threadIdx.x + u * T::NT;
			const int lx = 1 + t % IX, ly = 1 + t / IX;
			inb[u] = t < IN;
			qb[u] = lx + T::RX * ly;
			parb[u] = ((rx0 + lx + ry0 + ly) & 1) != 0;
		}
		const int lx = T::H + (threadIdx.x & 63), ly = T::H + (threadIdx.x >> 6);
		const int gx = rx0 + lx, gy = ry0 + ly;
		qc = lx + T::RX * ly;
		inc = gx < d.nx && gy < d.ny;
		gc = inc ? (unsigned)gx + (unsigned)d.nx * (unsigned)gy : 0u;
		parc = ((gx + gy) & 1) != 0;
	}
};

// Colour-packed positions of the z-marching stages.  Tile origins are even,
// so the colour of a cell is the parity of its tile-local x + y plus the
// plane.  Stage B updates the black cells of the region without its outer
// ring: IX/2 per row, one per thread for threads < BN (the other wavefronts
// skip the stage).
template <typename T>
struct ZmPack {
	static constexpr int IX = T::RX - 2, IY = T::RY - 2, BN = (IX / 2) * IY;
	static_assert(IX % 2 == 0 && BN <= T::NT, "colour packing");
	int bly, bm;  // stage B: region row (1-based) and index within the row
	__device__ ZmPack() {
		const int t = threadIdx.x;
		bly = 1 + t / (IX / 2);
		bm = t % (IX / 2);
	}
	// LDS index of the thread's stage-B black cell at plane parity par:
	// black means region-local x + y + plane odd
	__device__ int qb(int par) const { return 2 * bm + 1 + ((bly + par) & 1) + T::RX * bly; }
};

__device__ inline int ring_next(int s) { return s == ZRING - 1 ? 0 : s + 1; }
__device__ inline int ring_back(int s, int k) { return (s + ZRING - k) % ZRING; }

// Down leg (same sweeps as body_l0_down_brick): red from x = 0, black, residual.
// The stages are branch-free where lanes would diverge: plane loads read a
// clamped plane and are zeroed where used, colour and activity select
// between the new and the old value, and the division goes through the
// (diag, 1/diag) table.
template <int TY>
__device__ __forceinline__ void body_l0_down_zm(const L0DownP& P, long long blk, long long) {
	const Dims& d = P.d;
	const int kchunk = P.kchunk;
	const CgState* __restrict__ st = P.st;
	const uint8_t* __restrict__ code = P.code;
	const float* __restrict__ rin = P.rin;
	float* __restrict__ xout = P.xout;
	float* __restrict__ r0 = P.r0;
	using T = ZTile<TY>;
	// one LDS block in a fixed order: codes first (byte offsets from the
	// cell index), then the float rings contiguous, so every ring access is
	// an immediate offset (< 64 KiB) from one of two per-cell bases
	struct Lds {
		uint8_t cs[T::NS][T::RP];
		float xs[T::NS][T::RP];
		float bs[T::NS][T::RP];
		float4 dt[16];
	};
	static_assert((T::NS * T::RP) % 16 == 0 && 2 * sizeof(float) * T::NS * T::RP < 65536, "ring layout");
	__shared__ Lds lds;
	auto& cs = lds.cs;
	auto& xs = lds.xs;
	auto& bs = lds.bs;
	float4* dt = lds.dt;
	if (st->done) return;
	l0_tab_fill(dt);
	int ox, oy, k0, k1;
	zm_origin(blk, d, TY, kchunk, ox, oy, k0, k1);
	const int rx0 = ox - T::H, ry0 = oy - T::H;
	ZmCells<T> cells(d, rx0, ry0);
	const ZmPos<T> pos(d, ox, oy);
	const ZmPack<T> pk;
	// plane p's inputs are fetched into registers one plane ahead, so the
	// loads are in flight during the barriers and the sweeps of plane p-1;
	// a plane outside k0-2..k1+1 or the grid reads a clamped plane and pv
	// marks whether the values count
	const int pLo = max(k0 - 2, 0), pHi = min(k1 + 1, d.nz - 1);
	uint8_t pc[2];
	float pr[2];
	bool pv = false;
	auto fetch = [&](int p) {
		const long long pl = d.sz * min(max(p, pLo), pHi);
#pragma unroll
		for (int u = 0; u < 2; u++) {
			pc[u] = code[cells.g[u] + pl];
			pr[u] = rin[cells.g[u] + pl];
		}
		pv = p >= 0 && p < d.nz;
	};
	fetch(k0 - 2);
	__syncthreads();
	// One plane step; J = ring slot of plane p = (p - k0 + 2) mod ZRING, a
	// compile-time constant (the loop is unrolled by ZRING), so every slot
	// offset is an immediate of the LDS instructions.  k0 is even, so the
	// parity of p is that of J.
	auto step = [&](auto jc, int p) {
		constexpr int J = decltype(jc)::value;
		constexpr int sA = J, sB0 = (J + 4) % ZRING, sBm = (J + 3) % ZRING, sBu = (J + 5) % ZRING;
		constexpr int sC0 = (J + 2) % ZRING, sCm = (J + 1) % ZRING, sCu = (J + 3) % ZRING;
		constexpr int par = J & 1;
		// plane p into slot sA (after p = k1+1 the slot is not read again)
		{
			const bool pvp = pv;
#pragma unroll
			for (int u = 0; u < 2; u++) {
				const int q = pos.qa[u];
				if (q >= T::RP) continue;
				const bool ok = cells.ok[u] && pvp;
				const uint8_t cc = ok ? pc[u] : (uint8_t)0;
				const float4 e = l0_row(cc, dt);
				const float b = ok ? pr[u] * e.z : 0.f;
				cs[sA][q] = cc;
				bs[sA][q] = b;
				const float x = l0_div_r(b, e);
				xs[sA][q] = (l0_active(cc) && cells.par[u] == (par != 0)) ? x : 0.f;
			}
			fetch(p + 1);
		}
		const int pb = p - 2;
		if (pb >= k0 - 1 && pb <= k1 && pb >= 0 && pb < d.nz && threadIdx.x < ZmPack<T>::BN) {
			const int q = pk.qb(par);
			const uint8_t cc = cs[sB0][q];
			const float xo = xs[sB0][q];
			const float v =
			    l0_div_r(bs[sB0][q] + zm_nb_sum(cc, cs[sB0], xs[sB0], cs[sBm], xs[sBm], xs[sBu], q, T::RX), l0_row(cc, dt));
			xs[sB0][q] = l0_active(cc) ? v : xo;
		}
		const int pz = p - 4;
		if (pz >= k0 && pz < k1) {
			const int q = pos.qc;
			const uint8_t cc = cs[sC0][q];
			const float x = xs[sC0][q];
			const float rr = (bs[sC0][q] + zm_nb_sum(cc, cs[sC0], xs[sC0], cs[sCm], xs[sCm], xs[sCu], q, T::RX)) -
			                 l0_row(cc, dt).x * x;
			if (pos.inc) {
				const long long g = pos.gc + d.sz * pz;
				xout[g] = x;
				r0[g] = l0_active(cc) ? rr : 0.f;
			}
		}
		__syncthreads();
	};
	static_assert(ZRING == 6, "the plane loop is unrolled by the ring size");
	for (int p = k0 - 2; p <= k1 + 3; p += ZRING) {
		step(std::integral_constant<int, 0>{}, p);
		if (p + 1 <= k1 + 3) step(std::integral_constant<int, 1>{}, p + 1);
		if (p + 2 <= k1 + 3) step(std::integral_constant<int, 2>{}, p + 2);
		if (p + 3 <= k1 + 3) step(std::integral_constant<int, 3>{}, p + 3);
		if (p + 4 <= k1 + 3) step(std::integral_constant<int, 4>{}, p + 4);
		if (p + 5 <= k1 + 3) step(std::integral_constant<int, 5>{}, p + 5);
	}
}

// Up leg (same sweeps as body_l0_up_brick): x = xpre + I x_1, black, red,
// write z and the partial sums of fl(z*r).  Coarse planes are staged in a
// 2-slot ring: fine plane p needs coarse planes p>>1 and (p+1)>>1.  The x/y
// parents of the thread's region cells are fixed for the whole column, only
// the z parents change with the plane.  Branch-free in
// the same way as the down leg; the prolongation reads all four x/y parents
// of a coarse plane and adds each with a select.
template <int TY>
__device__ __forceinline__ void body_l0_up_zm(const L0UpP& P, long long blk, long long slot) {
	const Dims& d = P.d;
	const int kchunk = P.kchunk;
	const CgState* __restrict__ st = P.st;
	const uint8_t* __restrict__ code = P.code;
	const float* __restrict__ rin = P.rin;
	const float* __restrict__ xpre = P.xpre;
	float* __restrict__ z = P.z;
	const LvDev& c1 = P.c1;
	const float* __restrict__ xc = P.xc;
	double* __restrict__ part = P.part;
	using T = ZTile<TY>;
	constexpr int CRX = T::RX / 2 + 2, CRY = T::RY / 2 + 2, CRP = CRX * CRY;
	// LDS in a fixed order as in the down leg
	struct Lds {
		uint8_t cs[T::NS][T::RP];
		float xs[T::NS][T::RP];
		float rs[T::NS][T::RP];
		float xcs[2][CRP];
		double tab[kZmMaxGroups][TY];
		float4 dt[16];
	};
	static_assert((T::NS * T::RP) % 16 == 0 && 2 * sizeof(float) * T::NS * T::RP < 65536, "ring layout");
	__shared__ Lds lds;
	auto& cs = lds.cs;
	auto& xs = lds.xs;
	auto& rs = lds.rs;
	auto& xcs = lds.xcs;
	auto& tab = lds.tab;
	float4* dt = lds.dt;
	if (st->done) return;
	l0_tab_fill(dt);
	int ox, oy, k0, k1;
	zm_origin(blk, d, TY, kchunk, ox, oy, k0, k1);
	const int rx0 = ox - T::H, ry0 = oy - T::H;
	const int cx0 = rx0 >> 1, cy0 = ry0 >> 1;  // floor division
	static_assert(CRP <= T::NT, "one coarse value per thread");
	// this thread's vertex of the coarse-plane region; coarse planes outside
	// the level read a clamped plane and are zeroed where stored
	const int cq = threadIdx.x;
	const int cpx = cx0 + cq % CRX, cpy = cy0 + cq / CRX;
	const bool cok = cq < CRP && cpx >= 0 && cpy >= 0 && cpx < c1.nx && cpy < c1.ny;
	const unsigned cg = cok ? (unsigned)cpx + (unsigned)c1.sy * (unsigned)cpy : 0u;  // coarse n < 2^31
	auto cfetch = [&](int Pz) -> float { return xc[cg + (unsigned)c1.sz * (unsigned)min(max(Pz, 0), c1.nz - 1)]; };
	auto cvalid = [&](int Pz) -> bool { return cok && Pz >= 0 && Pz < c1.nz; };
	ZmCells<T> cells(d, rx0, ry0);
	const ZmPos<T> pos(d, ox, oy);
	const ZmPack<T> pk;
	// prolongation offsets (coarse-plane region) and x/y parities per region cell
	int po[2];
	bool pxo[2], pyo[2];
#pragma unroll
	for (int u = 0; u < 2; u++) {
		const int q = pos.qa[u];
		const int gx = rx0 + q % T::RX, gy = ry0 + q / T::RX;
		pxo[u] = gx & 1;
		pyo[u] = gy & 1;
		po[u] = ((gx >> 1) - cx0) + CRX * ((gy >> 1) - cy0);
	}
	const int pLo = max(k0 - 2, 0), pHi = min(k1 + 1, d.nz - 1);
	uint8_t pc[2];
	float pr[2], px[2];
	bool pv = false;
	auto fetch = [&](int p) {
		const long long pl = d.sz * min(max(p, pLo), pHi);
#pragma unroll
		for (int u = 0; u < 2; u++) {
			pc[u] = code[cells.g[u] + pl];
			pr[u] = rin[cells.g[u] + pl];
			px[u] = xpre[cells.g[u] + pl];
		}
		pv = p >= 0 && p < d.nz;
	};
	// coarse plane (k0-2)/2 goes to LDS now; later planes are fetched into a
	// register two fine planes before they are needed (k0 is even)
	if (cq < CRP) xcs[((k0 - 2) >> 1) & 1][cq] = cvalid((k0 - 2) >> 1) ? cfetch((k0 - 2) >> 1) : 0.f;
	float creg = cfetch(k0 >> 1);
	bool cregv = cvalid(k0 >> 1);
	fetch(k0 - 2);
	__syncthreads();
	double acc = 0.0;
	// One plane step, J = ring slot of plane p (compile-time, as in the down
	// leg; the parity of p is that of J)
	auto step = [&](auto jc, int p) {
		constexpr int J = decltype(jc)::value;
		constexpr int sA = J, sB0 = (J + 4) % ZRING, sBm = (J + 3) % ZRING, sBu = (J + 5) % ZRING;
		constexpr int sC0 = (J + 2) % ZRING, sCm = (J + 1) % ZRING, sCu = (J + 3) % ZRING;
		constexpr int par = J & 1;
		constexpr bool zo = (J & 1) != 0;
		// plane p into slot sA (after p = k1+1 the slot is not read again)
		{
			const float* c0 = xcs[(p >> 1) & 1];
			const float* c1p = xcs[((p >> 1) + 1) & 1];
			const bool pvp = pv;
#pragma unroll
			for (int u = 0; u < 2; u++) {
				const int q = pos.qa[u];
				if (q >= T::RP) continue;
				const bool ok = cells.ok[u] && pvp;
				const uint8_t cc = ok ? pc[u] : (uint8_t)0;
				const float rr = ok ? pr[u] : 0.f;
				float x = ok ? px[u] : 0.f;
				// parents in the reference's loop order: z, then y, then x
				const int o = po[u];
				const bool bx = pxo[u], by = pyo[u], bxy = pxo[u] && pyo[u];
				float sum = c0[o];
				const float a01 = c0[o + 1], a10 = c0[o + CRX], a11 = c0[o + CRX + 1];
				sum = bx ? sum + a01 : sum;
				sum = by ? sum + a10 : sum;
				sum = bxy ? sum + a11 : sum;
				if (zo) {
					const float b00 = c1p[o], b01 = c1p[o + 1], b10 = c1p[o + CRX], b11 = c1p[o + CRX + 1];
					sum += b00;
					sum = bx ? sum + b01 : sum;
					sum = by ? sum + b10 : sum;
					sum = bxy ? sum + b11 : sum;
				}
				// x/y weight 1/2 per odd coordinate (exact powers of two)
				const float pw = (bx ? 0.5f : 1.0f) * (by ? 0.5f : 1.0f);
				const float xn = x + (zo ? 0.5f * pw : pw) * sum;
				x = l0_active(cc) ? xn : x;
				cs[sA][q] = cc;
				rs[sA][q] = rr;
				xs[sA][q] = x;
			}
			if (!zo) {
				// coarse plane p/2+1 is first needed by plane p+1
				if (cq < CRP) xcs[((p >> 1) + 1) & 1][cq] = cregv ? creg : 0.f;
				creg = cfetch((p >> 1) + 2);
				cregv = cvalid((p >> 1) + 2);
			}
			fetch(p + 1);
		}
		const int pb = p - 2;
		// packed positions from the plane's run-time parity (equal to par;
		// recomputed per plane instead of holding both parities' positions
		// in registers across the loop)
		const int rpar = p & 1;
		if (pb >= k0 - 1 && pb <= k1 && pb >= 0 && pb < d.nz && threadIdx.x < ZmPack<T>::BN) {
			const int q = pk.qb(rpar);
			const uint8_t cc = cs[sB0][q];
			const float xo = xs[sB0][q];
			const float4 e = l0_row(cc, dt);
			const float v =
			    l0_div_r(rs[sB0][q] * e.z + zm_nb_sum(cc, cs[sB0], xs[sB0], cs[sBm], xs[sBm], xs[sBu], q, T::RX), e);
			xs[sB0][q] = l0_active(cc) ? v : xo;
		}
		const int pz = p - 4;
		if (pz >= k0 && pz < k1) {
			const int q = pos.qc;
			const uint8_t cc = cs[sC0][q];
			const float xo = xs[sC0][q];
			const float rq = rs[sC0][q];
			const float4 e = l0_row(cc, dt);
			const float v =
			    l0_div_r(rq * e.z + zm_nb_sum(cc, cs[sC0], xs[sC0], cs[sCm], xs[sCm], xs[sCu], q, T::RX), e);
			const float x = (pos.parc == (par != 0) && l0_active(cc)) ? v : xo;
			if (pos.inc) z[pos.gc + d.sz * pz] = x;
			const double prod = (double)(x * rq);
			acc = pos.inc ? acc + prod : acc;
			if ((pz + 1) % kDotGroup == 0 || pz + 1 == k1) zm_dot_flush<TY>(acc, tab, pz, k0);
		}
		__syncthreads();
	};
	for (int p = k0 - 2; p <= k1 + 3; p += ZRING) {
		step(std::integral_constant<int, 0>{}, p);
		if (p + 1 <= k1 + 3) step(std::integral_constant<int, 1>{}, p + 1);
		if (p + 2 <= k1 + 3) step(std::integral_constant<int, 2>{}, p + 2);
		if (p + 3 <= k1 + 3) step(std::integral_constant<int, 3>{}, p + 3);
		if (p + 4 <= k1 + 3) step(std::integral_constant<int, 4>{}, p + 4);
		if (p + 5 <= k1 + 3) step(std::integral_constant<int, 5>{}, p + 5);
	}
	(void)slot;
	zm_dot_store<TY>(tab, d, zm_column(blk, d, TY), k0, k1, part);
}

// coarse-level brick: 32x4x4 vertices in 3D, 64x8 in 2D, 512 threads (rows
// of 32/64 vertices keep the region loads in whole cache lines)
constexpr int kLcBX3 = 32, kLcBY3 = 4, kLcBZ3 = 4, kLcBX2 = 64, kLcBY2 = 8;
constexpr int kSmoothNT = 512;

// Coarse-level block Gauss-Seidel: one sweep over all 8 (4 in 2D) colours of
// the 27-point (9-point) Galerkin stencil inside an LDS brick; halo values
// are the pre-sweep ones (block-hybrid smoother).  Two modes (POST):
//   pre-smoothing:  x starts at 0, forward colour order;
//   post-smoothing: x = xin + I x_{l+1} (prolongation from the next coarser
//                   level on load; inactive coarse vertices hold 0 there, so
//                   no coarse type lookups are needed), reversed colour order,
// which keeps the V-cycle symmetric.  The result goes to xout, never to xin:
// neighbouring bricks read xin as their halo.
// Threads are assigned by colour: the 64 vertices of colour c are the lanes
// of wavefront c (in 2D, 128 vertices on wavefronts 2c and 2c+1), so each
// phase runs whole wavefronts.  Every thread loads its vertex's stencil once,
// before the sweep (canonical vertices take the uniform stencil), and the
// coarse parents of the brick are staged in LDS.  The stencil sum runs as
// three fused multiply-add chains, one per z layer.  Vertex indices are
// 32-bit (coarse levels hold fewer than 2^31 vertices).
template <bool IS3D, typename AT, bool POST>
__device__ __forceinline__ void body_smooth(const SmoothP& P, long long blk, long long) {
	const CgState* __restrict__ st = P.st;
	const LvDev& c = P.c;
	const AT* __restrict__ A = (const AT*)P.A;
	const uint8_t* __restrict__ type = P.type;
	const float* __restrict__ b = P.b;
	const float* __restrict__ xin = P.xin;
	float* __restrict__ xout = P.xout;
	const LvDev& cp = P.cp;
	const float* __restrict__ xcp = P.xcp;
	const Canon& cn = P.cn;
	constexpr int BX = IS3D ? kLcBX3 : kLcBX2, BY = IS3D ? kLcBY3 : kLcBY2, BZ = IS3D ? kLcBZ3 : 1;
	constexpr int RX = BX + 2, RY = BY + 2, RZ = IS3D ? BZ + 2 : 1, RN = RX * RY * RZ, RXY = RX * RY;
	constexpr int C = IS3D ? 13 : 4, NCOL = IS3D ? 8 : 4, NS = IS3D ? 27 : 9;
	// one colour's sub-lattice of the brick and the wavefronts that hold it
	constexpr int HX = BX / 2, HY = BY / 2, NWPC = kSmoothNT / 64 / NCOL;
	static_assert(BX * BY * BZ == kSmoothNT && HX * HY * (IS3D ? BZ / 2 : 1) == 64 * NWPC, "brick/thread mismatch");
	// coarse parents of the region: coarse coordinates o/2 - 1 .. (o + B)/2
	constexpr int CRX = BX / 2 + 2, CRY = BY / 2 + 2, CRZ = IS3D ? BZ / 2 + 2 : 1, CRXY = CRX * CRY;
	constexpr int CRN = CRXY * CRZ;
	__shared__ float xs[RN];
	__shared__ float xc[POST ? CRN : 1];
	if (st->done) return;
	int ox, oy, oz;
	brick_origin(blk, c.nx, c.ny, BX, BY, BZ, ox, oy, oz);
	const int t = threadIdx.x;
	const int hz = IS3D ? 1 : 0;
	const int sy = (int)c.sy, sz = (int)c.sz;
	// this thread's vertex
	const int col = (t >> 6) / NWPC, sidx = ((t >> 6) % NWPC) * 64 + (t & 63);
	const int tx = 2 * (sidx % HX) + (col & 1), ty = 2 * ((sidx / HX) % HY) + ((col >> 1) & 1);
	const int tz = IS3D ? 2 * (sidx / (HX * HY)) + (col >> 2) : 0;
	const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
	const bool valid = gx < c.nx && gy < c.ny && gz < c.nz;
	const int v = valid ? gx + sy * gy + sz * gz : 0;
	const uint8_t tv = valid ? type[v] : (uint8_t)0;
	const float bv = b[v];
	float a[NS];
	float diag = 1.f;
	if (tv == 2) {
		diag = cn.a[C];
#pragma unroll
		for (int s = 0; s < NS; s++) a[s] = s == C ? 0.f : cn.a[s];
	} else if (tv) {
		diag = ldA(A, v);
		int s = 0;
		for (int dz = (IS3D ? -1 : 0); dz <= (IS3D ? 1 : 0); dz++)
			for (int dy = -1; dy <= 1; dy++)
				for (int dx = -1; dx <= 1; dx++, s++) {
					float coef = 0.f;
					const int nx = gx + dx, ny = gy + dy, nz = gz + dz;
					if (s != C && nx >= 0 && ny >= 0 && nz >= 0 && nx < c.nx && ny < c.ny && nz < c.nz) {
						const int nn = nx + sy * ny + sz * nz;
						coef = s > C ? ldA(A, (long long)(s - C) * c.n + v) : ldA(A, (long long)(C - s) * c.n + nn);
					}
					a[s] = coef;
				}
	}
	if (POST) {
		const int psy = (int)cp.sy, psz = (int)cp.sz;
		const int cox = (ox >> 1) - 1, coy = (oy >> 1) - 1, coz = IS3D ? (oz >> 1) - 1 : 0;
		for (int q = t; q < CRN; q += kSmoothNT) {
			const int cx = q % CRX, cy = (q / CRX) % CRY, cz = q / CRXY;
			const int px = min(max(cox + cx, 0), cp.nx - 1), py = min(max(coy + cy, 0), cp.ny - 1);
			const int pz = IS3D ? min(max(coz + cz, 0), cp.nz - 1) : 0;
			xc[q] = xcp[px + psy * py + psz * pz];
		}
		__syncthreads();
	}
#pragma unroll
	for (int it = 0; it < (RN + kSmoothNT - 1) / kSmoothNT; it++) {
		const int q = t + kSmoothNT * it;
		if (q >= RN) break;
		const int rx = q % RX, ry = (q / RX) % RY, rz = q / RXY;
		const int hx = ox - 1 + rx, hy = oy - 1 + ry, hzz = oz - hz + rz;
		const bool inb = hx >= 0 && hy >= 0 && hzz >= 0 && hx < c.nx && hy < c.ny && hzz < c.nz;
		float x = 0.f;
		if (POST) {
			// branch-free: a valid address for every load, results selected
			const int g = inb ? hx + sy * hy + sz * hzz : 0;
			x = xin[g];
			const uint8_t tp = type[g];
			// parents [s/2, (s+1)/2] per axis; with o even, s = o - 1 + r has
			// parity (r + 1) & 1 and its first parent in LDS slot (r + 1) >> 1
			const int odx = (rx + 1) & 1, ody = (ry + 1) & 1, odz = IS3D ? (rz + 1) & 1 : 0;
			const int base = ((rx + 1) >> 1) + CRX * ((ry + 1) >> 1) + (IS3D ? CRXY * ((rz + 1) >> 1) : 0);
			const int ex = odx, ey = CRX * ody, ez = CRXY * odz;
			float pv[IS3D ? 8 : 4];
			pv[0] = xc[base];
			pv[1] = xc[base + ex];
			pv[2] = xc[base + ey];
			pv[3] = xc[base + ex + ey];
			if constexpr (IS3D) {
				pv[4] = xc[base + ez];
				pv[5] = xc[base + ex + ez];
				pv[6] = xc[base + ey + ez];
				pv[7] = xc[base + ex + ey + ez];
			}
			float sum = 0.f;
#pragma unroll
			for (int e = 0; e < (IS3D ? 8 : 4); e++) {
				const bool use = ((e & 1) == 0 || odx) && (((e >> 1) & 1) == 0 || ody) && ((e >> 2) == 0 || odz);
				if (use) sum += pv[e];
			}
			if (tp) x += wgt(hx, hy, hzz) * sum;
			if (!inb) x = 0.f;
		}
		xs[q] = x;
	}
	__syncthreads();
	const int me = (tx + 1) + RX * (ty + 1) + RXY * (tz + hz);
	for (int ph = 0; ph < NCOL; ph++) {
		if (tv && col == (POST ? NCOL - 1 - ph : ph)) {
			float acc[IS3D ? 3 : 1];
#pragma unroll
			for (int lz = 0; lz < (IS3D ? 3 : 1); lz++) {
				float sacc = 0.f;
#pragma unroll
				for (int dy = -1; dy <= 1; dy++)
#pragma unroll
					for (int dx = -1; dx <= 1; dx++) {
						const int s = (dx + 1) + 3 * (dy + 1) + 9 * lz;
						if (s != C) sacc = fmaf(a[s], xs[me + dx + RX * dy + RXY * (lz - hz)], sacc);
					}
				acc[lz] = sacc;
			}
			const float tot = IS3D ? (acc[0] + acc[1]) + acc[IS3D ? 2 : 0] : acc[0];
			xs[me] = (bv - tot) / diag;
		}
		__syncthreads();
	}
	if (valid) xout[v] = xs[me];
}

__device__ __forceinline__ void body_fill(const FillP& P, long long blk, long long) {
	if (P.st->done) return;
	const long long v = blk * 256 + threadIdx.x;
	if (v < P.n) P.x[v] = 0.f;
}

// coarsest level: x[cidx[i]] = sum_j G[i][j] * b[cidx[j]] (FP64), one wave per row;
// for a single-level hierarchy the rhs is the trivially scaled level-0 input
__device__ __forceinline__ void body_coarsest(const CoarsestP& P, long long blk, long long) {
	const CgState* __restrict__ st = P.st;
	const int nc = P.nc;
	const int* __restrict__ cidx = P.cidx;
	const double* __restrict__ G = P.G;
	const float* __restrict__ b = P.b;
	const uint8_t* __restrict__ code = P.code;
	float* __restrict__ x = P.x;
	if (st->done) return;
	const int lane = threadIdx.x & 63;
	const int row = (int)blk * 4 + (threadIdx.x >> 6);
	if (row >= nc) return;
	double acc = 0.0;
#pragma unroll 4
	for (int jj = lane; jj < nc; jj += 64) {
		const int v = cidx[jj];
		const double bv = code ? (double)(b[v] * l0_scale(code[v])) : (double)b[v];
		acc += G[(long long)row * nc + jj] * bv;
	}
	acc = wave_sum(acc);
	if (lane == 0) x[cidx[row]] = (float)acc;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
MgHierarchy::~MgHierarchy() {
	for (void* p : allocs) (void)hipFree(p);
}

void* MgHierarchy::alloc(size_t bytes) {
	void* p = nullptr;
	MFX_HIP(hipMalloc(&p, bytes < 16 ? 16 : bytes));
	allocs.push_back(p);
	return p;
}

static unsigned nblk(long long n) {
	long long b = (n + 255) / 256;
	if (b < 1) b = 1;
	if (b > 65535) b = 65535;
	return (unsigned)b;
}

void MgHierarchy::build(Scene* s, const uint8_t* code, int precision) {
	prec = precision;
	const Dims& d = s->d;
	is3d = d.is3d != 0;
	// the z-marching level-0 kernels keep in-plane offsets in 32 bits
	MFX_CHECK((long long)d.nx * d.ny < (1ll << 31), "multigrid: a grid plane has too many cells");
	S = is3d ? 14 : 5;
	L.clear();
	// level sizes (multigrid.cpp:242-262)
	MgLevel l0;
	l0.nx = d.nx;
	l0.ny = d.ny;
	l0.nz = d.nz;
	l0.n = d.n;
	l0.sy = d.nx;
	l0.sz = (long long)d.nx * d.ny;
	L.push_back(l0);
	long long n = d.n;
	for (int l = 1; l <= 100; l++) {
		const MgLevel& p = L.back();
		if (p.nx <= 5 && p.ny <= 5 && p.nz <= 5) break;
		if (n <= 1000) break;
		MgLevel c;
		c.nx = (p.nx + 2) / 2;
		c.ny = (p.ny + 2) / 2;
		c.nz = (p.nz + 2) / 2;
		c.n = (long long)c.nx * c.ny * c.nz;
		// the coarse-level kernels address vertices with 32-bit indices
		MFX_CHECK(c.n < (1ll << 31), "multigrid: coarse level has too many vertices");
		c.sy = c.nx;
		c.sz = (long long)c.nx * c.ny;
		n = c.n;
		L.push_back(c);
	}
	L[0].r = (float*)alloc(sizeof(float) * (size_t)d.n);
	L[0].x = (float*)alloc(sizeof(float) * (size_t)d.n);
	for (size_t l = 1; l < L.size(); l++) {
		MgLevel& c = L[l];
		c.A = (float*)alloc(sizeof(float) * (size_t)(c.n * S));
		c.x = (float*)alloc(sizeof(float) * (size_t)c.n);
		c.xp = (float*)alloc(sizeof(float) * (size_t)c.n);
		c.b = (float*)alloc(sizeof(float) * (size_t)c.n);
		c.r = (float*)alloc(sizeof(float) * (size_t)c.n);
		c.type = (uint8_t*)alloc((size_t)c.n);
		MFX_HIP(hipMemsetAsync(c.x, 0, sizeof(float) * (size_t)c.n, s->stream));
		const MgLevel& f = L[l - 1];
		hipLaunchKernelGGL(k_mg_activate, dim3(nblk(c.n)), dim3(256), 0, s->stream, lvdev(c), lvdev(f), (int)is3d,
		                   l == 1 ? code : (const uint8_t*)nullptr, f.type, c.type);
		hipLaunchKernelGGL(k_mg_cover, dim3(nblk(f.n)), dim3(256), 0, s->stream, lvdev(c), lvdev(f), (int)is3d,
		                   l == 1 ? code : (const uint8_t*)nullptr, f.type, c.type);
		hipLaunchKernelGGL(k_mg_cover_done, dim3(nblk(c.n)), dim3(256), 0, s->stream, lvdev(c), c.type);
		FineOp F;
		F.code = l == 1 ? code : nullptr;
		F.A = l == 1 ? nullptr : f.A;
		F.type = f.type;
		F.f = lvdev(f);
		F.S = S;
		F.C = S - 1;
		F.is3d = is3d;
		hipLaunchKernelGGL(k_mg_galerkin, dim3(nblk(c.n)), dim3(256), 0, s->stream, F, lvdev(c),
		                   (const uint8_t*)c.type, c.A, c.r);
		hipLaunchKernelGGL(k_mg_drop, dim3(nblk(c.n)), dim3(256), 0, s->stream, lvdev(c), c.type, (const float*)c.A,
		                   (const float*)c.r);
	}
	build_coarsest(s, code);
	// the coarsest level is solved through G; the others get a canonical stencil
	for (int l = 1; l + 1 < (int)L.size(); l++) {
		classify(s, l);
		if (prec == 1) {
			// mixed precision: smoother and residual read BF16 coefficients; the
			// canonical stencil is rounded the same way so that every coupling
			// has one value seen from both of its vertices (symmetric V-cycle)
			MgLevel& c = L[l];
			c.Ah = (uint16_t*)alloc(sizeof(uint16_t) * (size_t)(c.n * S));
			hipLaunchKernelGGL(k_to_bf16, dim3(nblk(c.n * S)), dim3(256), 0, s->stream, (const float*)c.A, c.Ah,
			                   c.n * S);
			for (int q = 0; q < 27; q++) c.canon.a[q] = from_bf16(to_bf16(c.canon.a[q]));
		}
	}
}

// Pick the most frequent full stencil among a lattice of sample vertices
// (all neighbours active) and mark the vertices that share it bitwise.
void MgHierarchy::classify(Scene* s, int l) {
	MgLevel& c = L[l];
	c.hasCanon = false;
	const int per = is3d ? 4 : 8;
	std::vector<long long> smp;
	for (int a = 0; a < per; a++)
		for (int b = 0; b < per; b++)
			for (int e = 0; e < (is3d ? per : 1); e++) {
				const int i = (int)((c.nx - 2) * (a + 0.5) / per) + 1, j = (int)((c.ny - 2) * (b + 0.5) / per) + 1;
				const int k = is3d ? (int)((c.nz - 2) * (e + 0.5) / per) + 1 : 0;
				if (i < 1 || j < 1 || i >= c.nx - 1 || j >= c.ny - 1 || (is3d && (k < 1 || k >= c.nz - 1))) continue;
				smp.push_back(i + c.sy * j + c.sz * k);
			}
	const int ns = (int)smp.size();
	if (ns == 0) return;
	long long* dsmp = nullptr;
	float* dout = nullptr;
	int* dok = nullptr;
	MFX_HIP(hipMalloc((void**)&dsmp, sizeof(long long) * ns));
	MFX_HIP(hipMalloc((void**)&dout, sizeof(float) * 27 * ns));
	MFX_HIP(hipMalloc((void**)&dok, sizeof(int) * ns));
	MFX_HIP(hipMemcpyAsync(dsmp, smp.data(), sizeof(long long) * ns, hipMemcpyHostToDevice, s->stream));
	hipLaunchKernelGGL(k_mg_sample, dim3((ns + 63) / 64), dim3(64), 0, s->stream, lvdev(c), (int)is3d, S,
	                   (const float*)c.A, (const uint8_t*)c.type, (const long long*)dsmp, ns, dout, dok);
	std::vector<float> out((size_t)27 * ns);
	std::vector<int> ok(ns);
	MFX_HIP(hipMemcpyAsync(out.data(), dout, sizeof(float) * 27 * ns, hipMemcpyDeviceToHost, s->stream));
	MFX_HIP(hipMemcpyAsync(ok.data(), dok, sizeof(int) * ns, hipMemcpyDeviceToHost, s->stream));
	sync(s);
	MFX_HIP(hipFree(dsmp));
	MFX_HIP(hipFree(dout));
	MFX_HIP(hipFree(dok));
	int best = -1, bestCount = 0;
	for (int a = 0; a < ns; a++) {
		if (!ok[a]) continue;
		int cnt = 0;
		for (int b = 0; b < ns; b++)
			if (ok[b] && !memcmp(&out[27 * (size_t)a], &out[27 * (size_t)b], sizeof(float) * 27)) cnt++;
		if (cnt > bestCount) {
			bestCount = cnt;
			best = a;
		}
	}
	if (best < 0) return;
	memcpy(c.canon.a, &out[27 * (size_t)best], sizeof(float) * 27);
	c.hasCanon = true;
	hipLaunchKernelGGL(k_mg_classify, dim3(nblk(c.n)), dim3(256), 0, s->stream, lvdev(c), (int)is3d, S,
	                   (const float*)c.A, c.type, c.canon);
}

// dense generalised inverse of the coarsest operator (host, FP64)
void MgHierarchy::build_coarsest(Scene* s, const uint8_t* code0) {
	const int lc = (int)L.size() - 1;
	const MgLevel& c = L[lc];
	std::vector<uint8_t> type((size_t)c.n), hcode;
	std::vector<float> A;
	if (lc == 0) {
		// single level: the dense operator comes from the level-0 matrix code
		hcode.resize((size_t)c.n);
		MFX_HIP(hipMemcpyAsync(hcode.data(), code0, (size_t)c.n, hipMemcpyDeviceToHost, s->stream));
		sync(s);
		for (long long v = 0; v < c.n; v++) type[v] = (hcode[v] & 7) ? 1 : 0;
	} else {
		MFX_HIP(hipMemcpyAsync(type.data(), c.type, (size_t)c.n, hipMemcpyDeviceToHost, s->stream));
		A.resize((size_t)(c.n * S));
		MFX_HIP(hipMemcpyAsync(A.data(), c.A, sizeof(float) * (size_t)(c.n * S), hipMemcpyDeviceToHost, s->stream));
		sync(s);
	}
	std::vector<int> idx, pos((size_t)c.n, -1);
	for (long long v = 0; v < c.n; v++)
		if (type[v]) {
			pos[v] = (int)idx.size();
			idx.push_back((int)v);
		}
	nc = (int)idx.size();
	const int m = nc;
	std::vector<double> M((size_t)m * m, 0.0);
	const int C = S - 1;
	for (int a = 0; a < m; a++) {
		const long long v = idx[a];
		const int i = (int)(v % c.nx), j = (int)((v / c.nx) % c.ny), k = (int)(v / c.sz);
		for (int dz = is3d ? -1 : 0; dz <= (is3d ? 1 : 0); dz++)
			for (int dy = -1; dy <= 1; dy++)
				for (int dx = -1; dx <= 1; dx++) {
					const int ni = i + dx, nj = j + dy, nk = k + dz;
					if (ni < 0 || nj < 0 || nk < 0 || ni >= c.nx || nj >= c.ny || nk >= c.nz) continue;
					const long long nn = ni + c.sy * nj + c.sz * nk;
					if (pos[nn] < 0) continue;
					double val;
					if (lc == 0) {
						const int nzc = (dx != 0) + (dy != 0) + (dz != 0);
						if (nzc > 1) continue;
						const uint8_t cv = hcode[v], cn = hcode[nn];
						if (nzc == 0) val = (cv & 128) ? (double)((float)(cv & 7) * 1e-6f) : (double)(cv & 7);
						else if (dx == 1) val = (cv & 8) ? -1.0 : 0.0;
						else if (dx == -1) val = (cn & 8) ? -1.0 : 0.0;
						else if (dy == 1) val = (cv & 16) ? -1.0 : 0.0;
						else if (dy == -1) val = (cn & 16) ? -1.0 : 0.0;
						else if (dz == 1) val = (cv & 32) ? -1.0 : 0.0;
						else val = (cn & 32) ? -1.0 : 0.0;
					} else {
						const int sidx = is3d ? (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1) : (dx + 1) + 3 * (dy + 1);
						val = sidx >= C ? A[(size_t)(sidx - C) * c.n + v] : A[(size_t)(C - sidx) * c.n + nn];
					}
					M[(size_t)a * m + pos[nn]] = val;
				}
	}
	// Pure-Neumann blocks (fluid pockets enclosed by obstacles, or a closed
	// domain) are singular; with float coefficients their last pivot is
	// round-off rather than zero and would put ~1e7-sized entries into G.
	// Pin one vertex of every connected component whose rows all sum to ~0.
	std::vector<char> dropped(m, 0);
	{
		std::vector<int> comp(m, -1), stack;
		for (int a0 = 0; a0 < m; a0++) {
			if (comp[a0] >= 0) continue;
			std::vector<int> members;
			comp[a0] = a0;
			stack.push_back(a0);
			while (!stack.empty()) {
				const int a = stack.back();
				stack.pop_back();
				members.push_back(a);
				for (int b = 0; b < m; b++)
					if (b != a && comp[b] < 0 && M[(size_t)a * m + b] != 0.0) {
						comp[b] = a0;
						stack.push_back(b);
					}
			}
			double maxSum = 0.0, maxDiag = 0.0;
			for (int a : members) {
				double rs = 0.0;
				for (int b : members) rs += M[(size_t)a * m + b];
				maxSum = std::max(maxSum, std::fabs(rs));
				maxDiag = std::max(maxDiag, M[(size_t)a * m + a]);
			}
			if (members.size() > 1 && maxSum <= 1e-5 * maxDiag) {
				int last = members[0];
				for (int a : members) last = std::max(last, a);
				dropped[last] = 1;
			}
		}
	}
	// Cholesky with dropped (pinned or non-positive) pivots
	std::vector<double> Lm((size_t)m * m, 0.0);
	for (int jj = 0; jj < m; jj++) {
		double dval = M[(size_t)jj * m + jj];
		for (int q = 0; q < jj; q++) dval -= Lm[(size_t)jj * m + q] * Lm[(size_t)jj * m + q];
		const double orig = M[(size_t)jj * m + jj];
		if (dropped[jj] || !(dval > 1e-10 * (orig > 0 ? orig : 1.0))) {
			dropped[jj] = 1;
			for (int q = 0; q < jj; q++) Lm[(size_t)jj * m + q] = 0.0;
			Lm[(size_t)jj * m + jj] = 1.0;
			continue;
		}
		const double ljj = std::sqrt(dval);
		Lm[(size_t)jj * m + jj] = ljj;
		for (int ii = jj + 1; ii < m; ii++) {
			double sv = M[(size_t)ii * m + jj];
			const double* li = &Lm[(size_t)ii * m];
			const double* lj = &Lm[(size_t)jj * m];
			for (int q = 0; q < jj; q++) sv -= li[q] * lj[q];
			Lm[(size_t)ii * m + jj] = sv / ljj;
		}
	}
	for (int jj = 0; jj < m; jj++)
		if (dropped[jj])
			for (int ii = jj + 1; ii < m; ii++) Lm[(size_t)ii * m + jj] = 0.0;
	// Linv (lower), then G = Linv^T Linv
	std::vector<double> Li((size_t)m * m, 0.0);
	for (int col = 0; col < m; col++) {
		Li[(size_t)col * m + col] = 1.0 / Lm[(size_t)col * m + col];
		for (int row = col + 1; row < m; row++) {
			double sv = 0.0;
			for (int q = col; q < row; q++) sv -= Lm[(size_t)row * m + q] * Li[(size_t)q * m + col];
			Li[(size_t)row * m + col] = sv / Lm[(size_t)row * m + row];
		}
	}
	std::vector<double> Gh((size_t)m * m, 0.0);
	for (int a = 0; a < m; a++) {
		if (dropped[a]) continue;
		for (int b = a; b < m; b++) {
			if (dropped[b]) continue;
			double sv = 0.0;
			for (int q = b; q < m; q++) sv += Li[(size_t)q * m + a] * Li[(size_t)q * m + b];
			Gh[(size_t)a * m + b] = sv;
			Gh[(size_t)b * m + a] = sv;
		}
	}
	G = (double*)alloc(sizeof(double) * (size_t)m * m);
	cidx = (int*)alloc(sizeof(int) * (size_t)m);
	MFX_HIP(hipMemcpyAsync(G, Gh.data(), sizeof(double) * (size_t)m * m, hipMemcpyHostToDevice, s->stream));
	MFX_HIP(hipMemcpyAsync(cidx, idx.data(), sizeof(int) * (size_t)m, hipMemcpyHostToDevice, s->stream));
	sync(s);
}

// ---------------------------------------------------------------------------
// launches (single problem and work lists)
// ---------------------------------------------------------------------------
constexpr int kZmTY = kZmTYHost;
long long l0_tiles(const Dims& d, int* kchunk, bool tuned, int force) {
	// 3D: z-marching 64 x kZmTY column tiles over chunks of a multiple of
	// kDotGroup planes (even, as the restriction pairs planes), halved while
	// fewer than about 8192 wavefronts (one per SIMD slot of the chip) cover
	// the grid; 2D: bricks of 64x16 cells
	const int BY = d.is3d ? kZmTY : 16;
	const long long cols = (long long)((d.nx + 63) / 64) * ((d.ny + BY - 1) / BY);
	if (!d.is3d) {
		*kchunk = 2;
		return cols;
	}
	int kc = zm_chunk(d.nz);
	while (kc > 16 && cols * ((d.nz + kc - 1) / kc) < 8192 / kZmTY) kc = zm_chunk(kc / 2);
	if (tuned && zm_tuned(d).l0 > 0) kc = zm_tuned(d).l0;
	if (force > 0) kc = force;
	*kchunk = kc;
	return cols * ((d.nz + kc - 1) / kc);
}
long long lcblk_tiles(const LvDev& c, bool is3d) {
	const long long bx = is3d ? kLcBX3 : kLcBX2, by = is3d ? kLcBY3 : kLcBY2, bz = is3d ? kLcBZ3 : 1;
	return ((c.nx + bx - 1) / bx) * ((c.ny + by - 1) / by) * ((c.nz + bz - 1) / bz);
}
void run_l0_down_one(hipStream_t s, bool is3d, unsigned nb, const L0DownP& p) {
	if (is3d) launch_one<L0DownP, body_l0_down_zm<kZmTY>, 64 * kZmTY, 8>(s, nb, p);
	else launch_one<L0DownP, body_l0_down_brick<64, 16, 1, false>, 1024>(s, nb, p);
}
void run_l0_down_many(hipStream_t s, bool is3d, unsigned nb, WorkList items, const L0DownP* p) {
	if (is3d) launch_many<L0DownP, body_l0_down_zm<kZmTY>, 64 * kZmTY, 8>(s, nb, items, p);
	else launch_many<L0DownP, body_l0_down_brick<64, 16, 1, false>, 1024>(s, nb, items, p);
}
void run_l0_up_many(hipStream_t s, bool is3d, unsigned nb, WorkList items, const L0UpP* p) {
	if (is3d) launch_many<L0UpP, body_l0_up_zm<kZmTY>, 64 * kZmTY, 8>(s, nb, items, p);
	else launch_many<L0UpP, body_l0_up_brick<64, 16, 1, false>, 1024>(s, nb, items, p);
}
void run_restrict_many(hipStream_t s, bool is3d, unsigned nb, WorkList items, const RestrictP* p) {
	if (is3d) launch_many<RestrictP, body_restrict<true>, 256>(s, nb, items, p);
	else launch_many<RestrictP, body_restrict<false>, 256>(s, nb, items, p);
}
template <bool POST>
static void run_smooth_mode(hipStream_t s, bool is3d, bool bf16, unsigned nb, WorkList items,
                            const SmoothP* p) {
	if (is3d && bf16) launch_many<SmoothP, body_smooth<true, uint16_t, POST>, kSmoothNT>(s, nb, items, p);
	else if (is3d) launch_many<SmoothP, body_smooth<true, float, POST>, kSmoothNT>(s, nb, items, p);
	else if (bf16) launch_many<SmoothP, body_smooth<false, uint16_t, POST>, kSmoothNT>(s, nb, items, p);
	else launch_many<SmoothP, body_smooth<false, float, POST>, kSmoothNT>(s, nb, items, p);
}
void run_smooth_many(hipStream_t s, bool is3d, bool bf16, bool post, unsigned nb, WorkList items,
                     const SmoothP* p) {
	if (post) run_smooth_mode<true>(s, is3d, bf16, nb, items, p);
	else run_smooth_mode<false>(s, is3d, bf16, nb, items, p);
}
void run_resid_many(hipStream_t s, bool is3d, bool bf16, unsigned nb, WorkList items, const ResidP* p) {
	if (is3d && bf16) launch_many<ResidP, body_resid<true, uint16_t>, 256>(s, nb, items, p);
	else if (is3d) launch_many<ResidP, body_resid<true, float>, 256>(s, nb, items, p);
	else if (bf16) launch_many<ResidP, body_resid<false, uint16_t>, 256>(s, nb, items, p);
	else launch_many<ResidP, body_resid<false, float>, 256>(s, nb, items, p);
}
void run_coarsest_many(hipStream_t s, unsigned nb, WorkList items, const CoarsestP* p) {
	launch_many<CoarsestP, body_coarsest, 256>(s, nb, items, p);
}
void run_fill_many(hipStream_t s, unsigned nb, WorkList items, const FillP* p) {
	launch_many<FillP, body_fill, 256>(s, nb, items, p);
}

// Parameters of every V-cycle step of this hierarchy (shared by the single
// launch sequence below and by batched work lists).
MgHierarchy::Steps MgHierarchy::steps(const Dims& d, const uint8_t* code, const float* r, float* z, const CgState* st,
                                      double* part, bool tuned, int forceChunk) const {
	Steps S_;
	const int nl = (int)L.size();
	S_.nl = nl;
	S_.l0Tiles = l0_tiles(d, &S_.kchunk, tuned, forceChunk);
	S_.nPart = is3d ? zm_dot_parts(d, kZmTY) : S_.l0Tiles;
	const void* noA = nullptr;
	if (nl == 1) {
		S_.fill = FillP{st, z, d.n};
		S_.coarsest = CoarsestP{st, nc, cidx, G, r, code, z};
		return S_;
	}
	S_.down = L0DownP{d, S_.kchunk, st, code, r, L[0].x, L[0].r};
	for (int l = 1; l < nl; l++) {
		const MgLevel& c = L[l];
		const MgLevel& f = L[l - 1];
		Steps::Level q;
		q.restrict_ = RestrictP{st, lvdev(c), lvdev(f), c.type, f.r, c.b};
		q.tiles = lc_tiles(lvdev(c));
		q.blkTiles = lcblk_tiles(lvdev(c), is3d);
		if (l < nl - 1) {
			const void* A = c.Ah ? (const void*)c.Ah : (const void*)c.A;
			q.pre = SmoothP{st, lvdev(c), A, c.type, c.b, nullptr, c.x, 1, 0, lvdev(c), nullptr, c.canon};
			q.resid = ResidP{st, lvdev(c), A, c.type, c.b, c.x, c.r, c.canon};
			const MgLevel& cc = L[l + 1];
			q.post = SmoothP{st, lvdev(c), A, c.type, c.b, c.x, c.xp, 0, 1, lvdev(cc), l + 1 == nl - 1 ? cc.x : cc.xp,
			                 c.canon};
		}
		S_.lv.push_back(q);
	}
	(void)noA;
	S_.coarsest = CoarsestP{st, nc, cidx, G, L[nl - 1].b, nullptr, L[nl - 1].x};
	const MgLevel& c1 = L[1];
	S_.up = L0UpP{d, S_.kchunk, st, code, r, L[0].x, z, lvdev(c1), nl == 2 ? c1.x : c1.xp, part};
	return S_;
}

int MgHierarchy::vcycle(Scene* s, const uint8_t* code, const float* r, float* z, const CgState* st) {
	const Dims& d = s->d;
	const Steps S_ = steps(d, code, r, z, st, s->dpart, true);
	hipStream_t q = s->stream;
	const bool bf = prec == 1;
	if (S_.nl == 1) {
		launch_one<FillP, body_fill, 256>(q, (d.n + 255) / 256, S_.fill);
		launch_one<CoarsestP, body_coarsest, 256>(q, coarsest_tiles(nc), S_.coarsest);
		return 0;
	}
	{
		KtScope kt("mg_l0_down", q);
		if (is3d) launch_one<L0DownP, body_l0_down_zm<kZmTY>, 64 * kZmTY, 8>(q, S_.l0Tiles, S_.down);
		else launch_one<L0DownP, body_l0_down_brick<64, 16, 1, false>, 1024>(q, S_.l0Tiles, S_.down);
	}
	const int nlv = (int)S_.lv.size();
	for (int i = 0; i < nlv; i++) {
		const Steps::Level& v = S_.lv[i];
		if (is3d) launch_one<RestrictP, body_restrict<true>, 256>(q, v.tiles, v.restrict_);
		else launch_one<RestrictP, body_restrict<false>, 256>(q, v.tiles, v.restrict_);
		if (i == nlv - 1) break;
		if (is3d && bf) {
			launch_one<SmoothP, body_smooth<true, uint16_t, false>, kSmoothNT>(q, v.blkTiles, v.pre);
			launch_one<ResidP, body_resid<true, uint16_t>, 256>(q, v.tiles, v.resid);
		} else if (is3d) {
			launch_one<SmoothP, body_smooth<true, float, false>, kSmoothNT>(q, v.blkTiles, v.pre);
			launch_one<ResidP, body_resid<true, float>, 256>(q, v.tiles, v.resid);
		} else if (bf) {
			launch_one<SmoothP, body_smooth<false, uint16_t, false>, kSmoothNT>(q, v.blkTiles, v.pre);
			launch_one<ResidP, body_resid<false, uint16_t>, 256>(q, v.tiles, v.resid);
		} else {
			launch_one<SmoothP, body_smooth<false, float, false>, kSmoothNT>(q, v.blkTiles, v.pre);
			launch_one<ResidP, body_resid<false, float>, 256>(q, v.tiles, v.resid);
		}
	}
	launch_one<CoarsestP, body_coarsest, 256>(q, coarsest_tiles(nc), S_.coarsest);
	for (int i = nlv - 2; i >= 0; i--) {
		const Steps::Level& v = S_.lv[i];
		if (is3d && bf) launch_one<SmoothP, body_smooth<true, uint16_t, true>, kSmoothNT>(q, v.blkTiles, v.post);
		else if (is3d) launch_one<SmoothP, body_smooth<true, float, true>, kSmoothNT>(q, v.blkTiles, v.post);
		else if (bf) launch_one<SmoothP, body_smooth<false, uint16_t, true>, kSmoothNT>(q, v.blkTiles, v.post);
		else launch_one<SmoothP, body_smooth<false, float, true>, kSmoothNT>(q, v.blkTiles, v.post);
	}
	{
		KtScope kt("mg_l0_up", q);
		if (is3d) launch_one<L0UpP, body_l0_up_zm<kZmTY>, 64 * kZmTY, 8>(q, S_.l0Tiles, S_.up);
		else launch_one<L0UpP, body_l0_up_brick<64, 16, 1, false>, 1024>(q, S_.l0Tiles, S_.up);
	}
	return (int)S_.nPart;
}

}  // namespace mfx
