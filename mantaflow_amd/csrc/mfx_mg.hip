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

#include "mfx_mg.h"
#include "mfx_prof.h"

namespace mfx {

__device__ __host__ inline int oddc(int i, int j, int k) { return (i & 1) + (j & 1) + (k & 1); }
__device__ inline float wgt(int i, int j, int k) { return 1.0f / (float)(1 << oddc(i, j, k)); }

struct LvDev {
	int nx, ny, nz;
	long long sy, sz, n;
};
static inline LvDev lvdev(const MgLevel& l) { return LvDev{l.nx, l.ny, l.nz, l.sy, l.sz, l.n}; }

__device__ inline bool l0_active(uint8_t c) { return (c & 7) != 0; }
__device__ inline float l0_diag(uint8_t c) { return (c & 128) ? (float)(c & 7) * 1e-6f : (float)(c & 7); }
__device__ inline float l0_scale(uint8_t c) { return (c & 128) ? 1e-6f : 1.0f; }

// ---------------------------------------------------------------------------
// hierarchy construction
// ---------------------------------------------------------------------------
// coarse vertex active iff an active fine vertex lies in [2V-1, 2V+1]
__global__ void k_mg_activate(LvDev c, LvDev f, int is3d, const uint8_t* __restrict__ code,
                              const uint8_t* __restrict__ ftype, uint8_t* __restrict__ ctype) {
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x) {
		const int i = (int)(v % c.nx), j = (int)((v / c.nx) % c.ny), k = (int)(v / c.sz);
		bool act = false;
		const int k0 = is3d ? max(0, 2 * k - 1) : 0, k1 = is3d ? min(f.nz - 1, 2 * k + 1) : 0;
		for (int uz = k0; uz <= k1 && !act; uz++)
			for (int uy = max(0, 2 * j - 1); uy <= min(f.ny - 1, 2 * j + 1) && !act; uy++)
				for (int ux = max(0, 2 * i - 1); ux <= min(f.nx - 1, 2 * i + 1); ux++) {
					const long long u = ux + f.sy * uy + f.sz * uz;
					if (code ? l0_active(code[u]) : (ftype[u] != 0)) {
						act = true;
						break;
					}
				}
		ctype[v] = act ? 1 : 0;
	}
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

// ---------------------------------------------------------------------------
// V-cycle kernels
// ---------------------------------------------------------------------------
// level 0 red-black Gauss-Seidel (colour = parity of i+j+k).  first: x = 0
// everywhere, then the colour-0 sweep (which then only reads zeros).
__global__ void __launch_bounds__(kBlock) k_l0_smooth(Dims d, Tiles t, const CgState* __restrict__ st,
                                                      const uint8_t* __restrict__ code, const float* __restrict__ rin,
                                                      float* __restrict__ x, int color, int first) {
	if (st->done) return;
	CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const long long c = d.idx(p.i, p.j, p.k);
	const int col = (p.i + p.j + p.k) & 1;
	const uint8_t cc = code[c];
	if (first) {
		x[c] = (col == 0 && l0_active(cc)) ? (rin[c] * l0_scale(cc)) / l0_diag(cc) : 0.f;
		return;
	}
	if (col != color || !l0_active(cc)) return;
	float sum = rin[c] * l0_scale(cc);
	if (p.i > 0 && (code[c - 1] & 8)) sum += x[c - 1];
	if (p.i < d.nx - 1 && (cc & 8)) sum += x[c + 1];
	if (p.j > 0 && (code[c - d.nx] & 16)) sum += x[c - d.nx];
	if (p.j < d.ny - 1 && (cc & 16)) sum += x[c + d.nx];
	if (d.is3d) {
		if (p.k > 0 && (code[c - d.sz] & 32)) sum += x[c - d.sz];
		if (p.k < d.nz - 1 && (cc & 32)) sum += x[c + d.sz];
	}
	x[c] = sum / l0_diag(cc);
}

// level 0 residual r0 = b - A x (0 on inactive vertices)
__global__ void __launch_bounds__(kBlock) k_l0_residual(Dims d, Tiles t, const CgState* __restrict__ st,
                                                        const uint8_t* __restrict__ code,
                                                        const float* __restrict__ rin, const float* __restrict__ x,
                                                        float* __restrict__ r0) {
	if (st->done) return;
	CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const long long c = d.idx(p.i, p.j, p.k);
	const uint8_t cc = code[c];
	if (!l0_active(cc)) {
		r0[c] = 0.f;
		return;
	}
	float sum = rin[c] * l0_scale(cc);
	if (p.i > 0 && (code[c - 1] & 8)) sum += x[c - 1];
	if (p.i < d.nx - 1 && (cc & 8)) sum += x[c + 1];
	if (p.j > 0 && (code[c - d.nx] & 16)) sum += x[c - d.nx];
	if (p.j < d.ny - 1 && (cc & 16)) sum += x[c + d.nx];
	if (d.is3d) {
		if (p.k > 0 && (code[c - d.sz] & 32)) sum += x[c - d.sz];
		if (p.k < d.nz - 1 && (cc & 32)) sum += x[c + d.sz];
	}
	sum -= l0_diag(cc) * x[c];
	r0[c] = sum;
}

// b_c[V] = sum_{U in [2V-1,2V+1]} w(U) r_f[U]   (r_f is 0 on inactive vertices)
__global__ void k_mg_restrict(const CgState* __restrict__ st, LvDev c, LvDev f, int is3d,
                              const uint8_t* __restrict__ ctype, const float* __restrict__ rf, float* __restrict__ bc) {
	if (st->done) return;
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x) {
		if (!ctype[v]) {
			bc[v] = 0.f;
			continue;
		}
		const int i = (int)(v % c.nx), j = (int)((v / c.nx) % c.ny), k = (int)(v / c.sz);
		float sum = 0.f;
		const int k0 = is3d ? max(0, 2 * k - 1) : 0, k1 = is3d ? min(f.nz - 1, 2 * k + 1) : 0;
		for (int uz = k0; uz <= k1; uz++)
			for (int uy = max(0, 2 * j - 1); uy <= min(f.ny - 1, 2 * j + 1); uy++)
				for (int ux = max(0, 2 * i - 1); ux <= min(f.nx - 1, 2 * i + 1); ux++)
					sum += wgt(ux, uy, uz) * rf[ux + f.sy * uy + f.sz * uz];
		bc[v] = sum;
	}
}

// coarse-level Gauss-Seidel on one colour (parity triple); first: zero x and
// sweep colour 0 from zeros
__global__ void k_lc_smooth(const CgState* __restrict__ st, LvDev c, int is3d, int S, const float* __restrict__ A,
                            const uint8_t* __restrict__ type, const float* __restrict__ b, float* __restrict__ x,
                            int cx, int cy, int cz, int first) {
	if (st->done) return;
	const int C = S - 1;
	if (first) {
		for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
		     v += (long long)gridDim.x * blockDim.x) {
			const int i = (int)(v % c.nx), j = (int)((v / c.nx) % c.ny), k = (int)(v / c.sz);
			const bool col0 = ((i & 1) == 0) && ((j & 1) == 0) && ((k & 1) == 0);
			x[v] = (col0 && type[v]) ? b[v] / A[v] : 0.f;
		}
		return;
	}
	const int hx = (c.nx - cx + 1) / 2, hy = (c.ny - cy + 1) / 2, hz = (c.nz - cz + 1) / 2;
	const long long m = (long long)hx * hy * hz;
	for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < m; q += (long long)gridDim.x * blockDim.x) {
		const int i = 2 * (int)(q % hx) + cx, j = 2 * (int)((q / hx) % hy) + cy, k = 2 * (int)(q / ((long long)hx * hy)) + cz;
		const long long v = i + c.sy * j + c.sz * k;
		if (!type[v]) continue;
		float sum = b[v];
		const int dz0 = is3d ? -1 : 0, dz1 = is3d ? 1 : 0;
		for (int dz = dz0; dz <= dz1; dz++)
			for (int dy = -1; dy <= 1; dy++)
				for (int dx = -1; dx <= 1; dx++) {
					const int s = is3d ? (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1) : (dx + 1) + 3 * (dy + 1);
					if (s == C) continue;
					const int ni = i + dx, nj = j + dy, nk = k + dz;
					if (ni < 0 || nj < 0 || nk < 0 || ni >= c.nx || nj >= c.ny || nk >= c.nz) continue;
					const long long nn = ni + c.sy * nj + c.sz * nk;
					if (!type[nn]) continue;
					const float a = s > C ? A[(long long)(s - C) * c.n + v] : A[(long long)(C - s) * c.n + nn];
					sum -= a * x[nn];
				}
		x[v] = sum / A[v];
	}
}

__global__ void k_lc_residual(const CgState* __restrict__ st, LvDev c, int is3d, int S, const float* __restrict__ A,
                              const uint8_t* __restrict__ type, const float* __restrict__ b,
                              const float* __restrict__ x, float* __restrict__ r) {
	if (st->done) return;
	const int C = S - 1;
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < c.n;
	     v += (long long)gridDim.x * blockDim.x) {
		if (!type[v]) {
			r[v] = 0.f;
			continue;
		}
		const int i = (int)(v % c.nx), j = (int)((v / c.nx) % c.ny), k = (int)(v / c.sz);
		float sum = b[v];
		const int dz0 = is3d ? -1 : 0, dz1 = is3d ? 1 : 0;
		for (int dz = dz0; dz <= dz1; dz++)
			for (int dy = -1; dy <= 1; dy++)
				for (int dx = -1; dx <= 1; dx++) {
					const int s = is3d ? (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1) : (dx + 1) + 3 * (dy + 1);
					const int ni = i + dx, nj = j + dy, nk = k + dz;
					if (ni < 0 || nj < 0 || nk < 0 || ni >= c.nx || nj >= c.ny || nk >= c.nz) continue;
					const long long nn = ni + c.sy * nj + c.sz * nk;
					if (!type[nn]) continue;
					const float a = s >= C ? A[(long long)(s - C) * c.n + v] : A[(long long)(C - s) * c.n + nn];
					sum -= a * x[nn];
				}
		r[v] = sum;
	}
}

// x_f[V] += w(V) * sum_{I in [V/2,(V+1)/2], active} x_c[I] on active fine vertices
__global__ void k_mg_prolong(const CgState* __restrict__ st, LvDev f, LvDev c, int is3d,
                             const uint8_t* __restrict__ fcode, const uint8_t* __restrict__ ftype,
                             const uint8_t* __restrict__ ctype, const float* __restrict__ xc, float* __restrict__ xf) {
	if (st->done) return;
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < f.n;
	     v += (long long)gridDim.x * blockDim.x) {
		const bool act = fcode ? l0_active(fcode[v]) : ftype[v] != 0;
		if (!act) continue;
		const int i = (int)(v % f.nx), j = (int)((v / f.nx) % f.ny), k = (int)(v / f.sz);
		float sum = 0.f;
		for (int nz = k / 2; nz <= (k + 1) / 2; nz++)
			for (int ny = j / 2; ny <= (j + 1) / 2; ny++)
				for (int nx = i / 2; nx <= (i + 1) / 2; nx++) {
					const long long q = nx + c.sy * ny + c.sz * nz;
					if (ctype[q]) sum += xc[q];
				}
		xf[v] += wgt(i, j, k) * sum;
	}
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

__device__ inline void brick_origin(int nx, int ny, int BX, int BY, int BZ, int& ox, int& oy, int& oz) {
	const int nbx = (nx + BX - 1) / BX, nby = (ny + BY - 1) / BY;
	const long long blk = xcd_tile(blockIdx.x, gridDim.x);
	ox = (int)(blk % nbx) * BX;
	const long long rest = blk / nbx;
	oy = (int)(rest % nby) * BY;
	oz = (int)(rest / nby) * BZ;
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
__global__ void __launch_bounds__(1024) k_l0_down(Dims d, const CgState* __restrict__ st,
                                                   const uint8_t* __restrict__ code, const float* __restrict__ rin,
                                                   float* __restrict__ z, float* __restrict__ r0) {
	using B = Brick<BX, BY, BZ>;
	__shared__ float xs[B::RN];
	__shared__ float bs[B::RN];
	__shared__ uint8_t cs[B::RN];
	if (st->done) return;
	int ox, oy, oz;
	brick_origin(d.nx, d.ny, BX, BY, BZ, ox, oy, oz);
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
__global__ void __launch_bounds__(1024) k_l0_up(Dims d, const CgState* __restrict__ st,
                                                 const uint8_t* __restrict__ code, const float* __restrict__ rin,
                                                 const float* __restrict__ xpre, float* __restrict__ z, LvDev c1,
                                                 const uint8_t* __restrict__ ctype, const float* __restrict__ xc,
                                                 double* __restrict__ part) {
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
	brick_origin(d.nx, d.ny, BX, BY, BZ, ox, oy, oz);
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
	if (threadIdx.x == 0) part[3 * blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// z-marching level-0 kernels (3D).  A workgroup owns a 64 x TY column tile
// and a chunk of planes [k0, k1).  It walks the planes once, staging each
// plane of the tile plus a 2-cell ring in a 5-slot LDS ring buffer, and runs
// the three dependent stages of red-black Gauss-Seidel as a pipeline:
//   A(p)   load plane p (code, rhs, x), first colour pointwise
//   B(p-1) second colour on plane p-1 (tile + 1-cell ring)
//   C(p-2) final output on the tile's cells of plane p-2
// so every input element is read from HBM once plus 4 halo planes per
// chunk.  Arithmetic and operation order equal the brick kernels above.
// ---------------------------------------------------------------------------
template <int TY>
struct ZTile {
	static constexpr int TX = 64, H = 2, RX = TX + 2 * H, RY = TY + 2 * H, RP = RX * RY, NS = 5, NT = TX * TY;
};

__device__ inline void zm_origin(const Dims& d, int TY, int kchunk, int& ox, int& oy, int& k0, int& k1) {
	const int ntx = (d.nx + 63) / 64, nty = (d.ny + TY - 1) / TY;
	const long long blk = xcd_tile(blockIdx.x, gridDim.x);
	ox = (int)(blk % ntx) * 64;
	const long long rest = blk / ntx;
	oy = (int)(rest % nty) * TY;
	k0 = (int)(rest / nty) * kchunk;
	k1 = min(d.nz, k0 + kchunk);
}

// l0_nb_sum with the plane below (cM, xM) and above (xU) in separate slots
__device__ inline float zm_nb_sum(const uint8_t* cP, const float* xP, const uint8_t* cM, const float* xM,
                                  const float* xU, int q, int RX) {
	const uint8_t cc = cP[q];
	float sum = 0.f;
	if (cP[q - 1] & 8) sum += xP[q - 1];
	if (cc & 8) sum += xP[q + 1];
	if (cP[q - RX] & 16) sum += xP[q - RX];
	if (cc & 16) sum += xP[q + RX];
	if (cM[q] & 32) sum += xM[q];
	if (cc & 32) sum += xU[q];
	return sum;
}

__device__ inline int zslot(int p) { return (p + 10) % 5; }  // p >= -2

// The (at most two) region cells of one thread: q = tid and tid + NT.
template <typename T>
struct ZmCells {
	static_assert(T::RP <= 2 * T::NT, "two region cells per thread");
	long long g[2];  // in-plane offset i + nx*j
	bool ok[2];      // inside the region and the grid
	int par[2];      // (i + j) & 1
	__device__ ZmCells(const Dims& d, int rx0, int ry0) {
#pragma unroll
		for (int u = 0; u < 2; u++) {
			const int q = threadIdx.x + u * T::NT;
			const int gx = rx0 + q % T::RX, gy = ry0 + q / T::RX;
			ok[u] = q < T::RP && gx >= 0 && gy >= 0 && gx < d.nx && gy < d.ny;
			g[u] = ok[u] ? (long long)gx + (long long)d.nx * gy : 0;
			par[u] = (gx + gy) & 1;
		}
	}
};

// Down leg (same result as k_l0_down): red from x = 0, black, residual.
template <int TY>
__global__ void __launch_bounds__(64 * TY) k_l0_down_zm(Dims d, int kchunk, const CgState* __restrict__ st,
                                                         const uint8_t* __restrict__ code,
                                                         const float* __restrict__ rin, float* __restrict__ xout,
                                                         float* __restrict__ r0) {
	using T = ZTile<TY>;
	__shared__ float xs[T::NS][T::RP];
	__shared__ float bs[T::NS][T::RP];
	__shared__ uint8_t cs[T::NS][T::RP];
	if (st->done) return;
	int ox, oy, k0, k1;
	zm_origin(d, TY, kchunk, ox, oy, k0, k1);
	const int rx0 = ox - T::H, ry0 = oy - T::H;
	ZmCells<T> cells(d, rx0, ry0);
	uint8_t pc[2];
	float pr[2];
	// plane p's inputs are fetched into registers one plane ahead, so the
	// loads are in flight during the barriers and the sweeps of plane p-1
	auto fetch = [&](int p) {
#pragma unroll
		for (int u = 0; u < 2; u++) {
			pc[u] = 0;
			pr[u] = 0.f;
			if (cells.ok[u] && p >= 0 && p < d.nz) {
				const long long g = cells.g[u] + d.sz * p;
				pc[u] = code[g];
				pr[u] = rin[g];
			}
		}
	};
	fetch(k0 - 2);
	for (int p = k0 - 2; p <= k1 + 1; p++) {
		const int sp = zslot(p);
#pragma unroll
		for (int u = 0; u < 2; u++) {
			const int q = threadIdx.x + u * T::NT;
			if (q >= T::RP) continue;
			const uint8_t cc = pc[u];
			const float b = pr[u] * l0_scale(cc);
			cs[sp][q] = cc;
			bs[sp][q] = b;
			xs[sp][q] = (l0_active(cc) && ((cells.par[u] + p) & 1) == 0) ? b / l0_diag(cc) : 0.f;
		}
		if (p <= k1) fetch(p + 1);
		__syncthreads();
		const int pb = p - 1;
		if (pb >= k0 - 1 && pb <= k1 && pb >= 0 && pb < d.nz) {
			const int s0 = zslot(pb), sm = zslot(pb - 1), su = zslot(pb + 1);
			constexpr int IX = T::RX - 2, IN = IX * (T::RY - 2);
			for (int t = threadIdx.x; t < IN; t += T::NT) {
				const int lx = 1 + t % IX, ly = 1 + t / IX;
				const int q = lx + T::RX * ly;
				const uint8_t cc = cs[s0][q];
				if (((rx0 + lx + ry0 + ly + pb) & 1) == 1 && l0_active(cc))
					xs[s0][q] = (bs[s0][q] + zm_nb_sum(cs[s0], xs[s0], cs[sm], xs[sm], xs[su], q, T::RX)) / l0_diag(cc);
			}
		}
		__syncthreads();
		const int pc = p - 2;
		if (pc >= k0 && pc < k1) {
			const int s0 = zslot(pc), sm = zslot(pc - 1), su = zslot(pc + 1);
			const int lx = T::H + (threadIdx.x & 63), ly = T::H + (threadIdx.x >> 6);
			const int gx = rx0 + lx, gy = ry0 + ly;
			if (gx < d.nx && gy < d.ny) {
				const int q = lx + T::RX * ly;
				const long long g = d.idx(gx, gy, pc);
				const uint8_t cc = cs[s0][q];
				const float x = xs[s0][q];
				xout[g] = x;
				r0[g] = l0_active(cc)
				            ? (bs[s0][q] + zm_nb_sum(cs[s0], xs[s0], cs[sm], xs[sm], xs[su], q, T::RX)) - l0_diag(cc) * x
				            : 0.f;
			}
		}
	}
}

// Up leg (same result as k_l0_up): x = xpre + I x_1, black, red, write z and
// the partial sums of fl(z*r).  Coarse planes are staged in a 2-slot ring:
// fine plane p needs coarse planes p>>1 and (p+1)>>1.
template <int TY>
__global__ void __launch_bounds__(64 * TY) k_l0_up_zm(Dims d, int kchunk, const CgState* __restrict__ st,
                                                       const uint8_t* __restrict__ code,
                                                       const float* __restrict__ rin, const float* __restrict__ xpre,
                                                       float* __restrict__ z, LvDev c1, const float* __restrict__ xc,
                                                       double* __restrict__ part) {
	using T = ZTile<TY>;
	constexpr int CRX = T::RX / 2 + 2, CRY = T::RY / 2 + 2, CRP = CRX * CRY;
	__shared__ float xs[T::NS][T::RP];
	__shared__ float rs[T::NS][T::RP];
	__shared__ uint8_t cs[T::NS][T::RP];
	__shared__ float xcs[2][CRP];
	__shared__ double sh[16];
	if (st->done) return;
	int ox, oy, k0, k1;
	zm_origin(d, TY, kchunk, ox, oy, k0, k1);
	const int rx0 = ox - T::H, ry0 = oy - T::H;
	const int cx0 = rx0 >> 1, cy0 = ry0 >> 1;  // floor division
	static_assert(CRP <= T::NT, "one coarse value per thread");
	// this thread's vertex of the coarse-plane region
	const int cq = threadIdx.x;
	const int cpx = cx0 + cq % CRX, cpy = cy0 + cq / CRX;
	const bool cok = cq < CRP && cpx >= 0 && cpy >= 0 && cpx < c1.nx && cpy < c1.ny;
	const long long cg = cok ? (long long)cpx + c1.sy * cpy : 0;
	auto cfetch = [&](int P) -> float { return (cok && P >= 0 && P < c1.nz) ? xc[cg + c1.sz * P] : 0.f; };
	ZmCells<T> cells(d, rx0, ry0);
	uint8_t pc[2];
	float pr[2], px[2];
	auto fetch = [&](int p) {
#pragma unroll
		for (int u = 0; u < 2; u++) {
			pc[u] = 0;
			pr[u] = 0.f;
			px[u] = 0.f;
			if (cells.ok[u] && p >= 0 && p < d.nz) {
				const long long g = cells.g[u] + d.sz * p;
				pc[u] = code[g];
				pr[u] = rin[g];
				px[u] = xpre[g];
			}
		}
	};
	// coarse plane (k0-2)/2 goes to LDS now; later planes are fetched into a
	// register two fine planes before they are needed (k0 is even)
	if (cq < CRP) xcs[((k0 - 2) >> 1) & 1][cq] = cfetch((k0 - 2) >> 1);
	float creg = cfetch(k0 >> 1);
	fetch(k0 - 2);
	__syncthreads();
	double acc = 0.0;
	for (int p = k0 - 2; p <= k1 + 1; p++) {
		const int sp = zslot(p);
#pragma unroll
		for (int u = 0; u < 2; u++) {
			const int q = threadIdx.x + u * T::NT;
			if (q >= T::RP) continue;
			const uint8_t cc = pc[u];
			float x = px[u];
			if (l0_active(cc)) {
				const int gx = rx0 + q % T::RX, gy = ry0 + q / T::RX;
				float sum = 0.f;
				for (int pz = p >> 1; pz <= (p + 1) >> 1; pz++)
					for (int py = gy >> 1; py <= (gy + 1) >> 1; py++)
						for (int qx = gx >> 1; qx <= (gx + 1) >> 1; qx++)
							sum += xcs[pz & 1][(qx - cx0) + CRX * (py - cy0)];
				x += wgt(gx, gy, p) * sum;
			}
			cs[sp][q] = cc;
			rs[sp][q] = pr[u];
			xs[sp][q] = x;
		}
		if (!(p & 1)) {
			// coarse plane p/2+1 is first needed by plane p+1
			if (cq < CRP) xcs[((p >> 1) + 1) & 1][cq] = creg;
			creg = cfetch((p >> 1) + 2);
		}
		if (p <= k1) fetch(p + 1);
		__syncthreads();
		const int pb = p - 1;
		if (pb >= k0 - 1 && pb <= k1 && pb >= 0 && pb < d.nz) {
			const int s0 = zslot(pb), sm = zslot(pb - 1), su = zslot(pb + 1);
			constexpr int IX = T::RX - 2, IN = IX * (T::RY - 2);
			for (int t = threadIdx.x; t < IN; t += T::NT) {
				const int lx = 1 + t % IX, ly = 1 + t / IX;
				const int q = lx + T::RX * ly;
				const uint8_t cc = cs[s0][q];
				if (((rx0 + lx + ry0 + ly + pb) & 1) == 1 && l0_active(cc))
					xs[s0][q] = (rs[s0][q] * l0_scale(cc) + zm_nb_sum(cs[s0], xs[s0], cs[sm], xs[sm], xs[su], q, T::RX)) /
					            l0_diag(cc);
			}
		}
		__syncthreads();
		const int pc = p - 2;
		if (pc >= k0 && pc < k1) {
			const int s0 = zslot(pc), sm = zslot(pc - 1), su = zslot(pc + 1);
			const int lx = T::H + (threadIdx.x & 63), ly = T::H + (threadIdx.x >> 6);
			const int gx = rx0 + lx, gy = ry0 + ly;
			if (gx < d.nx && gy < d.ny) {
				const int q = lx + T::RX * ly;
				const long long g = d.idx(gx, gy, pc);
				const uint8_t cc = cs[s0][q];
				float x = xs[s0][q];
				if (((gx + gy + pc) & 1) == 0 && l0_active(cc))
					x = (rs[s0][q] * l0_scale(cc) + zm_nb_sum(cs[s0], xs[s0], cs[sm], xs[sm], xs[su], q, T::RX)) /
					    l0_diag(cc);
				z[g] = x;
				acc += (double)(x * rs[s0][q]);
			}
		}
	}
	const double s = block_sum(acc, sh);
	if (threadIdx.x == 0) part[3 * blockIdx.x] = s;
}

// coarse-level brick: 32x4x4 vertices in 3D, 64x8 in 2D (512 threads; rows of
// 32/64 vertices keep the per-vertex stencil loads in whole cache lines)
constexpr int kLcBX3 = 32, kLcBY3 = 4, kLcBZ3 = 4, kLcBX2 = 64, kLcBY2 = 8;
static unsigned lc_blocks(const MgLevel& c, bool is3d) {
	const long long bx = is3d ? kLcBX3 : kLcBX2, by = is3d ? kLcBY3 : kLcBY2, bz = is3d ? kLcBZ3 : 1;
	return (unsigned)(((c.nx + bx - 1) / bx) * ((c.ny + by - 1) / by) * ((c.nz + bz - 1) / bz));
}

// Coarse-level block Gauss-Seidel: one sweep over all 8 (4 in 2D) colours of
// the 27-point (9-point) Galerkin stencil inside an LDS brick; halo values
// are the pre-sweep ones (block-hybrid smoother).  Forward colour order for
// pre-smoothing, reversed for post-smoothing, which keeps the V-cycle
// symmetric.  first: start from x = 0 (xin unused).  xcp != null: prolongate
// from the next coarser level on load (x = xin + I x_{l+1}).  The result goes
// to xout, never to xin: neighbouring bricks read xin as their halo.
template <bool IS3D>
__global__ void __launch_bounds__(512) k_lc_smooth_blk(const CgState* __restrict__ st, LvDev c, const float* __restrict__ A,
                                                       const uint8_t* __restrict__ type, const float* __restrict__ b,
                                                       const float* __restrict__ xin, float* __restrict__ xout, int first,
                                                       int reverse, LvDev cp, const uint8_t* __restrict__ ctp,
                                                       const float* __restrict__ xcp) {
	constexpr int BX = IS3D ? kLcBX3 : kLcBX2, BY = IS3D ? kLcBY3 : kLcBY2, BZ = IS3D ? kLcBZ3 : 1;
	constexpr int RX = BX + 2, RY = BY + 2, RZ = IS3D ? BZ + 2 : 1, RN = RX * RY * RZ, RXY = RX * RY;
	constexpr int S = IS3D ? 14 : 5, C = S - 1, NCOL = IS3D ? 8 : 4;
	__shared__ float xs[RN];
	if (st->done) return;
	int ox, oy, oz;
	brick_origin(c.nx, c.ny, BX, BY, BZ, ox, oy, oz);
	const int hz = IS3D ? 1 : 0;
	for (int q = threadIdx.x; q < RN; q += blockDim.x) {
		const int gx = ox - 1 + q % RX, gy = oy - 1 + (q / RX) % RY, gz = oz - hz + q / RXY;
		float v = 0.f;
		if (!first && gx >= 0 && gy >= 0 && gz >= 0 && gx < c.nx && gy < c.ny && gz < c.nz) {
			const long long g = gx + c.sy * gy + c.sz * gz;
			v = xin[g];
			if (xcp && type[g]) {
				float sum = 0.f;
				for (int pz = gz / 2; pz <= (gz + 1) / 2; pz++)
					for (int py = gy / 2; py <= (gy + 1) / 2; py++)
						for (int px = gx / 2; px <= (gx + 1) / 2; px++) {
							const long long pi = px + cp.sy * py + cp.sz * pz;
							if (ctp[pi]) sum += xcp[pi];
						}
				v += wgt(gx, gy, gz) * sum;
			}
		}
		xs[q] = v;
	}
	__syncthreads();
	const int tx = threadIdx.x % BX, ty = (threadIdx.x / BX) % BY, tz = threadIdx.x / (BX * BY);
	const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
	const bool valid = gx < c.nx && gy < c.ny && gz < c.nz;
	const long long v = valid ? gx + c.sy * gy + c.sz * gz : 0;
	const bool act = valid && type[v];
	const int me = (tx + 1) + RX * (ty + 1) + RXY * (tz + hz);
	float a[IS3D ? 27 : 9];
	float bv = 0.f, diag = 1.f;
	if (act) {
		bv = b[v];
		diag = A[v];
		int s = 0;
		for (int dz = (IS3D ? -1 : 0); dz <= (IS3D ? 1 : 0); dz++)
			for (int dy = -1; dy <= 1; dy++)
				for (int dx = -1; dx <= 1; dx++, s++) {
					float coef = 0.f;
					const int nx = gx + dx, ny = gy + dy, nz = gz + dz;
					if (s != C && nx >= 0 && ny >= 0 && nz >= 0 && nx < c.nx && ny < c.ny && nz < c.nz) {
						const long long nn = nx + c.sy * ny + c.sz * nz;
						coef = s > C ? A[(long long)(s - C) * c.n + v] : A[(long long)(C - s) * c.n + nn];
					}
					a[s] = coef;
				}
	}
	const int par = (gx & 1) | ((gy & 1) << 1) | ((gz & 1) << 2);
	for (int ph = 0; ph < NCOL; ph++) {
		const int col = reverse ? NCOL - 1 - ph : ph;
		if (act && par == col) {
			float sum = bv;
			int s = 0;
			for (int dz = (IS3D ? -1 : 0); dz <= (IS3D ? 1 : 0); dz++)
				for (int dy = -1; dy <= 1; dy++)
					for (int dx = -1; dx <= 1; dx++, s++)
						if (s != C) sum -= a[s] * xs[me + dx + RX * dy + RXY * dz];
			xs[me] = sum / diag;
		}
		__syncthreads();
	}
	if (valid) xout[v] = xs[me];
}

__global__ void k_fill_zero(const CgState* __restrict__ st, float* __restrict__ x, long long n) {
	if (st->done) return;
	for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
		x[v] = 0.f;
}

// coarsest level: x[cidx[i]] = sum_j G[i][j] * b[cidx[j]] (FP64), one wave per row;
// for a single-level hierarchy the rhs is the trivially scaled level-0 input
__global__ void k_mg_coarsest(const CgState* __restrict__ st, int nc, const int* __restrict__ cidx,
                              const double* __restrict__ G, const float* __restrict__ b,
                              const uint8_t* __restrict__ code, float* __restrict__ x) {
	if (st->done) return;
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
	if (row >= nc) return;
	double acc = 0.0;
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
	for (void* p : allocs) hipFree(p);
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
	(void)precision;
	const Dims& d = s->d;
	is3d = d.is3d != 0;
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
	// Cholesky with dropped (non-positive) pivots
	std::vector<double> Lm((size_t)m * m, 0.0);
	std::vector<char> dropped(m, 0);
	for (int jj = 0; jj < m; jj++) {
		double dval = M[(size_t)jj * m + jj];
		for (int q = 0; q < jj; q++) dval -= Lm[(size_t)jj * m + q] * Lm[(size_t)jj * m + q];
		const double orig = M[(size_t)jj * m + jj];
		if (!(dval > 1e-10 * (orig > 0 ? orig : 1.0))) {
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

int MgHierarchy::vcycle(Scene* s, const uint8_t* code, const float* r, float* z, const CgState* st) {
	const Dims& d = s->d;
	const int nl = (int)L.size();
	if (nl == 1) {
		hipLaunchKernelGGL(k_fill_zero, dim3(nblk(d.n)), dim3(256), 0, s->stream, st, z, d.n);
		const unsigned rows = (unsigned)((nc + 3) / 4);
		hipLaunchKernelGGL(k_mg_coarsest, dim3(rows > 0 ? rows : 1), dim3(256), 0, s->stream, st, nc,
		                   (const int*)cidx, (const double*)G, r, code, z);
		return 0;
	}
	// level-0 geometry: 3D z-marching 64x8 column tiles over chunks of an even
	// number of planes (about 1024 workgroups); 2D bricks of 64x16 cells
	constexpr int kZmTY = 8;
	const int BX = 64, BY = is3d ? kZmTY : 16;
	const long long cols = (long long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
	int kchunk = (d.nz + 1) & ~1;
	while (is3d && kchunk > 16 && cols * ((d.nz + kchunk - 1) / kchunk) < 1024) kchunk = ((kchunk / 2) + 1) & ~1;
	const unsigned nb0 = (unsigned)(is3d ? cols * ((d.nz + kchunk - 1) / kchunk) : cols);
	// down
	{
		KtScope kt("mg_l0_down", s->stream);
		if (is3d)
			hipLaunchKernelGGL(k_l0_down_zm<kZmTY>, dim3(nb0), dim3(64 * kZmTY), 0, s->stream, d, kchunk, st, code, r,
			                   L[0].x, L[0].r);
		else
			hipLaunchKernelGGL((k_l0_down<64, 16, 1, false>), dim3(nb0), dim3(1024), 0, s->stream, d, st, code, r, L[0].x,
			                   L[0].r);
	}
	for (int l = 1; l < nl; l++) {
		MgLevel& c = L[l];
		const MgLevel& f = L[l - 1];
		hipLaunchKernelGGL(k_mg_restrict, dim3(nblk(c.n)), dim3(256), 0, s->stream, st, lvdev(c), lvdev(f), (int)is3d,
		                   (const uint8_t*)c.type, (const float*)f.r, c.b);
		if (l == nl - 1) break;
		const unsigned nbl = lc_blocks(c, is3d);
		if (is3d)
			hipLaunchKernelGGL(k_lc_smooth_blk<true>, dim3(nbl), dim3(512), 0, s->stream, st, lvdev(c), (const float*)c.A,
			                   (const uint8_t*)c.type, (const float*)c.b, (const float*)nullptr, c.x, 1, 0, lvdev(c),
			                   (const uint8_t*)nullptr, (const float*)nullptr);
		else
			hipLaunchKernelGGL(k_lc_smooth_blk<false>, dim3(nbl), dim3(512), 0, s->stream, st, lvdev(c), (const float*)c.A,
			                   (const uint8_t*)c.type, (const float*)c.b, (const float*)nullptr, c.x, 1, 0, lvdev(c),
			                   (const uint8_t*)nullptr, (const float*)nullptr);
		hipLaunchKernelGGL(k_lc_residual, dim3(nblk(c.n)), dim3(256), 0, s->stream, st, lvdev(c), (int)is3d, S,
		                   (const float*)c.A, (const uint8_t*)c.type, (const float*)c.b, (const float*)c.x, c.r);
	}
	// coarsest
	{
		MgLevel& c = L[nl - 1];
		const unsigned rows = (unsigned)((nc + 3) / 4);
		hipLaunchKernelGGL(k_mg_coarsest, dim3(rows > 0 ? rows : 1), dim3(256), 0, s->stream, st, nc,
		                   (const int*)cidx, (const double*)G, (const float*)c.b, (const uint8_t*)nullptr, c.x);
	}
	// up: prolongation fused into each post-smoothing sweep
	for (int l = nl - 2; l >= 1; l--) {
		const MgLevel& c = L[l + 1];
		MgLevel& f = L[l];
		const unsigned nbl = lc_blocks(f, is3d);
		if (is3d)
			hipLaunchKernelGGL(k_lc_smooth_blk<true>, dim3(nbl), dim3(512), 0, s->stream, st, lvdev(f), (const float*)f.A,
			                   (const uint8_t*)f.type, (const float*)f.b, (const float*)f.x, f.xp, 0, 1, lvdev(c),
			                   (const uint8_t*)c.type, (const float*)(l + 1 == nl - 1 ? c.x : c.xp));
		else
			hipLaunchKernelGGL(k_lc_smooth_blk<false>, dim3(nbl), dim3(512), 0, s->stream, st, lvdev(f), (const float*)f.A,
			                   (const uint8_t*)f.type, (const float*)f.b, (const float*)f.x, f.xp, 0, 1, lvdev(c),
			                   (const uint8_t*)c.type, (const float*)(l + 1 == nl - 1 ? c.x : c.xp));
	}
	{
		KtScope kt("mg_l0_up", s->stream);
		const MgLevel& c = L[1];
		const float* xc1 = nl == 2 ? c.x : c.xp;
		if (is3d)
			hipLaunchKernelGGL(k_l0_up_zm<kZmTY>, dim3(nb0), dim3(64 * kZmTY), 0, s->stream, d, kchunk, st, code, r,
			                   (const float*)L[0].x, z, lvdev(c), xc1, s->dpart);
		else
			hipLaunchKernelGGL((k_l0_up<64, 16, 1, false>), dim3(nb0), dim3(1024), 0, s->stream, d, st, code, r,
			                   (const float*)L[0].x, z, lvdev(c), (const uint8_t*)c.type, (const float*)xc1, s->dpart);
	}
	return (int)nb0;
}

}  // namespace mfx
