// Semi-Lagrangian / MacCormack advection, wall boundary conditions and
// buoyancy on device-resident grids.
//
// Arithmetic mirrors the reference's C++ expressions operation by operation
// (compiled with -ffp-contract=off), so results are bit-identical to the
// reference's serial FP32 plugins:
//   advection.cpp:25-460, util/interpol.h:50-90, grid.h:424-470,
//   extforces.cpp:72-87,186-217.
// Kernels are laid out for gfx950: one 64-wide wavefront per x-row segment
// (coalesced SoA loads), 64x4 tiles, XCD-contiguous tile order; trilinear
// gathers are served from L1/L2.  The MacCormack backward step, correction
// and clamp are fused into one pass (the backward value of a cell is only
// needed at that cell), so order-2 advection is two full passes.
#include <type_traits>

#include "mfx_internal.h"
#include "mfx_prof.h"

namespace mfx {

struct Vel {
	const float* __restrict__ u;
	const float* __restrict__ v;
	const float* __restrict__ w;
};

__device__ inline bool interior(const Dims& d, int i, int j, int k) {
	return i >= 1 && i < d.nx - 1 && j >= 1 && j < d.ny - 1 && (d.is3d ? (k >= 1 && k < d.nz - 1) : true);
}

// Cell geometry of the advection kernels.  IS3D is the grid's dimension,
// fixed at compile time (2D grids have z stride 0, as the reference's grids
// do).  I is the index type: int when the grid has fewer than 2^31 cells and
// nx and ny*nz are below 2^24 (checked on the host; two 24-bit multiply-adds
// then give the exact index), long long otherwise.
template <bool IS3D, typename I>
struct Geo {
	int nx, ny, nz;
	I Y, Z;
	__device__ explicit Geo(const Dims& d) : nx(d.nx), ny(d.ny), nz(d.nz), Y((I)d.nx), Z(IS3D ? (I)d.sz : (I)0) {}
	__device__ I at(int i, int j, int k) const {
		if constexpr (sizeof(I) == 4) {
			const unsigned row = IS3D ? __umul24((unsigned)k, (unsigned)ny) + (unsigned)j : (unsigned)j;
			return (I)(__umul24(row, (unsigned)nx) + (unsigned)i);
		} else {
			return (I)i + Y * j + Z * k;
		}
	}
	__device__ bool interior(int i, int j, int k) const {
		return i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1 && (IS3D ? (k >= 1 && k < nz - 1) : true);
	}
};

inline bool index32_ok(const Dims& d) {
	return d.n < (1ll << 31) && d.nx < (1 << 24) && (long long)d.ny * d.nz < (1ll << 24);
}

// MACGrid::getCentered (grid.h:424-437)
template <bool IS3D, typename I>
__device__ inline void getCentered(const Vel& V, const Geo<IS3D, I>& g, I ix, float* o) {
	o[0] = 0.5f * (V.u[ix] + V.u[ix + 1]);
	o[1] = 0.5f * (V.v[ix] + V.v[ix + g.Y]);
	o[2] = 0.f;
	if (IS3D) o[2] = 0.5f * (V.w[ix] + V.w[ix + g.Z]);
}
// MACGrid::getAtMACX/Y/Z (grid.h:439-470); Z is not guarded by is3D (in 2D
// its z offsets are 0)
template <bool IS3D, typename I>
__device__ inline void getAtMACX(const Vel& V, const Geo<IS3D, I>& g, I ix, float* o) {
	const float* __restrict__ v = V.v + ix;
	const float* __restrict__ w = V.w + ix;
	const I Y = g.Y, Z = g.Z;
	o[0] = V.u[ix];
	o[1] = 0.25f * (v[0] + v[-1] + v[Y] + v[Y - 1]);
	o[2] = 0.f;
	if (IS3D) o[2] = 0.25f * (w[0] + w[-1] + w[Z] + w[Z - 1]);
}
template <bool IS3D, typename I>
__device__ inline void getAtMACY(const Vel& V, const Geo<IS3D, I>& g, I ix, float* o) {
	const float* __restrict__ u = V.u + ix;
	const float* __restrict__ w = V.w + ix;
	const I Y = g.Y, Z = g.Z;
	o[0] = 0.25f * (u[0] + u[-Y] + u[1] + u[1 - Y]);
	o[1] = V.v[ix];
	o[2] = 0.f;
	if (IS3D) o[2] = 0.25f * (w[0] + w[-Y] + w[Z] + w[Z - Y]);
}
template <bool IS3D, typename I>
__device__ inline void getAtMACZ(const Vel& V, const Geo<IS3D, I>& g, I ix, float* o) {
	const float* __restrict__ u = V.u + ix;
	const float* __restrict__ v = V.v + ix;
	const I Y = g.Y, Z = g.Z;
	o[0] = 0.25f * (u[0] + u[-Z] + u[1] + u[1 - Z]);
	o[1] = 0.25f * (v[0] + v[-Z] + v[Y] + v[Y - Z]);
	o[2] = V.w[ix];
}

// Two x-adjacent floats in one 8-byte load (addresses are 4-byte aligned:
// a dword-aligned multi-dword global load is a single global_load_dwordx2)
struct __attribute__((packed, aligned(4))) FPair {
	float a, b;
};
__device__ inline FPair ld_pair(const float* __restrict__ p) { return *reinterpret_cast<const FPair*>(p); }

// interpol<Real> / interpolComponent (interpol.h:50-90).  The eight corners
// are read at one base index plus constant strides, x-adjacent pairs in one
// load each.
template <bool IS3D, typename I>
__device__ inline float interp(const float* __restrict__ data, const Geo<IS3D, I>& g, float posx, float posy,
                               float posz) {
	float px = posx - 0.5f, py = posy - 0.5f, pz = posz - 0.5f;
	int xi = (int)px, yi = (int)py, zi = (int)pz;
	float s1 = px - (float)xi;
	float t1 = py - (float)yi;
	float f1 = pz - (float)zi;
	if (px < 0.f) { xi = 0; s1 = 0.0f; }
	if (py < 0.f) { yi = 0; t1 = 0.0f; }
	if (pz < 0.f) { zi = 0; f1 = 0.0f; }
	if (xi >= g.nx - 1) { xi = g.nx - 2; s1 = 1.0f; }
	if (yi >= g.ny - 1) { yi = g.ny - 2; t1 = 1.0f; }
	if (g.nz > 1) {
		if (zi >= g.nz - 1) { zi = g.nz - 2; f1 = 1.0f; }
	}
	// the reference sets (s0, s1) to (1, 0) / (0, 1) in the clamps and to
	// (1 - s1, s1) otherwise: 1 - s1 is exact in all three cases
	const float s0 = 1.0f - s1, t0 = 1.0f - t1, f0 = 1.0f - f1;
	const I Y = g.Y, Z = g.Z;
	const float* __restrict__ p = data + g.at(xi, yi, IS3D ? zi : 0);
	const FPair c00 = ld_pair(p), c10 = ld_pair(p + Y), c01 = ld_pair(p + Z), c11 = ld_pair(p + Y + Z);
	return ((c00.a * t0 + c10.a * t1) * s0 + (c00.b * t0 + c10.b * t1) * s1) * f0 +
	       ((c01.a * t0 + c11.a * t1) * s0 + (c01.b * t0 + c11.b * t1) * s1) * f1;
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ inline float clampf_ref(float v, float lo, float hi) {
	if (v < lo) return lo;
	if (v > hi) return hi;
	return v;
}
// getMinMax (advection.cpp): strict comparisons, first of equal values kept
__device__ inline void min_max(float v, float& mn, float& mx) {
	if (v < mn) mn = v;
	if (v > mx) mx = v;
}

#define CHECKFLAG(ix) (f8[ix] & (kFluid | kEmpty))

// ---------------------------------------------------------------------------
// forward SL step for centered grids (1 or 3 components): dst = 0 on the
// outer layer (the reference writes a zeroed temporary with bnd=1)
// ---------------------------------------------------------------------------
template <int NC, bool IS3D, typename I>
__global__ void __launch_bounds__(kBlock) k_sl_centered(Vel V, Dims d, Tiles t, float dt, const float* __restrict__ s0,
                                                        const float* __restrict__ s1, const float* __restrict__ s2,
                                                        float* __restrict__ o0, float* __restrict__ o1,
                                                        float* __restrict__ o2) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(p.i, p.j, p.k);
	if (!g.interior(p.i, p.j, p.k)) {
		o0[ix] = 0.f;
		if (NC == 3) {
			o1[ix] = 0.f;
			o2[ix] = 0.f;
		}
		return;
	}
	float c[3];
	getCentered(V, g, ix, c);
	const float px = ((float)p.i + 0.5f) - c[0] * dt;
	const float py = ((float)p.j + 0.5f) - c[1] * dt;
	const float pz = ((float)p.k + 0.5f) - c[2] * dt;
	o0[ix] = interp(s0, g, px, py, pz);
	if (NC == 3) {
		o1[ix] = interp(s1, g, px, py, pz);
		o2[ix] = interp(s2, g, px, py, pz);
	}
}

// SemiLagrangeMAC (advection.cpp:44-57) of one cell from the face velocities
// vx/vy/vz (getAtMACX/Y/Z at the cell)
template <bool IS3D, typename I>
__device__ inline void sl_mac_cell(const Geo<IS3D, I>& g, int i, int j, int k, float dt, const float* vx,
                                   const float* vy, const float* vz, const float* __restrict__ su,
                                   const float* __restrict__ sv, const float* __restrict__ sw, float* out) {
	const float p0x = (float)i + 0.5f, p0y = (float)j + 0.5f, p0z = (float)k + 0.5f;
	out[0] = interp(su, g, p0x - vx[0] * dt, p0y - vx[1] * dt, p0z - vx[2] * dt);
	out[1] = interp(sv, g, p0x - vy[0] * dt, p0y - vy[1] * dt, p0z - vy[2] * dt);
	out[2] = interp(sw, g, p0x - vz[0] * dt, p0y - vz[1] * dt, p0z - vz[2] * dt);
}

template <bool IS3D, typename I>
__global__ void __launch_bounds__(kBlock) k_sl_mac(Vel V, Dims d, Tiles t, float dt, const float* __restrict__ su,
                                                   const float* __restrict__ sv, const float* __restrict__ sw,
                                                   float* __restrict__ ou, float* __restrict__ ov,
                                                   float* __restrict__ ow) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(p.i, p.j, p.k);
	if (!g.interior(p.i, p.j, p.k)) {
		ou[ix] = 0.f;
		ov[ix] = 0.f;
		ow[ix] = 0.f;
		return;
	}
	float vx[3], vy[3], vz[3], out[3];
	getAtMACX(V, g, ix, vx);
	getAtMACY(V, g, ix, vy);
	getAtMACZ(V, g, ix, vz);
	sl_mac_cell(g, p.i, p.j, p.k, dt, vx, vy, vz, su, sv, sw, out);
	ou[ix] = out[0];
	ov[ix] = out[1];
	ow[ix] = out[2];
}

// ---------------------------------------------------------------------------
// orderSpace 2 (Catmull-Rom cubic lookups) and orderTrace 2 (explicit
// midpoint trace).  Rarely used options: one plain gather kernel per SL step,
// the MacCormack correction and clamp then read the backward values from
// memory (k_maccormack_* with PRE).
// ---------------------------------------------------------------------------

// cubicInterp<Real> (util/interpolHigh.h:22-39): the coefficient expressions
// are double (Real * double) and stored as float; the polynomial is float
__device__ inline float cubic_real(float t, const float* p) {
	const float d0 = (float)((double)(p[2] - p[0]) * 0.5);
	const float d1 = (float)((double)(p[3] - p[1]) * 0.5);
	const float dk = p[2] - p[1];
	const float a2 = (float)(3.0 * (double)dk - 2.0 * (double)d0 - (double)d1);
	const float a3 = (float)(-2.0 * (double)dk + (double)d0 + (double)d1);
	const float sq = t * t, cu = sq * t;
	return a3 * cu + a2 * sq + d0 * t + p[1];
}
// cubicInterp<Vec3> on one component: every Vec3 * double product is
// narrowed on its own and Vec3 sums are float (util/vectorbase.h:235-279)
__device__ inline float cubic_vec(float t, const float* p) {
	const float d0 = (float)((double)(p[2] - p[0]) * 0.5);
	const float d1 = (float)((double)(p[3] - p[1]) * 0.5);
	const float dk = p[2] - p[1];
	const float a2 = ((float)(3.0 * (double)dk) - (float)(2.0 * (double)d0)) - d1;
	const float a3 = ((float)(-2.0 * (double)dk) + d0) + d1;
	const float sq = t * t, cu = sq * t;
	return a3 * cu + a2 * sq + d0 * t + p[1];
}

// interpolCubic / interpolCubic2D (util/interpolHigh.h:41-170) of one
// component; VEC selects the Vec3 rounding rules.  A 4x4(x4) support that
// leaves the grid falls back to interpol (:58-60, :97-99).  The support test
// is written on x1 = (int)px itself (x1 - 1 >= 0 and x1 + 2 < nx), which is
// the reference's test without its integer overflow for huge positions.
template <bool VEC, bool IS3D, typename I>
__device__ inline float cubic_lookup(const float* __restrict__ data, const Geo<IS3D, I>& g, float posx, float posy,
                                     float posz) {
	const float px = posx - 0.5f, py = posy - 0.5f;
	const int x1 = (int)px, y1 = (int)py;
	const I Y = g.Y;
	if (!IS3D) {
		if (x1 < 1 || y1 < 1 || x1 > g.nx - 3 || y1 > g.ny - 3) return interp(data, g, posx, posy, posz);
		const float xt = px - (float)x1, yt = py - (float)y1;
		const float* __restrict__ b = data + g.at(x1 - 1, y1 - 1, 0);
		float rows[4];
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const float pts[4] = {b[r * Y], b[r * Y + 1], b[r * Y + 2], b[r * Y + 3]};
			rows[r] = VEC ? cubic_vec(xt, pts) : cubic_real(xt, pts);
		}
		return VEC ? cubic_vec(yt, rows) : cubic_real(yt, rows);
	}
	const float pz = posz - 0.5f;
	const int z1 = (int)pz;
	if (x1 < 1 || y1 < 1 || z1 < 1 || x1 > g.nx - 3 || y1 > g.ny - 3 || z1 > g.nz - 3)
		return interp(data, g, posx, posy, posz);
	const float xt = px - (float)x1, yt = py - (float)y1, zt = pz - (float)z1;
	const I Z = g.Z;
	const float* __restrict__ b = data + g.at(x1 - 1, y1 - 1, z1 - 1);
	float slabs[4];
#pragma unroll
	for (int c = 0; c < 4; c++) {
		float rows[4];
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const float* __restrict__ q = b + c * Z + r * Y;
			const float pts[4] = {q[0], q[1], q[2], q[3]};
			rows[r] = VEC ? cubic_vec(xt, pts) : cubic_real(xt, pts);
		}
		slabs[c] = VEC ? cubic_vec(yt, rows) : cubic_real(yt, rows);
	}
	return VEC ? cubic_vec(zt, slabs) : cubic_real(zt, slabs);
}

// BUILD_INDEX of one axis (util/interpol.h:50-66): cell and upper weight
__device__ inline void build_axis(float p, int n, bool clampHi, int& xi, float& s1) {
	xi = (int)p;
	s1 = p - (float)xi;
	if (p < 0.f) {
		xi = 0;
		s1 = 0.f;
	}
	if (clampHi && xi >= n - 1) {
		xi = n - 2;
		s1 = 1.f;
	}
}

// interpolMAC (util/interpol.h:121-157): every component of a MAC grid at
// pos, each with the shifted (BUILD_INDEX_SHIFT) weights along its own axis.
// In 2D the lookups stay in plane 0 (the reference would read outside the
// grid for a z position of 1.5 or more there).
template <bool IS3D, typename I>
__device__ inline void interp_mac(const Vel& S, const Geo<IS3D, I>& g, float x, float y, float z, float* out) {
	int xi, yi, zi, sx, sy, sz;
	float s1, t1, f1, ss1, st1, sf1;
	build_axis(x - 0.5f, g.nx, true, xi, s1);
	build_axis(y - 0.5f, g.ny, true, yi, t1);
	build_axis(z - 0.5f, g.nz, g.nz > 1, zi, f1);
	build_axis(x, g.nx, true, sx, ss1);
	build_axis(y, g.ny, true, sy, st1);
	build_axis(z, g.nz, g.nz > 1, sz, sf1);
	const float s0 = 1.0f - s1, t0 = 1.0f - t1, f0 = 1.0f - f1;
	const float ss0 = 1.0f - ss1, st0 = 1.0f - st1, sf0 = 1.0f - sf1;
	const I Y = g.Y, Z = g.Z;
	{
		const float* __restrict__ r = S.u + g.at(sx, yi, IS3D ? zi : 0);
		out[0] = f0 * ((r[0] * t0 + r[Y] * t1) * ss0 + (r[1] * t0 + r[1 + Y] * t1) * ss1) +
		         f1 * ((r[Z] * t0 + r[Z + Y] * t1) * ss0 + (r[1 + Z] * t0 + r[1 + Y + Z] * t1) * ss1);
	}
	{
		const float* __restrict__ r = S.v + g.at(xi, sy, IS3D ? zi : 0);
		out[1] = f0 * ((r[0] * st0 + r[Y] * st1) * s0 + (r[1] * st0 + r[1 + Y] * st1) * s1) +
		         f1 * ((r[Z] * st0 + r[Z + Y] * st1) * s0 + (r[1 + Z] * st0 + r[1 + Y + Z] * st1) * s1);
	}
	{
		const float* __restrict__ r = S.w + g.at(xi, yi, IS3D ? sz : 0);
		out[2] = sf0 * ((r[0] * t0 + r[Y] * t1) * s0 + (r[1] * t0 + r[1 + Y] * t1) * s1) +
		         sf1 * ((r[Z] * t0 + r[Z + Y] * t1) * s0 + (r[1 + Z] * t0 + r[1 + Y + Z] * t1) * s1);
	}
}

// SemiLagrange<T> with orderTrace 1/2 and orderSpace 1/2 (advection.cpp:25-41,
// grid.h:146-151): the midpoint trace takes p1 = p0 - getCentered*dt*0.5 and
// the interpolated velocity there; NC == 3 uses the Vec3 cubic rules
template <int NC, bool IS3D, typename I>
__global__ void __launch_bounds__(kBlock) k_sl_centered_hi(Vel V, Dims d, Tiles t, float dt, int orderSpace,
                                                           int orderTrace, const float* __restrict__ s0,
                                                           const float* __restrict__ s1, const float* __restrict__ s2,
                                                           float* __restrict__ o0, float* __restrict__ o1,
                                                           float* __restrict__ o2) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(p.i, p.j, p.k);
	float* dst[3] = {o0, o1, o2};
	if (!g.interior(p.i, p.j, p.k)) {
		for (int c = 0; c < NC; c++) dst[c][ix] = 0.f;
		return;
	}
	float v[3];
	getCentered(V, g, ix, v);
	const float p0[3] = {(float)p.i + 0.5f, (float)p.j + 0.5f, (float)p.k + 0.5f};
	float pos[3];
	if (orderTrace == 2) {
		float p1[3], u[3];
		for (int a = 0; a < 3; a++) p1[a] = p0[a] - (float)((double)(v[a] * dt) * 0.5);
		interp_mac(V, g, p1[0], p1[1], p1[2], u);
		for (int a = 0; a < 3; a++) pos[a] = p0[a] - u[a] * dt;
	} else {
		for (int a = 0; a < 3; a++) pos[a] = p0[a] - v[a] * dt;
	}
	const float* src[3] = {s0, s1, s2};
	for (int c = 0; c < NC; c++)
		dst[c][ix] = orderSpace == 2 ? cubic_lookup<NC == 3>(src[c], g, pos[0], pos[1], pos[2])
		                             : interp(src[c], g, pos[0], pos[1], pos[2]);
}

// SemiLagrangeMAC with orderTrace 1/2 and orderSpace 1/2 (advection.cpp:44-77).
// The midpoint trace starts at the face centre and takes the face velocity
// and the interpolated velocity from S, the field being advected (:59-73).
// orderSpace 2 is interpolCubicMAC(pos)[c] = interpolCubic<Vec3>(pos + e_c/2)[c],
// 0 for z in 2D (grid.h:268-273, interpolHigh.h:172-179).
template <bool IS3D, typename I>
__global__ void __launch_bounds__(kBlock) k_sl_mac_hi(Vel V, Dims d, Tiles t, float dt, int orderSpace,
                                                      int orderTrace, Vel S, float* __restrict__ ou,
                                                      float* __restrict__ ov, float* __restrict__ ow) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(p.i, p.j, p.k);
	if (!g.interior(p.i, p.j, p.k)) {
		ou[ix] = 0.f;
		ov[ix] = 0.f;
		ow[ix] = 0.f;
		return;
	}
	const Vel& F = orderTrace == 2 ? S : V;
	float vf[3][3];
	getAtMACX(F, g, ix, vf[0]);
	getAtMACY(F, g, ix, vf[1]);
	getAtMACZ(F, g, ix, vf[2]);
	const float h[3] = {(float)p.i + 0.5f, (float)p.j + 0.5f, (float)p.k + 0.5f};
	const float cell[3] = {(float)p.i, (float)p.j, (float)p.k};
	const float* src[3] = {S.u, S.v, S.w};
	float out[3];
	for (int c = 0; c < 3; c++) {
		float pos[3];
		if (orderTrace == 2) {
			float p1[3], u[3];
			for (int a = 0; a < 3; a++) p1[a] = (a == c ? cell[a] : h[a]) - (float)((double)(vf[c][a] * dt) * 0.5);
			interp_mac(S, g, p1[0], p1[1], p1[2], u);
			for (int a = 0; a < 3; a++) pos[a] = h[a] - u[a] * dt;
		} else {
			for (int a = 0; a < 3; a++) pos[a] = h[a] - vf[c][a] * dt;
		}
		if (orderSpace == 2) {
			if (c == 2 && !IS3D) {
				out[c] = 0.f;
			} else {
				pos[c] = pos[c] + 0.5f;
				out[c] = cubic_lookup<true>(src[c], g, pos[0], pos[1], pos[2]);
			}
		} else {
			out[c] = interp(src[c], g, pos[0], pos[1], pos[2]);
		}
	}
	ou[ix] = out[0];
	ov[ix] = out[1];
	ow[ix] = out[2];
}

// ---------------------------------------------------------------------------
// MacCormack for centered grids: backward SL (fwd, -dt), correction
// (MacCormackCorrect, all idx) and clamp (MacCormackClamp, bnd=1) fused.
// ---------------------------------------------------------------------------
// With PRE the backward values come from pb0..pb2 (a backward SL step of
// orderSpace / orderTrace 2 run as its own pass, zero on the outer layer).
template <int NC, bool IS3D, typename I, bool PRE = false>
__global__ void __launch_bounds__(kBlock) k_maccormack_centered(Vel V, Dims d, Tiles t, float dt, float strength,
                                                                int clampMode, const uint8_t* __restrict__ f8,
                                                                const float* o0, const float* o1, const float* o2,
                                                                const float* __restrict__ fw0,
                                                                const float* __restrict__ fw1,
                                                                const float* __restrict__ fw2, float* __restrict__ n0,
                                                                float* __restrict__ n1, float* __restrict__ n2,
                                                                const float* __restrict__ pb0 = nullptr,
                                                                const float* __restrict__ pb1 = nullptr,
                                                                const float* __restrict__ pb2 = nullptr) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(p.i, p.j, p.k);
	const bool inner = g.interior(p.i, p.j, p.k);
	const float* orig[3] = {o0, o1, o2};
	const float* fwd[3] = {fw0, fw1, fw2};
	float* dst[3] = {n0, n1, n2};
	float vc[3] = {0.f, 0.f, 0.f};
	float px = 0.f, py = 0.f, pz = 0.f;
	if (inner) {
		getCentered(V, g, ix, vc);
		// bwd position uses -dt: pos - v*(-dt)
		const float mdt = -dt;
		px = ((float)p.i + 0.5f) - vc[0] * mdt;
		py = ((float)p.j + 0.5f) - vc[1] * mdt;
		pz = ((float)p.k + 0.5f) - vc[2] * mdt;
	}
	const bool fl = f8[ix] & kFluid;
	float val[3], fv[3];
#pragma unroll
	for (int c = 0; c < NC; c++) {
		float bwd;
		if constexpr (PRE) bwd = (c == 0 ? pb0 : (c == 1 ? pb1 : pb2))[ix];
		else bwd = inner ? interp(fwd[c], g, px, py, pz) : 0.f;
		fv[c] = fwd[c][ix];
		float x = fv[c];
		if (fl) {
			if (NC == 1) {
				// float "+=" double expression (advection.cpp:89)
				x = (float)((double)x + (double)strength * 0.5 * (double)(orig[c][ix] - bwd));
			} else {
				// Vec3: double product narrowed per component, float add
				x = x + (float)((double)strength * 0.5 * (double)(orig[c][ix] - bwd));
			}
		}
		val[c] = x;
	}
	if (inner) {
		// doClampComponent (advection.cpp:144-186) with pos = (i,j,k)
		const float vdt[3] = {vc[0] * dt, vc[1] * dt, vc[2] * dt};
		const float pos[3] = {(float)p.i, (float)p.j, (float)p.k};
		const int gux = g.nx - 1, guy = g.ny - 1, guz = g.nz - 1;
		float mn[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
		float mx[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
		bool haveFl = false;
		const I Y = g.Y, Z = g.Z;
#pragma unroll
		for (int l = 0; l < 2; ++l) {
			if (l == 1 && clampMode != 1) break;
			int P[3];
			for (int a = 0; a < 3; a++) P[a] = (int)(l == 0 ? pos[a] - vdt[a] : pos[a] + vdt[a]);
			const int i0 = clampi(P[0], 0, gux - 1), j0 = clampi(P[1], 0, guy - 1);
			const I c0 = g.at(i0, j0, IS3D ? clampi(P[2], 0, guz - 1) : 0);
			// corners (i0|i1, j0|j1, k0|k1) in the reference's order; every
			// corner address is inside the grid, so flags and values are all
			// loaded before the flag tests
			constexpr int NQ = IS3D ? 8 : 4;
			const I off[8] = {0, 1, Y, 1 + Y, Z, 1 + Z, Y + Z, 1 + Y + Z};
			bool cf[NQ];
			float cv[NQ][NC];
#pragma unroll
			for (int q = 0; q < NQ; q++) cf[q] = CHECKFLAG(c0 + off[q]);
#pragma unroll
			for (int q = 0; q < NQ; q += 2) {  // x-adjacent corner pairs
#pragma unroll
				for (int c = 0; c < NC; c++) {
					const FPair e = ld_pair(orig[c] + c0 + off[q]);
					cv[q][c] = e.a;
					cv[q + 1][c] = e.b;
				}
			}
#pragma unroll
			for (int q = 0; q < NQ; q++) {
				if (cf[q]) {
					for (int c = 0; c < NC; c++) min_max(cv[q][c], mn[c], mx[c]);
					haveFl = true;
				}
			}
		}
		if (!haveFl) {
			for (int c = 0; c < NC; c++) val[c] = fv[c];
		} else if (clampMode == 1) {
			for (int c = 0; c < NC; c++) val[c] = clampf_ref(val[c], mn[c], mx[c]);
		} else {
			bool out = false;
			for (int c = 0; c < NC; c++) out = out || (val[c] < mn[c]) || (val[c] > mx[c]);
			if (out)
				for (int c = 0; c < NC; c++) val[c] = fv[c];
		}
		if (clampMode == 1) {
			int pf[3], pb[3];
			for (int a = 0; a < 3; a++) {
				const float base = pos[a] + 0.5f;
				pf[a] = (int)(base - vdt[a]);
				pb[a] = (int)(base + vdt[a]);
			}
			if (pf[0] < 0 || pf[1] < 0 || pf[2] < 0 || pb[0] < 0 || pb[1] < 0 || pb[2] < 0 || pf[0] > gux ||
			    pf[1] > guy || ((pf[2] > guz) && IS3D) || pb[0] > gux || pb[1] > guy || ((pb[2] > guz) && IS3D) ||
			    (f8[g.at(pf[0], pf[1], IS3D ? pf[2] : 0)] & kObstacle) ||
			    (f8[g.at(pb[0], pb[1], IS3D ? pb[2] : 0)] & kObstacle)) {
				for (int c = 0; c < NC; c++) val[c] = fv[c];
			}
		}
	}
	for (int c = 0; c < NC; c++) dst[c][ix] = val[c];
}

// k_maccormack_centered for clampMode 2 on grids of at least 3 cells per
// axis: the same arithmetic without divergent branches, as k_maccormack_mac2
// (the trace runs at the interior-clamped cell and is discarded outside the
// interior; the clamp corners load with the backward trace).  The bounds
// take only fluid / empty corners, as getMinMax behind CHECKFLAG does: a
// corner that fails the flag test contributes +-FLT_MAX, the seeds, which
// leave fminf / fmaxf unchanged.  In clampMode 2 the bounds are only
// compared with, so the sign of a zero bound (the one thing fminf / fmaxf may
// choose differently from the strict comparisons) is never seen.
template <int NC, bool IS3D, typename I>
__global__ void __launch_bounds__(kBlock) k_maccormack_centered2(Vel V, Dims d, Tiles t, float dt, float strength,
                                                                 const uint8_t* __restrict__ f8, const float* o0,
                                                                 const float* o1, const float* o2,
                                                                 const float* __restrict__ fw0,
                                                                 const float* __restrict__ fw1,
                                                                 const float* __restrict__ fw2, float* __restrict__ n0,
                                                                 float* __restrict__ n1, float* __restrict__ n2) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(p.i, p.j, p.k);
	const bool inner = g.interior(p.i, p.j, p.k);
	const int ic = clampi(p.i, 1, g.nx - 2), jc = clampi(p.j, 1, g.ny - 2), kc = IS3D ? clampi(p.k, 1, g.nz - 2) : p.k;
	const I ixc = g.at(ic, jc, kc);
	const float* orig[3] = {o0, o1, o2};
	const float* fwd[3] = {fw0, fw1, fw2};
	float* dst[3] = {n0, n1, n2};
	const bool fl = f8[ix] & kFluid;
	float fv[NC], org[NC];
#pragma unroll
	for (int c = 0; c < NC; c++) {
		fv[c] = fwd[c][ix];
		org[c] = orig[c][ix];
	}
	float vc[3];
	getCentered(V, g, ixc, vc);
	// bwd position uses -dt: pos - v*(-dt)
	const float mdt = -dt;
	const float px = ((float)ic + 0.5f) - vc[0] * mdt;
	const float py = ((float)jc + 0.5f) - vc[1] * mdt;
	const float pz = ((float)kc + 0.5f) - vc[2] * mdt;
	// doClampComponent (advection.cpp:144-186) corners at toVec3i(pos - v dt)
	const float vdt[3] = {vc[0] * dt, vc[1] * dt, vc[2] * dt};
	const int P0 = (int)((float)ic - vdt[0]), P1 = (int)((float)jc - vdt[1]), P2 = (int)((float)kc - vdt[2]);
	const int i0 = min(max(P0, 0), g.nx - 2), j0 = min(max(P1, 0), g.ny - 2);
	const I c0 = g.at(i0, j0, IS3D ? min(max(P2, 0), g.nz - 2) : 0);
	constexpr int NQ = IS3D ? 8 : 4;
	const I Y = g.Y, Z = g.Z;
	const I off[8] = {0, 1, Y, 1 + Y, Z, 1 + Z, Y + Z, 1 + Y + Z};
	bool cf[NQ];
	float cv[NQ][NC];
#pragma unroll
	for (int q = 0; q < NQ; q++) cf[q] = CHECKFLAG(c0 + off[q]);
#pragma unroll
	for (int q = 0; q < NQ; q += 2) {
#pragma unroll
		for (int c = 0; c < NC; c++) {
			const FPair e = ld_pair(orig[c] + c0 + off[q]);
			cv[q][c] = e.a;
			cv[q + 1][c] = e.b;
		}
	}
	float bwd[NC];
#pragma unroll
	for (int c = 0; c < NC; c++) bwd[c] = interp(fwd[c], g, px, py, pz);
	bool haveFl = false;
	float mn[NC], mx[NC];
#pragma unroll
	for (int c = 0; c < NC; c++) {
		mn[c] = 3.402823466e+38f;
		mx[c] = -3.402823466e+38f;
	}
#pragma unroll
	for (int q = 0; q < NQ; q++) {
		haveFl = haveFl || cf[q];
#pragma unroll
		for (int c = 0; c < NC; c++) {
			mn[c] = fminf(mn[c], cf[q] ? cv[q][c] : 3.402823466e+38f);
			mx[c] = fmaxf(mx[c], cf[q] ? cv[q][c] : -3.402823466e+38f);
		}
	}
	float val[NC];
	bool out = false;
#pragma unroll
	for (int c = 0; c < NC; c++) {
		const float b = inner ? bwd[c] : 0.f;
		float x;
		if (NC == 1) {
			// float "+=" double expression (advection.cpp:89)
			x = (float)((double)fv[c] + (double)strength * 0.5 * (double)(org[c] - b));
		} else {
			// Vec3: double product narrowed per component, float add
			x = fv[c] + (float)((double)strength * 0.5 * (double)(org[c] - b));
		}
		// evaluated for every lane and selected; the empty asm keeps the
		// select from becoming a branch around the load of org
		asm("" : "+v"(x));
		x = fl ? x : fv[c];
		val[c] = x;
		out = out || (x < mn[c]) || (x > mx[c]);
	}
	// clamp (interior only): no fluid corner, or any component outside its
	// bounds, reverts to the forward values
	const bool revert = inner && (!haveFl || out);
#pragma unroll
	for (int c = 0; c < NC; c++) dst[c][ix] = revert ? fv[c] : val[c];
}

// doClampComponentMAC<C> (advection.cpp:191-235); vel = face velocity * dt
template <int C, bool IS3D, typename I>
__device__ inline float clamp_comp_mac(const Geo<IS3D, I>& g, const uint8_t* __restrict__ f8, I ix, float dst,
                                       const float* __restrict__ origc, float fwd, int i, int j, int k,
                                       const float* vel, int clampMode) {
	const I Y = g.Y, Z = g.Z;
	if (clampMode == 2) {
		const I nb = ix - (C == 0 ? (I)1 : (C == 1 ? Y : Z));
		if (!(CHECKFLAG(ix) && CHECKFLAG(nb))) return fwd;
	}
	const int gux = g.nx - 1, guy = g.ny - 1, guz = g.nz - 1;
	const float pos[3] = {(float)i, (float)j, (float)k};
	float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
#pragma unroll
	for (int l = 0; l < 2; ++l) {
		if (l == 1 && clampMode != 1) break;
		int P[3];
		for (int a = 0; a < 3; a++) P[a] = (int)(l == 0 ? pos[a] - vel[a] : pos[a] + vel[a]);
		const int i0 = clampi(P[0], 0, gux - 1), j0 = clampi(P[1], 0, guy - 1);
		const float* __restrict__ o = origc + g.at(i0, j0, IS3D ? clampi(P[2], 0, guz - 1) : 0);
		const FPair e00 = ld_pair(o), e10 = ld_pair(o + Y);
		min_max(e00.a, mn, mx);
		min_max(e00.b, mn, mx);
		min_max(e10.a, mn, mx);
		min_max(e10.b, mn, mx);
		if (IS3D) {
			const FPair e01 = ld_pair(o + Z), e11 = ld_pair(o + Y + Z);
			min_max(e01.a, mn, mx);
			min_max(e01.b, mn, mx);
			min_max(e11.a, mn, mx);
			min_max(e11.b, mn, mx);
		}
	}
	if (clampMode == 1) return clampf_ref(dst, mn, mx);
	if (dst < mn || dst > mx) return fwd;
	return dst;
}

// MacCormack for MAC grids: backward SemiLagrangeMAC (fwd, -dt), correction
// (MacCormackCorrectMAC, all ijk) and clamp (MacCormackClampMAC, bnd=1) fused.
// The face velocities of the cell are read once and serve both the backward
// trace and the clamp.
// With PRE the backward values come from pbu/pbv/pbw (as k_maccormack_centered).
template <bool IS3D, typename I, bool PRE = false>
__global__ void __launch_bounds__(kBlock) k_maccormack_mac(Vel V, Dims d, Tiles t, float dt, float strength,
                                                           int clampMode, const uint8_t* __restrict__ f8,
                                                           const float* ou, const float* ov, const float* ow,
                                                           const float* __restrict__ fu, const float* __restrict__ fv,
                                                           const float* __restrict__ fw, float* __restrict__ nu,
                                                           float* __restrict__ nv, float* __restrict__ nw,
                                                           const float* __restrict__ pbu = nullptr,
                                                           const float* __restrict__ pbv = nullptr,
                                                           const float* __restrict__ pbw = nullptr) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const int i = p.i, j = p.j, k = p.k;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(i, j, k);
	const bool inner = g.interior(i, j, k);
	float vx[3], vy[3], vz[3];
	float bwd[3] = {0.f, 0.f, 0.f};
	if (inner) {
		getAtMACX(V, g, ix, vx);
		getAtMACY(V, g, ix, vy);
		getAtMACZ(V, g, ix, vz);
		if constexpr (PRE) {
			bwd[0] = pbu[ix];
			bwd[1] = pbv[ix];
			bwd[2] = pbw[ix];
		} else {
			sl_mac_cell(g, i, j, k, -dt, vx, vy, vz, fu, fv, fw, bwd);
		}
	}
	// MacCormackCorrectMAC skip logic (isMAC = true)
	bool skip[3] = {false, false, false};
	if (!(f8[ix] & kFluid)) skip[0] = skip[1] = skip[2] = true;
	if ((i > 0) && !(f8[ix - 1] & kFluid)) skip[0] = true;
	if ((j > 0) && !(f8[ix - g.Y] & kFluid)) skip[1] = true;
	if (IS3D && (k > 0) && !(f8[ix - g.Z] & kFluid)) skip[2] = true;
	const float* orig[3] = {ou, ov, ow};
	const float* fwd[3] = {fu, fv, fw};
	float val[3], fvv[3];
#pragma unroll
	for (int c = 0; c < 3; c++) {
		fvv[c] = fwd[c][ix];
		if (skip[c]) val[c] = fvv[c];
		else val[c] = (float)((double)fvv[c] + (double)strength * 0.5 * (double)(orig[c][ix] - bwd[c]));
	}
	if (inner) {
		float vdt[3];
		for (int a = 0; a < 3; a++) vdt[a] = vx[a] * dt;
		val[0] = clamp_comp_mac<0>(g, f8, ix, val[0], ou, fvv[0], i, j, k, vdt, clampMode);
		for (int a = 0; a < 3; a++) vdt[a] = vy[a] * dt;
		val[1] = clamp_comp_mac<1>(g, f8, ix, val[1], ov, fvv[1], i, j, k, vdt, clampMode);
		if (IS3D) {
			for (int a = 0; a < 3; a++) vdt[a] = vz[a] * dt;
			val[2] = clamp_comp_mac<2>(g, f8, ix, val[2], ow, fvv[2], i, j, k, vdt, clampMode);
		}
	}
	nu[ix] = val[0];
	nv[ix] = val[1];
	nw[ix] = val[2];
}

// Clamp bounds of doClampComponentMAC (advection.cpp:191-235), clampMode 2:
// min and max of the component over the 8 (4) corners at toVec3i(pos - vel),
// the same corners as clamp_comp_mac.  In clampMode 2 the bounds are only
// compared with, so fminf / fmaxf (v_min / v_max, which skip NaN operands as
// getMinMax's strict comparisons do) give getMinMax's values; only the sign
// of a zero bound could differ, which no comparison sees.
template <bool IS3D, typename I>
__device__ inline void clamp_bounds_mac(const Geo<IS3D, I>& g, const float* __restrict__ origc, int i, int j, int k,
                                        const float* vel, float& mn, float& mx) {
	const I Y = g.Y, Z = g.Z;
	const int P0 = (int)((float)i - vel[0]), P1 = (int)((float)j - vel[1]), P2 = (int)((float)k - vel[2]);
	const int i0 = min(max(P0, 0), g.nx - 2), j0 = min(max(P1, 0), g.ny - 2);
	const float* __restrict__ o = origc + g.at(i0, j0, IS3D ? min(max(P2, 0), g.nz - 2) : 0);
	const FPair e00 = ld_pair(o), e10 = ld_pair(o + Y);
	// seeded with +-FLT_MAX as getMinMax is (so infinite and NaN corners give
	// its values too)
	mn = fminf(fminf(3.402823466e+38f, fminf(e00.a, e00.b)), fminf(e10.a, e10.b));
	mx = fmaxf(fmaxf(-3.402823466e+38f, fmaxf(e00.a, e00.b)), fmaxf(e10.a, e10.b));
	if (IS3D) {
		const FPair e01 = ld_pair(o + Z), e11 = ld_pair(o + Y + Z);
		mn = fminf(mn, fminf(fminf(e01.a, e01.b), fminf(e11.a, e11.b)));
		mx = fmaxf(mx, fmaxf(fmaxf(e01.a, e01.b), fmaxf(e11.a, e11.b)));
	}
}

// k_maccormack_mac for clampMode 2 on grids of at least 3 cells per axis, the
// same arithmetic in an order with two dependent memory rounds instead of one
// per conditional: every lane evaluates the trace at its cell clamped into
// the interior (equal to its own cell where the reference evaluates it; the
// other lanes' results are discarded), so there is no divergent branch; the
// cell's flags, forward and original values load with the face velocities,
// and the clamp bounds (which depend on the face velocities and the original
// field only) load with the backward trace.
template <bool IS3D, typename I>
__global__ void __launch_bounds__(kBlock) k_maccormack_mac2(Vel V, Dims d, Tiles t, float dt, float strength,
                                                            const uint8_t* __restrict__ f8, const float* ou,
                                                            const float* ov, const float* ow,
                                                            const float* __restrict__ fu, const float* __restrict__ fv,
                                                            const float* __restrict__ fw, float* __restrict__ nu,
                                                            float* __restrict__ nv, float* __restrict__ nw) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const int i = p.i, j = p.j, k = p.k;
	const Geo<IS3D, I> g(d);
	const I ix = g.at(i, j, k);
	const bool inner = g.interior(i, j, k);
	const int ic = clampi(i, 1, g.nx - 2), jc = clampi(j, 1, g.ny - 2), kc = IS3D ? clampi(k, 1, g.nz - 2) : k;
	const I ixc = g.at(ic, jc, kc);
	// the cell's flags (lower neighbours clamped into the grid; used only
	// where they exist), forward and original values
	constexpr int FM = kFluid | kEmpty;
	const uint8_t f0 = f8[ix];
	const uint8_t fxm = f8[i > 0 ? ix - 1 : ix];
	const uint8_t fym = f8[j > 0 ? ix - g.Y : ix];
	const uint8_t fzm = IS3D ? f8[k > 0 ? ix - g.Z : ix] : (uint8_t)0;
	const float* orig[3] = {ou, ov, ow};
	const float* fwd[3] = {fu, fv, fw};
	float fvv[3], org[3];
#pragma unroll
	for (int c = 0; c < 3; c++) {
		fvv[c] = fwd[c][ix];
		org[c] = orig[c][ix];
	}
	float vx[3], vy[3], vz[3];
	getAtMACX(V, g, ixc, vx);
	getAtMACY(V, g, ixc, vy);
	getAtMACZ(V, g, ixc, vz);
	float mn[3], mx[3];
	{
		float vdt[3];
		for (int a = 0; a < 3; a++) vdt[a] = vx[a] * dt;
		clamp_bounds_mac(g, ou, ic, jc, kc, vdt, mn[0], mx[0]);
		for (int a = 0; a < 3; a++) vdt[a] = vy[a] * dt;
		clamp_bounds_mac(g, ov, ic, jc, kc, vdt, mn[1], mx[1]);
		if (IS3D) {
			for (int a = 0; a < 3; a++) vdt[a] = vz[a] * dt;
			clamp_bounds_mac(g, ow, ic, jc, kc, vdt, mn[2], mx[2]);
		}
	}
	float bwd[3];
	sl_mac_cell(g, ic, jc, kc, -dt, vx, vy, vz, fu, fv, fw, bwd);
	if (!inner) bwd[0] = bwd[1] = bwd[2] = 0.f;
	// MacCormackCorrectMAC skip logic (isMAC = true)
	const bool sk = !(f0 & kFluid);
	const bool skip[3] = {sk || ((i > 0) && !(fxm & kFluid)), sk || ((j > 0) && !(fym & kFluid)),
	                      sk || (IS3D && (k > 0) && !(fzm & kFluid))};
	// clampMode 2: no clamp unless the cell and its lower face neighbour are
	// fluid or empty
	const bool cf[3] = {(f0 & FM) && (fxm & FM), (f0 & FM) && (fym & FM), (f0 & FM) && (fzm & FM)};
	float val[3];
#pragma unroll
	for (int c = 0; c < 3; c++) {
		// evaluated for every lane and selected; the empty asm keeps the
		// compiler from turning the select back into a branch around the
		// load of org
		float corr = (float)((double)fvv[c] + (double)strength * 0.5 * (double)(org[c] - bwd[c]));
		asm("" : "+v"(corr));
		float x = skip[c] ? fvv[c] : corr;
		if (inner && (c < 2 || IS3D)) {
			if (!cf[c]) x = fvv[c];
			else if (x < mn[c] || x > mx[c]) x = fvv[c];
		}
		val[c] = x;
	}
	nu[ix] = val[0];
	nv[ix] = val[1];
	nw[ix] = val[2];
}

// ---------------------------------------------------------------------------
// outflow extrapolation (advection.cpp:326-391): only outflow cells do work;
// velDst is kept in a scratch Vec grid and copied back in a second pass.
// ---------------------------------------------------------------------------
__device__ inline bool inb(const Dims& d, int i, int j, int k) {
	return i >= 0 && j >= 0 && k >= 0 && i < d.nx && j < d.ny && k < d.nz;
}

__device__ inline void outflow_extrap_cell(const Dims& d, const uint8_t* __restrict__ f8, const float* __restrict__ vu,
                                           const float* __restrict__ vv, const float* __restrict__ vw,
                                           const float* __restrict__ pu, const float* __restrict__ pv,
                                           const float* __restrict__ pw, float timeStep, float* __restrict__ du,
                                           float* __restrict__ dv, float* __restrict__ dw, int i, int j, int k,
                                           long long ix) {
	const float* vel[3] = {vu, vv, vw};
	const float* prev[3] = {pu, pv, pw};
	// getBulkVel
	float avg[3] = {0.f, 0.f, 0.f};
	int count = 0;
	const int nmax = d.is3d ? 1 : 0;
	for (int nn = -nmax; nn <= nmax; nn++)
		for (int m = -1; m <= 1; m++)
			for (int l = -1; l <= 1; l++) {
				if (inb(d, i + l, j + m, k + nn)) {
					const long long q = d.idx(i + l, j + m, k + nn);
					if ((f8[q] & kFluid) || (f8[q] & kOutflow)) {
						for (int c = 0; c < 3; c++) avg[c] += vel[c][q];
						count++;
					}
				}
			}
	float bulk[3];
	for (int c = 0; c < 3; c++) bulk[c] = count > 0 ? avg[c] / (float)count : avg[c];
	const int dim = d.is3d ? 3 : 2;
	const int cur[3] = {i, j, k};
	int cnt = 0;
	float vd[3] = {0.f, 0.f, 0.f};
	for (int c = 0; c < dim; c++) {
		int low[3], up[3], flLow[3], flUp[3];
		for (int a = 0; a < 3; a++) low[a] = up[a] = flLow[a] = flUp[a] = cur[a];
		const float mb = (1.0f < bulk[c]) ? bulk[c] : 1.0f;
		const float factor = timeStep * mb;
		low[c] = flLow[c] = cur[c] - 1;
		up[c] = flUp[c] = cur[c] + 1;
		for (int dd = 0; dd < 2; dd++) {
			const bool fromLower = inb(d, flLow[0], flLow[1], flLow[2]) && (f8[d.idx(flLow[0], flLow[1], flLow[2])] & kFluid);
			const bool fromUpper = inb(d, flUp[0], flUp[1], flUp[2]) && (f8[d.idx(flUp[0], flUp[1], flUp[2])] & kFluid);
			if (fromLower || fromUpper) {
				if (fromLower) {
					const long long q = d.idx(low[0], low[1], low[2]);
					for (int a = 0; a < 3; a++) vd[a] += ((vel[a][ix] - prev[a][ix]) / factor) + vel[a][q];
					cnt++;
				}
				if (fromUpper) {
					const long long q = d.idx(up[0], up[1], up[2]);
					for (int a = 0; a < 3; a++) vd[a] += ((vel[a][ix] - prev[a][ix]) / factor) + vel[a][q];
					cnt++;
				}
				break;
			}
			flLow[c]--;
			flUp[c]++;
		}
	}
	if (cnt > 0)
		for (int a = 0; a < 3; a++) vd[a] /= (float)cnt;
	du[ix] = vd[0];
	dv[ix] = vd[1];
	dw[ix] = vd[2];
}

// applyOutflowBC's extrapolation at the outflow cells listed by outflow_scan
// (one cell per thread, grid-stride over the device count)
__global__ void __launch_bounds__(256) k_outflow_extrap(Dims d, const unsigned* __restrict__ list,
                                                        const unsigned* __restrict__ count,
                                                        const uint8_t* __restrict__ f8, const float* __restrict__ vu,
                                                        const float* __restrict__ vv, const float* __restrict__ vw,
                                                        const float* __restrict__ pu, const float* __restrict__ pv,
                                                        const float* __restrict__ pw, float timeStep,
                                                        float* __restrict__ du, float* __restrict__ dv,
                                                        float* __restrict__ dw) {
	const unsigned m = *count;
	const unsigned nxy = (unsigned)(d.nx * d.ny);
	for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x) {
		const unsigned ix = list[q];
		const unsigned k = ix / nxy, rest = ix - k * nxy;
		const unsigned j = rest / (unsigned)d.nx, i = rest - j * (unsigned)d.nx;
		outflow_extrap_cell(d, f8, vu, vv, vw, pu, pv, pw, timeStep, du, dv, dw, (int)i, (int)j, (int)k,
		                    (long long)ix);
	}
}
// and the copy back of the extrapolated values
__global__ void __launch_bounds__(256) k_outflow_copy(const unsigned* __restrict__ list,
                                                      const unsigned* __restrict__ count, const float* __restrict__ du,
                                                      const float* __restrict__ dv, const float* __restrict__ dw,
                                                      float* __restrict__ vu, float* __restrict__ vv,
                                                      float* __restrict__ vw) {
	const unsigned m = *count;
	for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x) {
		const unsigned ix = list[q];
		vu[ix] = du[ix];
		vv[ix] = dv[ix];
		vw[ix] = dw[ix];
	}
}


// ---------------------------------------------------------------------------
// KnSetWallBcs (extforces.cpp:186-217), all ijk.  kLightKZ planes per thread:
// every flag (and obstacle velocity) of all planes is loaded first, from
// offsets that are 0 where the reference would not read the neighbour, then
// each face's final value is written once (the reference's later writes
// overwrite its earlier ones; no face is read in between).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_set_wall_bcs(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                         float* __restrict__ u, float* __restrict__ v,
                                                         float* __restrict__ w, const float* __restrict__ ou,
                                                         const float* __restrict__ ov, const float* __restrict__ ow) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const int i = p.i, j = p.j;
	const long long oxm = i > 0 ? 1 : 0, oxp = i < d.nx - 1 ? 1 : 0;
	const long long oym = j > 0 ? d.nx : 0, oyp = j < d.ny - 1 ? d.nx : 0;
	constexpr int KZ = kLightKZ;
	uint8_t f[KZ], fxm[KZ], fxp[KZ], fym[KZ], fyp[KZ], fzm[KZ], fzp[KZ];
	float ob[KZ][3];
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = min(p.k + q, d.nz - 1);  // planes past the end are loaded, never used
		const long long ix = d.idx(i, j, k);
		const long long ozm = (d.is3d && k > 0) ? d.sz : 0, ozp = (d.is3d && k < d.nz - 1) ? d.sz : 0;
		f[q] = f8[ix];
		fxm[q] = f8[ix - oxm];
		fxp[q] = f8[ix + oxp];
		fym[q] = f8[ix - oym];
		fyp[q] = f8[ix + oyp];
		fzm[q] = f8[ix - ozm];
		fzp[q] = f8[ix + ozp];
		ob[q][0] = ob[q][1] = ob[q][2] = 0.f;
		if (ou) {
			ob[q][0] = ou[ix];
			ob[q][1] = ov[ix];
			if (d.is3d) ob[q][2] = ow[ix];
		}
	}
	// The per-cell rules (wall_rule) as boolean algebra and selects, no
	// branch per condition.
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = p.k + q;
		if (k >= d.nz) break;
		const long long ix = d.idx(i, j, k);
		const WallRule r = wall_rule(d, i, j, k, f[q], fxm[q], fxp[q], fym[q], fyp[q], fzm[q], fzp[q]);
		const int is3 = d.is3d ? 1 : 0;
		// value masks: all ones keeps the obstacle velocity, zero gives 0.0
		const float nu = __int_as_float(__float_as_int(ob[q][0]) & ((r.sy | r.sz) - 1));
		const float nv = __int_as_float(__float_as_int(ob[q][1]) & ((r.sx | r.sz) - 1));
		const float nw = __int_as_float(__float_as_int(ob[q][2]) & ((r.sx | r.sy | (is3 ^ 1)) - 1));
		if (r.wu) u[ix] = nu;
		if (r.wv) v[ix] = nv;
		if (r.ww) w[ix] = nw;
	}
}

// setWallBcs (no obstacle velocity) followed by addBuoyancy in one pass:
// each face is first zeroed where KnSetWallBcs writes it, then gets
// KnAddBuoyancy's increment (extforces.cpp:72-80, bnd=1, evaluated in
// double), and is written once if that changed its bits.  Same values as
// k_set_wall_bcs then k_add_buoyancy; kLightKZ planes per thread, every
// load issued first.
__global__ void __launch_bounds__(kBlock) k_wall_buoyancy(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                          const float* __restrict__ den, float* __restrict__ u,
                                                          float* __restrict__ v, float* __restrict__ w, float fx,
                                                          float fy, float fz) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const int i = p.i, j = p.j;
	const long long oxm = i > 0 ? 1 : 0, oxp = i < d.nx - 1 ? 1 : 0;
	const long long oym = j > 0 ? d.nx : 0, oyp = j < d.ny - 1 ? d.nx : 0;
	constexpr int KZ = kLightKZ;
	uint8_t f[KZ], fxm[KZ], fxp[KZ], fym[KZ], fyp[KZ], fzm[KZ], fzp[KZ];
	float dc[KZ], dxm[KZ], dym[KZ], dzm[KZ], uc[KZ], vc[KZ], wc[KZ];
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = min(p.k + q, d.nz - 1);  // planes past the end are loaded, never used
		const long long ix = d.idx(i, j, k);
		const long long ozm = (d.is3d && k > 0) ? d.sz : 0, ozp = (d.is3d && k < d.nz - 1) ? d.sz : 0;
		f[q] = f8[ix];
		fxm[q] = f8[ix - oxm];
		fxp[q] = f8[ix + oxp];
		fym[q] = f8[ix - oym];
		fyp[q] = f8[ix + oyp];
		fzm[q] = f8[ix - ozm];
		fzp[q] = f8[ix + ozp];
		dc[q] = den[ix];
		dxm[q] = den[ix - oxm];
		dym[q] = den[ix - oym];
		dzm[q] = den[ix - ozm];
		uc[q] = u[ix];
		vc[q] = v[ix];
		wc[q] = w[ix];
	}
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = p.k + q;
		if (k >= d.nz) break;
		const long long ix = d.idx(i, j, k);
		const WallRule r = wall_rule(d, i, j, k, f[q], fxm[q], fxp[q], fym[q], fyp[q], fzm[q], fzp[q]);
		float nu = r.wu ? 0.f : uc[q], nv = r.wv ? 0.f : vc[q], nw = r.ww ? 0.f : wc[q];
		const bool bi = interior(d, i, j, k) && (f[q] & kFluid);
		const bool bu = bi && (fxm[q] & kFluid), bv = bi && (fym[q] & kFluid), bw = bi && d.is3d && (fzm[q] & kFluid);
		if (bu) nu = (float)((double)nu + (0.5 * (double)fx) * (double)(dc[q] + dxm[q]));
		if (bv) nv = (float)((double)nv + (0.5 * (double)fy) * (double)(dc[q] + dym[q]));
		if (bw) nw = (float)((double)nw + (0.5 * (double)fz) * (double)(dc[q] + dzm[q]));
		// a face is stored only where its bits change (an increment of 0 from
		// a zero gravity component changes only a -0); the others already
		// hold the result
		if (__float_as_uint(nu) != __float_as_uint(uc[q])) u[ix] = nu;
		if (__float_as_uint(nv) != __float_as_uint(vc[q])) v[ix] = nv;
		if (__float_as_uint(nw) != __float_as_uint(wc[q])) w[ix] = nw;
	}
}

// KnAddBuoyancy (extforces.cpp:72-80), bnd=1; face update evaluated in
// double.  kLightKZ planes per thread, all loads issued first.
__global__ void __launch_bounds__(kBlock) k_add_buoyancy(Dims d, Tiles t, const uint8_t* __restrict__ f8,
                                                         const float* __restrict__ den, float* __restrict__ u,
                                                         float* __restrict__ v, float* __restrict__ w, float fx,
                                                         float fy, float fz) {
	const CellPos p = tile_cell(t);
	if (p.i >= d.nx || p.j >= d.ny) return;
	const int i = p.i, j = p.j;
	constexpr int KZ = kLightKZ;
	// neighbour offsets, valid for every cell (interior cells use the true ones)
	const long long ox = i > 0 ? 1 : 0, oy = j > 0 ? d.nx : 0;
	uint8_t f[KZ], gx[KZ], gy[KZ], gz[KZ];
	float dc[KZ], dxm[KZ], dym[KZ], dzm[KZ], uc[KZ], vc[KZ], wc[KZ];
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = min(p.k + q, d.nz - 1);
		const long long ix = d.idx(i, j, k);
		f[q] = f8[ix];
		gx[q] = f8[ix - ox];
		gy[q] = f8[ix - oy];
		dc[q] = den[ix];
		dxm[q] = den[ix - ox];
		dym[q] = den[ix - oy];
		uc[q] = u[ix];
		vc[q] = v[ix];
		wc[q] = w[ix];
	}
	// z-neighbours: the plane below the first is loaded; for the others it is
	// the thread's previous plane (planes past nz are clamped and never used)
	{
		const int k0 = min(p.k, d.nz - 1);
		const long long c0 = d.idx(i, j, k0) - (k0 > 0 ? d.sz : 0);
		gz[0] = f8[c0];
		dzm[0] = den[c0];
	}
#pragma unroll
	for (int q = 1; q < KZ; q++) {
		gz[q] = f[q - 1];
		dzm[q] = dc[q - 1];
	}
#pragma unroll
	for (int q = 0; q < KZ; q++) {
		const int k = p.k + q;
		if (k >= d.nz) break;
		if (!interior(d, i, j, k) || !(f[q] & kFluid)) continue;
		const long long ix = d.idx(i, j, k);
		if (gx[q] & kFluid) u[ix] = (float)((double)uc[q] + (0.5 * (double)fx) * (double)(dc[q] + dxm[q]));
		if (gy[q] & kFluid) v[ix] = (float)((double)vc[q] + (0.5 * (double)fy) * (double)(dc[q] + dym[q]));
		if (d.is3d && (gz[q] & kFluid))
			w[ix] = (float)((double)wc[q] + (0.5 * (double)fz) * (double)(dc[q] + dzm[q]));
	}
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline Vel vel_of(Grid* g) { return Vel{g->comp(0), g->comp(1), g->comp(2)}; }

// ---------------------------------------------------------------------------
// setWallBcs with fractions and phiObs (extforces.cpp:220-315):
// KnSetWallBcsFrac into a second buffer, then swapped in.  At interior fluid
// or obstacle cells, a face touching an obstacle (the cell or its lower
// neighbour) loses the component of its getAtMAC velocity along the
// normalised obstacle-SDF gradient; every other value is copied.  Float and
// double roundings follow the reference (normalize: util/vectorbase.h:416-429).
// ---------------------------------------------------------------------------
__device__ inline void frac_normalize(float* v) {
	const float l = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
	const float eps2 = 1e-6f * 1e-6f;
	if (fabs((double)l - 1.) < (double)eps2) return;
	if (l > eps2) {
		const float norm = (float)sqrt((double)l);
		const float sc = (float)(1. / (double)norm);
		v[0] *= sc;
		v[1] *= sc;
		v[2] *= sc;
	} else {
		v[0] = v[1] = v[2] = 0.f;
	}
}

__global__ void __launch_bounds__(256) k_set_wall_bcs_frac(Dims d, const uint8_t* __restrict__ f8,
                                                           const float* __restrict__ u, const float* __restrict__ v,
                                                           const float* __restrict__ w, const float* __restrict__ P,
                                                           float* __restrict__ ou, float* __restrict__ ov,
                                                           float* __restrict__ ow) {
	const long long X = 1, Y = d.nx, Z = d.sz;
	for (long long c = blockIdx.x * 256ll + threadIdx.x; c < d.n; c += (long long)gridDim.x * 256) {
		const long long row = c / d.nx;
		const int i = (int)(c - row * d.nx);
		const int j = d.is3d ? (int)(row % d.ny) : (int)row;
		const int k = d.is3d ? (int)(row / d.ny) : 0;
		float nu = u[c], nv = v[c], nw = w[c];
		const uint8_t f = f8[c];
		const bool curObs = (f & kObstacle) != 0;
		const bool inb = i >= 1 && j >= 1 && i < d.nx - 1 && j < d.ny - 1 && (!d.is3d || (k >= 1 && k < d.nz - 1));
		if (((f & kFluid) || curObs) && inb) {
			if (curObs || (f8[c - X] & kObstacle)) {
				float dphi[3] = {0.f, 0.f, 0.f};
				const float tmp1 = (float)((P[c] + P[c - X]) * .5);
				float tmp2 = (float)((P[c + Y] + P[c - X + Y]) * .5);
				float phi1 = (float)((tmp1 + tmp2) * .5);
				tmp2 = (float)((P[c - Y] + P[c - X - Y]) * .5);
				float phi2 = (float)((tmp1 + tmp2) * .5);
				dphi[0] = P[c] - P[c - X];
				dphi[1] = phi1 - phi2;
				if (d.is3d) {
					tmp2 = (float)((P[c + Z] + P[c - X + Z]) * .5);
					phi1 = (float)((tmp1 + tmp2) * .5);
					tmp2 = (float)((P[c - Z] + P[c - X - Z]) * .5);
					phi2 = (float)((tmp1 + tmp2) * .5);
					dphi[2] = phi1 - phi2;
				}
				frac_normalize(dphi);
				// getAtMACX (grid.h:437-448)
				const float m0 = u[c];
				const float m1 = 0.25f * (v[c] + v[c - 1] + v[c + Y] + v[c + Y - 1]);
				const float m2 = d.is3d ? 0.25f * (w[c] + w[c - 1] + w[c + Z] + w[c + Z - 1]) : 0.f;
				const float dt = dphi[0] * m0 + dphi[1] * m1 + dphi[2] * m2;
				nu = m0 - dt * dphi[0];
			}
			if (curObs || (f8[c - Y] & kObstacle)) {
				float dphi[3] = {0.f, 0.f, 0.f};
				const float tmp1 = (float)((P[c] + P[c - Y]) * .5);
				float tmp2 = (float)((P[c + X] + P[c + X - Y]) * .5);
				float phi1 = (float)((tmp1 + tmp2) * .5);
				tmp2 = (float)((P[c - X] + P[c - X - Y]) * .5);
				float phi2 = (float)((tmp1 + tmp2) * .5);
				dphi[0] = phi1 - phi2;
				dphi[1] = P[c] - P[c - Y];
				if (d.is3d) {
					tmp2 = (float)((P[c + Z] + P[c - Y + Z]) * .5);
					phi1 = (float)((tmp1 + tmp2) * .5);
					tmp2 = (float)((P[c - Z] + P[c - Y - Z]) * .5);
					phi2 = (float)((tmp1 + tmp2) * .5);
					dphi[2] = phi1 - phi2;
				}
				frac_normalize(dphi);
				// getAtMACY (grid.h:450-459)
				const float m0 = 0.25f * (u[c] + u[c - Y] + u[c + 1] + u[c + 1 - Y]);
				const float m1 = v[c];
				const float m2 = d.is3d ? 0.25f * (w[c] + w[c - Y] + w[c + Z] + w[c + Z - Y]) : 0.f;
				const float dt = dphi[0] * m0 + dphi[1] * m1 + dphi[2] * m2;
				nv = m1 - dt * dphi[1];
			}
			if (d.is3d && (curObs || (f8[c - Z] & kObstacle))) {
				float dphi[3] = {0.f, 0.f, 0.f};
				const float tmp1 = (float)((P[c] + P[c - Z]) * .5);
				float tmp2 = (float)((P[c + X] + P[c + X - Z]) * .5);
				float phi1 = (float)((tmp1 + tmp2) * .5);
				tmp2 = (float)((P[c - X] + P[c - X - Z]) * .5);
				float phi2 = (float)((tmp1 + tmp2) * .5);
				dphi[0] = phi1 - phi2;
				tmp2 = (float)((P[c + Y] + P[c + Y - Z]) * .5);
				phi1 = (float)((tmp1 + tmp2) * .5);
				tmp2 = (float)((P[c - Y] + P[c - Y - Z]) * .5);
				phi2 = (float)((tmp1 + tmp2) * .5);
				dphi[1] = phi1 - phi2;
				dphi[2] = P[c] - P[c - Z];
				frac_normalize(dphi);
				// getAtMACZ (grid.h:461-470)
				const float m0 = 0.25f * (u[c] + u[c - Z] + u[c + 1] + u[c + 1 - Z]);
				const float m1 = 0.25f * (v[c] + v[c - Z] + v[c + Y] + v[c + Y - Z]);
				const float m2 = w[c];
				const float dt = dphi[0] * m0 + dphi[1] * m1 + dphi[2] * m2;
				nw = m2 - dt * dphi[2];
			}
		}
		ou[c] = nu;
		ov[c] = nv;
		ow[c] = nw;
	}
}

void op_set_wall_bcs_frac(Scene* s, Grid* flags, Grid* vel, Grid* phiObs) {
	float* out = s->getVec();
	hipLaunchKernelGGL(k_set_wall_bcs_frac, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, s->d,
	                   (const uint8_t*)flags->f8, (const float*)vel->comp(0), (const float*)vel->comp(1),
	                   (const float*)vel->comp(2), (const float*)phiObs->real(), out, out + s->cstride,
	                   out + 2 * s->cstride);
	// vel.swap(tmpvel)
	s->putVec(vel->real());
	vel->data = out;
}

void op_set_wall_bcs(Scene* s, Grid* flags, Grid* vel, Grid* obvel) {
	Tiles t = tiles_for(s->d, kLightKZ);
	hipLaunchKernelGGL(k_set_wall_bcs, dim3((unsigned)t.count()), dim3(kBlock), 0, s->stream, s->d, t, flags->f8,
	                   vel->comp(0), vel->comp(1), vel->comp(2), obvel ? obvel->comp(0) : nullptr,
	                   obvel ? obvel->comp(1) : nullptr, obvel ? obvel->comp(2) : nullptr);
}

void op_wall_buoyancy(Scene* s, Grid* flags, Grid* density, Grid* vel, const float f[3]) {
	Tiles t = tiles_for(s->d, kLightKZ);
	hipLaunchKernelGGL(k_wall_buoyancy, dim3((unsigned)t.count()), dim3(kBlock), 0, s->stream, s->d, t, flags->f8,
	                   (const float*)density->real(), vel->comp(0), vel->comp(1), vel->comp(2), f[0], f[1], f[2]);
}

// addBuoyancy's force (extforces.cpp:83-87): float gridScale = scale ?
// getDx() : 1; f = -gravity * dt / gridScale * coefficient (Vec3 float ops,
// one rounding each)
void buoyancy_strength(const Scene* s, const float gravity[3], float coefficient, int scale, float fv[3]) {
	int mx = s->d.nx > s->d.ny ? s->d.nx : s->d.ny;
	if (s->d.nz > mx) mx = s->d.nz;
	const float dx = (float)(1.0 / mx);
	volatile float gridScale = scale ? dx : 1.f;
	for (int c = 0; c < 3; c++) {
		volatile float a = -gravity[c];
		volatile float b = a * s->dt;
		volatile float q = b / gridScale;
		fv[c] = q * coefficient;
	}
}

void op_add_buoyancy(Scene* s, Grid* flags, Grid* density, Grid* vel, const float f[3]) {
	Tiles t = tiles_for(s->d, kLightKZ);
	hipLaunchKernelGGL(k_add_buoyancy, dim3((unsigned)t.count()), dim3(kBlock), 0, s->stream, s->d, t, flags->f8,
	                   (const float*)density->real(), vel->comp(0), vel->comp(1), vel->comp(2), f[0], f[1], f[2]);
}

// Testing hook: route every grid through the 64-bit index instantiations
// (otherwise only grids of 2^31 cells and more reach them).
static int g_force_index64 = 0;

// Calls f(std::integral_constant<bool, IS3D>, I{}) with the grid's dimension
// and the narrowest index type that addresses it.
template <typename F>
static void with_geo(const Dims& d, F&& f) {
	const bool i32 = index32_ok(d) && !g_force_index64;
	if (d.is3d) {
		if (i32) f(std::true_type{}, int{});
		else f(std::true_type{}, (long long){});
	} else {
		if (i32) f(std::false_type{}, int{});
		else f(std::false_type{}, (long long){});
	}
}

// applyOutflowBC(flags, result, orig, dt) (advection.cpp:387-391):
// timeStep = max(1.0, dt*4); the extrapolated values go to a scratch grid and
// are copied back into result
static void outflow_bc(Scene* s, Grid* flags, Grid* grid, float* result) {
	const Dims d = s->d;
	const double tsd = (double)s->dt * 4;
	const float ts = (float)((1.0 < tsd) ? tsd : 1.0);
	float* vdst = s->getVec();
	float* r3[3] = {result, result + s->cstride, result + 2 * s->cstride};
	float* d3[3] = {vdst, vdst + s->cstride, vdst + 2 * s->cstride};
	const unsigned* count = nullptr;
	const unsigned* list = outflow_list(flags, &count);
	hipLaunchKernelGGL(k_outflow_extrap, dim3(kOutflowListBlocks), dim3(256), 0, s->stream, d, list, count, flags->f8, (const float*)r3[0], (const float*)r3[1], (const float*)r3[2],
	                   (const float*)grid->comp(0), (const float*)grid->comp(1), (const float*)grid->comp(2), ts, d3[0],
	                   d3[1], d3[2]);
	hipLaunchKernelGGL(k_outflow_copy, dim3(kOutflowListBlocks), dim3(256), 0, s->stream, list, count,
	                   (const float*)d3[0], (const float*)d3[1], (const float*)d3[2], r3[0], r3[1], r3[2]);
	s->putVec(vdst);
}

// fnAdvectSemiLagrange with orderSpace or orderTrace 2 (advection.cpp:292-321,
// 406-436): forward SL, backward SL of the forward result and the fused
// correction + clamp as three passes
static void op_advect_hi(Scene* s, Grid* flags, Grid* vel, Grid* grid, int order, float strength, int clampMode,
                         int orderSpace, int orderTrace) {
	const Dims d = s->d;
	const Tiles t = tiles_for(d);
	const dim3 gb((unsigned)t.count()), tb(kBlock);
	const float dt = s->dt;
	const Vel V = vel_of(vel);
	const bool mac = grid->isMAC();
	const int nc = grid->isVec() ? 3 : 1;
	auto take = [&]() { return nc == 3 ? s->getVec() : s->getReal(); };
	auto give = [&](float* q) {
		if (nc == 3) s->putVec(q);
		else s->putReal(q);
	};
	auto comps = [&](float* q, float** c3) {
		c3[0] = q;
		c3[1] = nc == 3 ? q + s->cstride : nullptr;
		c3[2] = nc == 3 ? q + 2 * s->cstride : nullptr;
	};
	// one SL step of src into dst with time step h
	auto sl = [&](float* const* src, float* const* dst, float h) {
		with_geo(d, [&](auto D3, auto ix) {
			constexpr bool D = decltype(D3)::value;
			using I = decltype(ix);
			if (mac)
				hipLaunchKernelGGL((k_sl_mac_hi<D, I>), gb, tb, 0, s->stream, V, d, t, h, orderSpace, orderTrace,
				                   Vel{src[0], src[1], src[2]}, dst[0], dst[1], dst[2]);
			else if (nc == 1)
				hipLaunchKernelGGL((k_sl_centered_hi<1, D, I>), gb, tb, 0, s->stream, V, d, t, h, orderSpace,
				                   orderTrace, (const float*)src[0], (const float*)src[1], (const float*)src[2],
				                   dst[0], dst[1], dst[2]);
			else
				hipLaunchKernelGGL((k_sl_centered_hi<3, D, I>), gb, tb, 0, s->stream, V, d, t, h, orderSpace,
				                   orderTrace, (const float*)src[0], (const float*)src[1], (const float*)src[2],
				                   dst[0], dst[1], dst[2]);
		});
	};
	float* orig[3];
	comps(grid->real(), orig);
	float* fwd = take();
	float* fw[3];
	comps(fwd, fw);
	sl(orig, fw, dt);
	float* result = fwd;
	if (order == 2) {
		float* bwd = take();
		float* ng = take();
		float *bw[3], *nw[3];
		comps(bwd, bw);
		comps(ng, nw);
		sl(fw, bw, -dt);
		with_geo(d, [&](auto D3, auto ix) {
			constexpr bool D = decltype(D3)::value;
			using I = decltype(ix);
			if (mac)
				hipLaunchKernelGGL((k_maccormack_mac<D, I, true>), gb, tb, 0, s->stream, V, d, t, dt, strength,
				                   clampMode, flags->f8, (const float*)orig[0], (const float*)orig[1],
				                   (const float*)orig[2], (const float*)fw[0], (const float*)fw[1],
				                   (const float*)fw[2], nw[0], nw[1], nw[2], (const float*)bw[0], (const float*)bw[1],
				                   (const float*)bw[2]);
			else if (nc == 1)
				hipLaunchKernelGGL((k_maccormack_centered<1, D, I, true>), gb, tb, 0, s->stream, V, d, t, dt,
				                   strength, clampMode, flags->f8, (const float*)orig[0], (const float*)orig[1],
				                   (const float*)orig[2], (const float*)fw[0], (const float*)fw[1],
				                   (const float*)fw[2], nw[0], nw[1], nw[2], (const float*)bw[0], (const float*)bw[1],
				                   (const float*)bw[2]);
			else
				hipLaunchKernelGGL((k_maccormack_centered<3, D, I, true>), gb, tb, 0, s->stream, V, d, t, dt,
				                   strength, clampMode, flags->f8, (const float*)orig[0], (const float*)orig[1],
				                   (const float*)orig[2], (const float*)fw[0], (const float*)fw[1],
				                   (const float*)fw[2], nw[0], nw[1], nw[2], (const float*)bw[0], (const float*)bw[1],
				                   (const float*)bw[2]);
		});
		give(bwd);
		give(fwd);
		result = ng;
	}
	if (mac) outflow_bc(s, flags, grid, result);
	give(grid->real());
	grid->data = result;
}

void op_advect(Scene* s, Grid* flags, Grid* vel, Grid* grid, int order, float strength, int clampMode, int orderSpace,
               int orderTrace) {
	if (orderSpace != 1 || orderTrace != 1) {
		op_advect_hi(s, flags, vel, grid, order, strength, clampMode, orderSpace, orderTrace);
		return;
	}
	const Dims d = s->d;
	const Tiles t = tiles_for(d);
	const dim3 gb((unsigned)t.count()), tb(kBlock);
	const float dt = s->dt;
	const Vel V = vel_of(vel);
	if (grid->isMAC()) {
		float* fwd = s->getVec();
		float* fw[3] = {fwd, fwd + s->cstride, fwd + 2 * s->cstride};
		{
			KtScope kt("advect_sl_mac", s->stream);
			with_geo(d, [&](auto D3, auto ix) {
				hipLaunchKernelGGL((k_sl_mac<decltype(D3)::value, decltype(ix)>), gb, tb, 0, s->stream, V, d, t, dt,
				                   (const float*)grid->comp(0), (const float*)grid->comp(1),
				                   (const float*)grid->comp(2), fw[0], fw[1], fw[2]);
			});
		}
		float* result = fwd;
		if (order == 2) {
			float* ng = s->getVec();
			float* nw[3] = {ng, ng + s->cstride, ng + 2 * s->cstride};
			KtScope kt("advect_maccormack_mac", s->stream);
			const bool fast = clampMode == 2 && d.nx >= 3 && d.ny >= 3 && (!d.is3d || d.nz >= 3);
			with_geo(d, [&](auto D3, auto ix) {
				if (fast)
					hipLaunchKernelGGL((k_maccormack_mac2<decltype(D3)::value, decltype(ix)>), gb, tb, 0, s->stream, V,
					                   d, t, dt, strength, flags->f8, (const float*)grid->comp(0),
					                   (const float*)grid->comp(1), (const float*)grid->comp(2), (const float*)fw[0],
					                   (const float*)fw[1], (const float*)fw[2], nw[0], nw[1], nw[2]);
				else
					hipLaunchKernelGGL((k_maccormack_mac<decltype(D3)::value, decltype(ix)>), gb, tb, 0, s->stream, V,
					                   d, t, dt, strength, clampMode, flags->f8, (const float*)grid->comp(0),
					                   (const float*)grid->comp(1), (const float*)grid->comp(2), (const float*)fw[0],
					                   (const float*)fw[1], (const float*)fw[2], nw[0], nw[1], nw[2]);
			});
			s->putVec(fwd);
			result = ng;
		}
		outflow_bc(s, flags, grid, result);
		// orig.swap(result)
		s->putVec(grid->real());
		grid->data = result;
		return;
	}
	const int nc = grid->isVec() ? 3 : 1;
	float* fwd = nc == 3 ? s->getVec() : s->getReal();
	float* fw[3] = {fwd, nc == 3 ? fwd + s->cstride : nullptr, nc == 3 ? fwd + 2 * s->cstride : nullptr};
	const float* src[3] = {grid->comp(0), nc == 3 ? grid->comp(1) : nullptr, nc == 3 ? grid->comp(2) : nullptr};
	with_geo(d, [&](auto D3, auto ix) {
		constexpr bool D = decltype(D3)::value;
		using I = decltype(ix);
		if (nc == 1)
			hipLaunchKernelGGL((k_sl_centered<1, D, I>), gb, tb, 0, s->stream, V, d, t, dt, src[0], src[1], src[2],
			                   fw[0], fw[1], fw[2]);
		else
			hipLaunchKernelGGL((k_sl_centered<3, D, I>), gb, tb, 0, s->stream, V, d, t, dt, src[0], src[1], src[2],
			                   fw[0], fw[1], fw[2]);
	});
	float* result = fwd;
	if (order == 2) {
		float* ng = nc == 3 ? s->getVec() : s->getReal();
		float* nw[3] = {ng, nc == 3 ? ng + s->cstride : nullptr, nc == 3 ? ng + 2 * s->cstride : nullptr};
		const bool fast = clampMode == 2 && d.nx >= 3 && d.ny >= 3 && (!d.is3d || d.nz >= 3);
		with_geo(d, [&](auto D3, auto ix) {
			constexpr bool D = decltype(D3)::value;
			using I = decltype(ix);
			if (fast && nc == 1)
				hipLaunchKernelGGL((k_maccormack_centered2<1, D, I>), gb, tb, 0, s->stream, V, d, t, dt, strength,
				                   flags->f8, src[0], src[1], src[2], (const float*)fw[0], (const float*)fw[1],
				                   (const float*)fw[2], nw[0], nw[1], nw[2]);
			else if (fast)
				hipLaunchKernelGGL((k_maccormack_centered2<3, D, I>), gb, tb, 0, s->stream, V, d, t, dt, strength,
				                   flags->f8, src[0], src[1], src[2], (const float*)fw[0], (const float*)fw[1],
				                   (const float*)fw[2], nw[0], nw[1], nw[2]);
			else if (nc == 1)
				hipLaunchKernelGGL((k_maccormack_centered<1, D, I>), gb, tb, 0, s->stream, V, d, t, dt, strength,
				                   clampMode, flags->f8, src[0], src[1], src[2], (const float*)fw[0],
				                   (const float*)fw[1], (const float*)fw[2], nw[0], nw[1], nw[2]);
			else
				hipLaunchKernelGGL((k_maccormack_centered<3, D, I>), gb, tb, 0, s->stream, V, d, t, dt, strength,
				                   clampMode, flags->f8, src[0], src[1], src[2], (const float*)fw[0],
				                   (const float*)fw[1], (const float*)fw[2], nw[0], nw[1], nw[2]);
		});
		if (nc == 3) s->putVec(fwd);
		else s->putReal(fwd);
		result = ng;
	}
	if (nc == 3) s->putVec(grid->real());
	else s->putReal(grid->real());
	grid->data = result;
}

}  // namespace mfx

using namespace mfx;

// Not part of the public interface (include/mfx.h): tests use it to run the
// advection kernels' 64-bit index path on small grids.  Returns the previous
// setting.
extern "C" int mfx_testing_force_index64(int on) {
	const int prev = g_force_index64;
	g_force_index64 = on ? 1 : 0;
	return prev;
}

struct mfx_grid : public Grid {};
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

static void same_scene(const Grid* a, const Grid* b) {
	MFX_CHECK(a->s == b->s, "grids belong to different solvers");
}

extern "C" {

int mfx_advect_semi_lagrange(const mfx_grid* flags, const mfx_grid* vel, mfx_grid* grid, int order, float strength,
                             int orderSpace, int clampMode, int orderTrace) {
	API_BEGIN
	MFX_CHECK(flags && vel && grid, "advectSemiLagrange: flags, vel and grid are required");
	MFX_CHECK(order == 1 || order == 2, "AdvectSemiLagrange: Only order 1 (regular SL) and 2 (MacCormack) supported");
	// the reference asserts "Unknown interpolation order" / "Unknown
	// backtracing order" (grid.h:150, advection.cpp:39)
	MFX_CHECK(orderSpace == 1 || orderSpace == 2, "Unknown interpolation order " + std::to_string(orderSpace));
	MFX_CHECK(orderTrace == 1 || orderTrace == 2, "Unknown backtracing order " + std::to_string(orderTrace));
	Grid *f = G(flags), *v = G(vel), *g = G(grid);
	MFX_CHECK(f->isFlags(), "advectSemiLagrange: flags must be a FlagGrid");
	MFX_CHECK(v->isMAC(), "advectSemiLagrange: vel must be a MACGrid");
	same_scene(f, v);
	same_scene(f, g);
	MFX_CHECK(g->isReal() || g->isVec(), "AdvectSemiLagrange: Grid Type is not supported (only Real, Vec3, MAC, Levelset)");
	op_advect(f->s, f, v, g, order, strength, clampMode, orderSpace, orderTrace);
	API_END
}

int mfx_set_wall_bcs(const mfx_grid* flags, mfx_grid* vel, const mfx_grid* obvel) {
	API_BEGIN
	MFX_CHECK(flags && vel, "setWallBcs: flags and vel are required");
	Grid *f = G(flags), *v = G(vel);
	MFX_CHECK(f->isFlags() && v->isMAC(), "setWallBcs: needs a FlagGrid and a MACGrid");
	same_scene(f, v);
	if (obvel) {
		MFX_CHECK(G(obvel)->isVec(), "setWallBcs: obvel must be a MACGrid");
		same_scene(f, G(obvel));
	}
	op_set_wall_bcs(f->s, f, v, obvel ? G(obvel) : nullptr);
	API_END
}

int mfx_set_wall_bcs_frac(const mfx_grid* flags, mfx_grid* vel, const mfx_grid* obvel, const mfx_grid* fractions,
                          const mfx_grid* phiObs, int boundaryWidth) {
	API_BEGIN
	MFX_CHECK(flags && vel, "setWallBcs: flags and vel are required");
	Grid *f = G(flags), *v = G(vel);
	MFX_CHECK(f->isFlags() && v->isMAC(), "setWallBcs: needs a FlagGrid and a MACGrid");
	same_scene(f, v);
	if (obvel) {
		MFX_CHECK(G(obvel)->isVec(), "setWallBcs: obvel must be a MACGrid");
		same_scene(f, G(obvel));
	}
	(void)boundaryWidth;  // unused by KnSetWallBcsFrac (extforces.cpp:220-303)
	if (!fractions || !phiObs) {
		// setWallBcs without both (extforces.cpp:308-309)
		op_set_wall_bcs(f->s, f, v, obvel ? G(obvel) : nullptr);
	} else {
		MFX_CHECK(G(fractions)->isVec(), "setWallBcs: fractions must be a MACGrid");
		MFX_CHECK(G(phiObs)->isReal(), "setWallBcs: phiObs must be a Real or Levelset grid");
		same_scene(f, G(fractions));
		same_scene(f, G(phiObs));
		op_set_wall_bcs_frac(f->s, f, v, G(phiObs));
	}
	API_END
}

int mfx_add_buoyancy(const mfx_grid* flags, const mfx_grid* density, mfx_grid* vel, const float gravity[3],
                     float coefficient, int scale) {
	API_BEGIN
	MFX_CHECK(flags && density && vel && gravity, "addBuoyancy: flags, density, vel and gravity are required");
	Grid *f = G(flags), *den = G(density), *v = G(vel);
	MFX_CHECK(f->isFlags() && den->isReal() && v->isMAC(), "addBuoyancy: needs FlagGrid, RealGrid, MACGrid");
	same_scene(f, v);
	same_scene(f, den);
	Scene* s = f->s;
	float fv[3];
	buoyancy_strength(s, gravity, coefficient, scale, fv);
	op_add_buoyancy(s, f, den, v, fv);
	API_END
}

int mfx_set_wall_bcs_buoyancy(const mfx_grid* flags, mfx_grid* vel, const mfx_grid* density, const float gravity[3],
                              float coefficient, int scale) {
	API_BEGIN
	MFX_CHECK(flags && density && vel && gravity, "setWallBcs+addBuoyancy: flags, vel, density and gravity are required");
	Grid *f = G(flags), *den = G(density), *v = G(vel);
	MFX_CHECK(f->isFlags() && den->isReal() && v->isMAC(), "addBuoyancy: needs FlagGrid, RealGrid, MACGrid");
	same_scene(f, v);
	same_scene(f, den);
	float fv[3];
	buoyancy_strength(f->s, gravity, coefficient, scale, fv);
	op_wall_buoyancy(f->s, f, den, v, fv);
	API_END
}

}  // extern "C"
