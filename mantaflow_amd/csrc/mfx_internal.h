// Internal declarations of libmfx: scene/grid objects, launch helpers and
// the device-side cell-flag helpers shared by all kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mfx.h"

namespace mfx {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error : public std::runtime_error {
	explicit Error(const std::string& s) : std::runtime_error(s) {}
};

void set_last_error(const std::string& s);

#define MFX_HIP(call)                                                                  \
	do {                                                                               \
		hipError_t _e = (call);                                                        \
		if (_e != hipSuccess)                                                          \
			throw ::mfx::Error(std::string("HIP error: ") + hipGetErrorString(_e) +   \
			                   " at " + __FILE__ + ":" + std::to_string(__LINE__));   \
	} while (0)

#define MFX_CHECK(cond, msg)                      \
	do {                                          \
		if (!(cond)) throw ::mfx::Error(msg);     \
	} while (0)

// ---------------------------------------------------------------------------
// cell types (grid.h:292-304); the device keeps a uint8 copy of bits 0..6
// ---------------------------------------------------------------------------
enum : int {
	kFluid = 1, kObstacle = 2, kEmpty = 4, kInflow = 8, kOutflow = 16, kOpen = 32, kStick = 64
};

// grid type bits (grid.h:29)
enum : int { kTypeReal = 1, kTypeInt = 2, kTypeVec3 = 4, kTypeMAC = 8, kTypeLevelset = 16, kTypeFlags = 32 };

// ---------------------------------------------------------------------------
// geometry passed by value to kernels
// ---------------------------------------------------------------------------
struct Dims {
	int nx, ny, nz;
	int is3d;
	long long n;   // cells
	long long sz;  // z stride: nx*ny in 3D, 0 in 2D (grid.cpp:55)
	__host__ __device__ inline long long idx(int i, int j, int k) const {
		return (long long)i + (long long)nx * j + sz * k;
	}
};

// ---------------------------------------------------------------------------
// scene / grid objects
// ---------------------------------------------------------------------------
struct Grid;
class MgHierarchy;

struct Scene {
	Dims d;
	int dim;
	float dt = 1.0f;
	hipStream_t stream = nullptr;
	int device = 0;
	size_t cstride = 0;  // per-component stride of Vec3/MAC grids (floats)
	// pooled temporaries (mirror of FluidSolver::GridStorage)
	std::vector<float*> freeReal;
	std::vector<float*> freeVec;
	std::vector<void*> owned;
	// small device workspace for reductions
	double* dpart = nullptr;    // block partials (doubles)
	size_t dpartCap = 0;
	float* fpart = nullptr;     // block partials (floats, min/max)
	size_t fpartCap = 0;
	void* dscal = nullptr;      // device scalar block
	void* hscal = nullptr;      // pinned host mirror
	hipEvent_t ev = nullptr;    // cross-stream ordering (batched solves)
	hipEvent_t evState = nullptr;  // a CG state copy the host waits for
	void* arena = nullptr;      // work lists of batched solves led by this scene
	size_t arenaCap = 0;
	void* hstage = nullptr;     // pinned staging of the arena upload
	size_t hstageCap = 0;
	// two-iteration graph of the last batch composition led by this scene
	// that was seen twice in a row (key: launch layout)
	hipGraphExec_t bexec = nullptr;
	std::vector<long long> bkey, blastKey;
	std::unique_ptr<MgHierarchy> mgStatic;
	// multigrid hierarchy standing in for PcMIC (the reference's mIC keeps
	// nothing between solves): rebuilt whenever the matrix code changes,
	// checked by the flag grid's generation and a hash of the code
	std::unique_ptr<MgHierarchy> mgMic;
	const Grid* micFlags = nullptr;
	uint64_t micGen = 0, micHash = 0;
	long long micFix = -1;
	uint64_t* dhash = nullptr;  // device word of the code hash
	uint64_t* hhash = nullptr;  // its pinned host mirror
	uint8_t* code = nullptr;    // per-cell Laplace matrix code of the current solve
	std::vector<Grid*> grids;
	~Scene();
	float* getReal();
	void putReal(float* p);
	float* getVec();
	void putVec(float* p);
	void ensurePartials(size_t nd, size_t nf);
};

struct Grid {
	Scene* s;
	int type;
	void* data = nullptr;           // float[n] / int32[n] / float[3*cstride] (SoA)
	uint8_t* f8 = nullptr;          // flag grids: packed bits 0..6
	uint64_t gen = 0;               // flag grids: bumped by every write (flags_touched)
	// flag grids: indices of the outflow cells (outflow_list), valid while
	// olistGen == ogen; ogen is bumped by every write that can change which
	// cells carry the outflow flag (all but resetOutflow's)
	unsigned* olist = nullptr;
	unsigned* ocount = nullptr;
	uint64_t ogen = 1, olistGen = 0;
	bool external = false;
	float* comp(int c) const { return (float*)data + (size_t)c * s->cstride; }
	float* real() const { return (float*)data; }
	int* ints() const { return (int*)data; }
	bool isVec() const { return (type & kTypeVec3) != 0; }
	bool isMAC() const { return (type & kTypeMAC) != 0; }
	bool isFlags() const { return (type & kTypeFlags) != 0; }
	bool isInt() const { return (type & kTypeInt) != 0; }
	bool isReal() const { return (type & kTypeReal) != 0; }
};

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
constexpr int kTX = 64;  // threads along x per tile (one wavefront row)
constexpr int kTY = 4;   // rows per tile
constexpr int kBlock = kTX * kTY;

// Unsigned division by a divisor fixed at launch, for dividends below 2^31:
// q = (mulhi(n, m) + n) >> l with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, i.e. the 33-bit "round-up" multiplier
// of Granlund & Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994), Thm 4.2 (exact for n < 2^32; n < 2^31 keeps
// the 32-bit sum from overflowing).  Three scalar instructions instead of a
// 64-bit division expansion in every tile decode.
struct FastDiv {
	unsigned d = 1, m = 1, l = 0;
};
inline FastDiv fastdiv_make(unsigned d) {
	if (d == 0) throw std::runtime_error("fastdiv: zero divisor");
	FastDiv f;
	f.d = d;
	while ((1ull << f.l) < d) f.l++;
	f.m = (unsigned)(((1ull << 32) * ((1ull << f.l) - d)) / d + 1);
	return f;
}
__host__ __device__ inline unsigned fastdiv(unsigned n, const FastDiv& f) {
	return ((unsigned)(((unsigned long long)n * f.m) >> 32) + n) >> f.l;
}

struct Tiles {
	int tx, ty, tz;  // tiles per axis (tz = ceil(nz / kz))
	int kz = 1;      // planes per tile
	FastDiv dx, dy;  // division by tx and ty
	long long count() const { return (long long)tx * ty * tz; }
};
inline Tiles tiles_for(const Dims& d, int kz = 1) {
	Tiles t;
	t.tx = (d.nx + kTX - 1) / kTX;
	t.ty = (d.ny + kTY - 1) / kTY;
	t.kz = kz;
	t.tz = (d.nz + kz - 1) / kz;
	if (t.count() >= (1ll << 31)) throw std::runtime_error("grid has too many tiles for one launch");
	t.dx = fastdiv_make((unsigned)t.tx);
	t.dy = fastdiv_make((unsigned)t.ty);
	return t;
}

// Map the hardware workgroup id to a logical tile so that each of the 8 XCDs
// (dispatched round-robin by workgroup id) processes one contiguous run of
// tiles: neighbouring tiles then share halo lines in the same L2.
__device__ inline unsigned xcd_tile(unsigned wg, unsigned nwg) {
	const unsigned q = nwg >> 3, r = nwg & 7;
	const unsigned x = wg & 7, slot = wg >> 3;
	return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// KnSetWallBcs (extforces.cpp:186-217) at cell (i,j,k) as 0/1 integer
// predicates, from the cell's flags and its six neighbours' (a neighbour's
// flags are used only where it exists: the reference's i>0 / i<size-1
// tests).  w* = the face is written: a face next to an obstacle, or an
// obstacle face next to fluid, takes the obstacle velocity; s* = a stick
// wall across that axis, which zeroes the two other components of a fluid
// cell and overrides the obstacle value (the stick tests come last in the
// reference).  In 2D the z face of every fluid or obstacle cell is written
// with 0.
struct WallRule {
	int wu, wv, ww, sx, sy, sz;
};
__device__ inline WallRule wall_rule(const Dims& d, int i, int j, int k, int fl, int xm, int xp, int ym, int yp,
                                     int zm, int zp) {
	const int curF = fl & 1, curO = (fl >> 1) & 1;  // kFluid = 1, kObstacle = 2, kStick = 64
	const int act = curF | curO;
	const int gi = min(i, 1), li = min(d.nx - 1 - i, 1), gj = min(j, 1), lj = min(d.ny - 1 - j, 1);
	const int gk = min(k, 1), lk = min(d.nz - 1 - k, 1), is3 = d.is3d ? 1 : 0;
	const int ou = act & gi & (((xm >> 1) & 1) | (curO & xm & 1));
	const int ov = act & gj & (((ym >> 1) & 1) | (curO & ym & 1));
	const int ow = act & ((is3 ^ 1) | (gk & (((zm >> 1) & 1) | (curO & zm & 1))));
	WallRule r;
	r.sx = curF & ((gi & (xm >> 6)) | (li & (xp >> 6))) & 1;
	r.sy = curF & ((gj & (ym >> 6)) | (lj & (yp >> 6))) & 1;
	r.sz = curF & is3 & ((gk & (zm >> 6)) | (lk & (zp >> 6))) & 1;
	r.wu = ou | r.sy | r.sz;
	r.wv = ov | r.sx | r.sz;
	r.ww = ow | r.sx | r.sy;
	return r;
}

// Decode the calling thread's cell (i,j,k) from a 1-D tile launch; k is the
// first of the tile's kz planes.
struct CellPos {
	int i, j, k;
};
// cells of logical tile `tile` (the thread's column of kTX x kTY rows)
__device__ inline CellPos tile_cell_of(const Tiles& t, unsigned tile) {
	const unsigned rest = fastdiv(tile, t.dx);
	const unsigned bx = tile - rest * (unsigned)t.tx;
	const unsigned k = fastdiv(rest, t.dy);
	const unsigned by = rest - k * (unsigned)t.ty;
	CellPos p;
	p.i = (int)bx * kTX + (int)(threadIdx.x & (kTX - 1));
	p.j = (int)by * kTY + (int)(threadIdx.x >> 6);
	p.k = (int)k * t.kz;
	return p;
}
__device__ inline CellPos tile_cell(const Tiles& t) { return tile_cell_of(t, xcd_tile(blockIdx.x, gridDim.x)); }

// Planes per thread of the light per-cell kernels: a thread walks kLightKZ
// planes of its 64x4 tile and issues the loads of all of them before the
// first use, so their memory latencies overlap.
constexpr int kLightKZ = 4;

// flat launch over n elements with a grid-stride loop
inline unsigned flat_blocks(long long n, int per = 256, unsigned cap = 65536u) {
	long long b = (n + per - 1) / per;
	if (b < 1) b = 1;
	return (unsigned)(b > cap ? cap : b);
}

// ---------------------------------------------------------------------------
// reductions (block level, deterministic)
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}
// Sum over the wavefront of a double, valid in lane 63 only: shift-adds
// within rows of 16 lanes (DPP row_shr 1, 2, 4, 8), then rows into rows
// (row_bcast15 into rows 1 and 3, row_bcast31 into rows 2 and 3).  Lanes
// that a DPP move does not write read 0.0.  Data-path moves instead of the
// LDS crossbar of __shfl_xor; the summation order is fixed.
template <int CTRL, int ROWS>
__device__ inline double dpp_move_d(double v) {
	const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
	const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, ROWS, 0xf, false);
	const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, ROWS, 0xf, false);
	return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ inline double wave_sum_to_last(double v) {
	v += dpp_move_d<0x111, 0xf>(v);  // row_shr:1
	v += dpp_move_d<0x112, 0xf>(v);  // row_shr:2
	v += dpp_move_d<0x114, 0xf>(v);  // row_shr:4
	v += dpp_move_d<0x118, 0xf>(v);  // row_shr:8
	v += dpp_move_d<0x142, 0xa>(v);  // row_bcast:15
	v += dpp_move_d<0x143, 0xc>(v);  // row_bcast:31
	return v;
}
__device__ inline float wave_min(float v) {
	for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
	return v;
}
__device__ inline float wave_max(float v) {
	for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
	return v;
}

// Sum over a block of kBlock threads; result valid in thread 0.
template <typename T>
__device__ inline T block_sum(T v, T* sh) {
	v = wave_sum(v);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) sh[w] = v;
	__syncthreads();
	T r = 0;
	if (threadIdx.x == 0) {
		for (int q = 0; q < (int)(blockDim.x >> 6); q++) r += sh[q];
	}
	__syncthreads();
	return r;
}

// ---------------------------------------------------------------------------
// host-side API helpers (defined in mfx_grid.hip)
// ---------------------------------------------------------------------------
Grid* make_grid(Scene* s, int type);
void free_grid(Grid* g);
void flags_repack(Grid* flags);      // int32 -> uint8 mirror (also flags_touched)
void flags_touched(Grid* flags);     // new generation number for a written flag grid
// every live scene (releaseMG() without a solver releases the hierarchies of all)
std::vector<Scene*> live_scenes();
void sync(Scene* s);
double reduce_sum_double(Scene* s, const float* a, int mode, long long n);

// kernels / ops implemented in the various .hip files
void op_set_wall_bcs(Scene* s, Grid* flags, Grid* vel, Grid* obvel);
// Indices of the flag grid's outflow cells, in no particular order, and
// their number (*count, on the device), rebuilt on the scene's stream only
// when a write may have changed which cells are outflow cells.  The kernels
// that only touch outflow cells run over the list (grid-stride over *count)
// instead of over every cell.
const unsigned* outflow_list(Grid* flags, const unsigned** count);
constexpr unsigned kOutflowListBlocks = 1024;
void op_add_buoyancy(Scene* s, Grid* flags, Grid* density, Grid* vel, const float f[3]);
// setWallBcs (no obstacle velocity) then addBuoyancy, one pass
void op_wall_buoyancy(Scene* s, Grid* flags, Grid* density, Grid* vel, const float f[3]);
void buoyancy_strength(const Scene* s, const float gravity[3], float coefficient, int scale, float fv[3]);
void op_advect(Scene* s, Grid* flags, Grid* vel, Grid* grid, int order, float strength, int clampMode, int orderSpace,
               int orderTrace);

}  // namespace mfx
