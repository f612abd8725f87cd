// Batched launches: one kernel launch covers the tiles of many independent
// problems (pressure systems of different sizes).  Every loop kernel of the
// CG and the multigrid V-cycle is written as a __device__ body taking a
// parameter struct, a logical tile index and a partial-sum slot.  Two
// wrappers instantiate it:
//   k_one   - one problem; tile = slot = XCD-aware remap of blockIdx.x
//   k_many  - a work list: the problems' tiles concatenated in problem order
//             (problem i owns logical tiles [start[i], start[i+1])); block b
//             runs logical tile XCD-remap(b) with its problem's parameters
// so the batched path runs exactly the single-problem arithmetic.
#pragma once

#include <vector>

#include "mfx_internal.h"
#include "mfx_mg.h"

namespace mfx {

// A work list on the device: start[0..np] (start[np] = total tiles).
struct WorkList {
	const int* start;
	int np;
};

// The partial-sum slot is the logical tile in both wrappers, so a problem's
// partials are reduced in the same order alone and in a batch (bit-identical
// results).
// WPE: waves per SIMD the kernel must allow (limits its registers; 1 =
// no limit beyond NT)
template <typename P, void (*BODY)(const P&, long long, long long), int NT, int WPE = 1>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) k_one(P p) {
	const long long t = xcd_tile(blockIdx.x, gridDim.x);
	BODY(p, t, t);
}

// Problem of logical tile lt: the number of problems whose start is <= lt,
// minus one (starts are nondecreasing; empty problems share their start with
// the next one and are skipped).  One load per lane and a ballot per 64
// problems, identical in every wavefront of the block.
__device__ inline int worklist_prob(const WorkList& wl, unsigned lt) {
	int cnt = 0;
	for (int base = 0; base < wl.np; base += 64) {
		const int idx = base + (int)(threadIdx.x & 63);
		const bool le = idx < wl.np && (unsigned)wl.start[idx] <= lt;
		const int c = (int)__popcll(__ballot(le));
		cnt += c;
		if (c < 64) break;
	}
	return cnt - 1;
}

template <typename P, void (*BODY)(const P&, long long, long long), int NT, int WPE = 1>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) k_many(WorkList wl,
                                                                                       const P* __restrict__ params) {
	const unsigned lt = xcd_tile(blockIdx.x, gridDim.x);
	const int prob = worklist_prob(wl, lt);
	const long long t = (long long)lt - wl.start[prob];
	BODY(params[prob], t, t);
}

// launch helpers (instantiated in the file that defines BODY)
template <typename P, void (*BODY)(const P&, long long, long long), int NT, int WPE = 1>
inline void launch_one(hipStream_t s, long long nb, const P& p) {
	if (nb > 0) hipLaunchKernelGGL((k_one<P, BODY, NT, WPE>), dim3((unsigned)nb), dim3(NT), 0, s, p);
}
template <typename P, void (*BODY)(const P&, long long, long long), int NT, int WPE = 1>
inline void launch_many(hipStream_t s, long long nb, WorkList items, const P* p) {
	if (nb > 0) hipLaunchKernelGGL((k_many<P, BODY, NT, WPE>), dim3((unsigned)nb), dim3(NT), 0, s, items, p);
}

// ---------------------------------------------------------------------------
// CG loop kernels (mfx_pressure.hip)
// ---------------------------------------------------------------------------
// A: p_new = z + beta p_old, q = A p_new, partial sums of fl(q p_new)
struct MatvecP {
	Dims d;
	int kchunk;
	const CgState* st;
	const uint8_t* code;
	const float* z;
	const float* pold;
	float* pnew;
	float* q;
	double* part;
	// scaled Laplace matrices of the other GridCg clients (cgSolveDiffusion,
	// cgSolveWE): diagonal fl(fl(cnt * scale) + 1), off-diagonals -scale
	int scaled = 0;
	float scale = 0.f;
};
// B: x += alpha p, r -= alpha q, tiles striding over the float4 groups
struct UpdateP {
	Dims d;
	long long per;  // number of tiles of this problem (stride in 256-thread units)
	const CgState* st;
	float* x;
	float* r;
	const float* p;
	const float* q;
	int withDot;
	double* part;
	float* fpart;
};
// single-block reduction of one problem's partials + scalar update
struct FinP {
	CgState* st;
	const double* part;
	const float* fpart;
	int nb;
	int phase;
	int pcNone;
};
// sum fl(z r) over a range of 256 cells
struct DotP {
	Dims d;
	const CgState* st;
	const float* z;
	const float* r;
	double* part;
};
// end of a solve: ||b - A x||^2 partials, their sum, the velocity correction
struct TrueResP {
	Dims d;
	Tiles t;
	const uint8_t* code;
	const float* b;
	const float* x;
	double* part;
	float scale;
};
struct TrueRedP {
	CgState* st;
	const double* part;
	int nb;
};
struct CorrVelP {
	Dims d;
	Tiles t;
	const uint8_t* f8;
	const float* pr;
	float *u, *v, *w;
};

// FIN_BETA0: FIN_BETA of the initial V-cycle, which also marks the next
// iteration as the first (p = z)
enum { FIN_INIT = 0, FIN_ALPHA = 1, FIN_RES = 2, FIN_BETA = 3, FIN_BETA0 = 4 };
constexpr int kMvTY = 16;

// Plane chunks of the z-marching kernels (CG matvec, level-0 multigrid) for
// one grid shape, picked by autotune_zmarch (mfx_pressure.hip) the first
// time a 3D shape is solved alone; 0 = not tuned (a heuristic chunk is
// used).  3D chunks are multiples of kDotGroup planes, at most kZmMaxChunk.
struct ZmTune {
	int mv = 0, l0 = 0;
};
ZmTune zm_tuned(const Dims& d);
void zm_set_tuned(const Dims& d, ZmTune t);
// claims the tuning of a shape for the calling thread: false if the shape is
// tuned already or another thread is tuning it (that thread publishes the
// winners with zm_set_tuned; the others use the heuristic chunk meanwhile)
bool zm_claim_tuning(const Dims& d);

// Dot products of the z-marching kernels (fl(q p) of the matvec, fl(z r) of
// the level-0 up-sweep) leave one partial per tile column and group of
// kDotGroup planes: a thread sums its cells of the group in plane order, the
// wavefront sums its lanes (fixed DPP shift-add order), the tile sums its
// wavefronts in order.  With chunks that start at multiples of kDotGroup this partition
// does not depend on the chunk, so every chunk gives the same bits.
constexpr int kDotGroup = 4;
constexpr int kZmMaxChunk = 64;
constexpr int kZmMaxGroups = kZmMaxChunk / kDotGroup;
inline long long zm_dot_parts(const Dims& d, int ty) {
	return (long long)((d.nx + 63) / 64) * ((d.ny + ty - 1) / ty) * ((d.nz + kDotGroup - 1) / kDotGroup);
}
// a 3D chunk of at least kc planes that the dot partition allows
inline int zm_chunk(int kc) {
	kc = (kc + kDotGroup - 1) / kDotGroup * kDotGroup;
	return kc < kDotGroup ? kDotGroup : (kc > kZmMaxChunk ? kZmMaxChunk : kc);
}

// the calling block's tile column (its tile index modulo the columns)
__device__ inline long long zm_column(long long blk, const Dims& d, int ty) {
	const long long cols = (long long)((d.nx + 63) / 64) * ((d.ny + ty - 1) / ty);
	return blk % cols;
}
// end of a dot group at plane k: every thread of the block calls this; the
// wavefront sum goes to tab[group - first group of the tile][wavefront]
template <int TY>
__device__ inline void zm_dot_flush(double& acc, double (*tab)[TY], int k, int k0) {
	const double s = wave_sum_to_last(acc);
	if ((threadIdx.x & 63) == 63) tab[k / kDotGroup - k0 / kDotGroup][threadIdx.x >> 6] = s;
	acc = 0.0;
}
// after a barrier that follows the last flush: one thread per group of the
// tile writes the group's partial
template <int TY>
__device__ inline void zm_dot_store(const double (*tab)[TY], const Dims& d, long long col, int k0, int k1,
                                    double* part) {
	const int g0 = k0 / kDotGroup, ng = (k1 - 1) / kDotGroup - g0 + 1;
	const int t = threadIdx.x;
	if (t < ng) {
		double s = 0.0;
		for (int w = 0; w < TY; w++) s += tab[t][w];
		const long long groups = (d.nz + kDotGroup - 1) / kDotGroup;
		part[3 * (col * groups + g0 + t)] = s;
	}
}

// tiles of each CG kernel for one problem (and the plane chunk / range
// size); tuned = the autotuned chunk of the shape where there is one (single
// solves), else the heuristic (batched solves)
long long matvec_tiles(const Dims& d, int* kchunk, bool tuned, int force = 0);
long long update_tiles(const Dims& d, long long* per);
inline long long dot_tiles(const Dims& d) { return (d.n + kBlock - 1) / kBlock; }

// launches: *_one with a host parameter struct, *_many over a device work list
void run_matvec_one(hipStream_t s, unsigned nb, const MatvecP& p);
void run_matvec_many(hipStream_t s, unsigned nb, WorkList items, const MatvecP* p);
void run_update_one(hipStream_t s, unsigned nb, const UpdateP& p);
void run_update_many(hipStream_t s, unsigned nb, WorkList items, const UpdateP* p);
void run_finalize_one(hipStream_t s, const FinP& p);
void run_finalize_many(hipStream_t s, unsigned nprob, WorkList items, const FinP* p);
void run_dot_one(hipStream_t s, unsigned nb, const DotP& p);
void run_dot_many(hipStream_t s, unsigned nb, WorkList items, const DotP* p);

// ---------------------------------------------------------------------------
// V-cycle kernels (mfx_mg.hip)
// ---------------------------------------------------------------------------
struct L0DownP {  // level-0 pre-smooth from zero + residual
	Dims d;
	int kchunk;
	const CgState* st;
	const uint8_t* code;
	const float* rin;
	float* xout;
	float* r0;
};
struct L0UpP {  // prolongation + level-0 post-smooth + partial sums of fl(z r)
	Dims d;
	int kchunk;
	const CgState* st;
	const uint8_t* code;
	const float* rin;
	const float* xpre;
	float* z;
	LvDev c1;
	const float* xc;
	double* part;
};
struct RestrictP {
	const CgState* st;
	LvDev c, f;
	const uint8_t* ctype;
	const float* rf;
	float* bc;
};
struct SmoothP {
	const CgState* st;
	LvDev c;
	const void* A;  // float or BF16 by the kernel's template argument
	const uint8_t* type;
	const float* b;
	const float* xin;
	float* xout;
	int first, reverse;  // the mode of the launch: pre-smoothing 1/0, post-smoothing 0/1
	LvDev cp;
	const float* xcp;
	Canon cn;
};
struct ResidP {
	const CgState* st;
	LvDev c;
	const void* A;
	const uint8_t* type;
	const float* b;
	const float* x;
	float* r;
	Canon cn;
};
struct CoarsestP {
	const CgState* st;
	int nc;
	const int* cidx;
	const double* G;
	const float* b;
	const uint8_t* code;
	float* x;
};
struct FillP {
	const CgState* st;
	float* x;
	long long n;
};

// parameters of every step of one problem's V-cycle
struct MgHierarchy::Steps {
	int nl = 0;
	int kchunk = 1;
	long long l0Tiles = 0;
	long long nPart = 0;  // partials of fl(z r) left by the up-sweep
	L0DownP down{};
	L0UpP up{};
	struct Level {  // levels 1..nl-1
		RestrictP restrict_{};
		SmoothP pre{}, post{};  // not used on the coarsest level
		ResidP resid{};
		long long tiles = 0, blkTiles = 0;
	};
	std::vector<Level> lv;
	CoarsestP coarsest{};
	FillP fill{};  // single-level hierarchies: z = 0 before the coarsest solve
};

// tiles of the V-cycle kernels for one problem/level
// force > 0 uses that plane chunk (the autotuner's trials) instead of the
// heuristic or tuned one
long long l0_tiles(const Dims& d, int* kchunk, bool tuned, int force = 0);  // z-marching (3D) or brick (2D)
long long lc_tiles(const LvDev& c);               // 64x4 vertex tiles per plane
long long lcblk_tiles(const LvDev& c, bool is3d); // smoother bricks
inline long long coarsest_tiles(int nc) { return (nc + 3) / 4; }

void run_l0_down_many(hipStream_t s, bool is3d, unsigned nb, WorkList items, const L0DownP* p);
void run_l0_down_one(hipStream_t s, bool is3d, unsigned nb, const L0DownP& p);
constexpr int kZmTYHost = 16;  // rows per z-marching level-0 tile (wavefronts per workgroup)
void run_l0_up_many(hipStream_t s, bool is3d, unsigned nb, WorkList items, const L0UpP* p);
void run_restrict_many(hipStream_t s, bool is3d, unsigned nb, WorkList items, const RestrictP* p);
// post: post-smoothing (prolongation on load, reversed colours), else pre-smoothing
void run_smooth_many(hipStream_t s, bool is3d, bool bf16, bool post, unsigned nb, WorkList items,
                     const SmoothP* p);
void run_resid_many(hipStream_t s, bool is3d, bool bf16, unsigned nb, WorkList items, const ResidP* p);
void run_coarsest_many(hipStream_t s, unsigned nb, WorkList items, const CoarsestP* p);
void run_fill_many(hipStream_t s, unsigned nb, WorkList items, const FillP* p);

}  // namespace mfx
