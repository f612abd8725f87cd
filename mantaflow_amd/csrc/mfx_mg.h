// Device-side CG state and the multigrid preconditioner hierarchy.
#pragma once

#include "mfx_internal.h"

namespace mfx {

// scalars of one CG solve, kept in device memory (GridCg members
// mSigma/mResNorm/mIterations plus convergence control)
struct CgState {
	double bnorm2;   // ||b||_2^2 (relative criterion)
	double rr;       // last sum r^2
	double relTol;
	float sigma, alpha, beta, resNorm, acc;
	int iter, maxIter, done, converged, diverged, useL2, first, pad;
	float hist[64];  // residual norm per iteration (batch sizing)
	double trueRR;   // ||b - A x||^2 with FP64 row sums, after the solve
	// zero-pressure fixing (pressure.cpp:349-388), chosen on the device
	long long fix;                 // pinned cell, -1 if none
	unsigned long long fixFirst;   // first fluid cell of the bnd=1 interior (~0: none)
	int fixAnyEmpty;               // some cell is empty: nothing is pinned
	int fixPad;
};

// canonical full stencil of a coarse level (27 entries in 3D, 9 in 2D,
// entry S-1 = diagonal); vertices of type 2 use it instead of their arrays
struct Canon {
	float a[27];
};

// level geometry as the kernels see it
struct LvDev {
	int nx, ny, nz;
	long long sy, sz, n;
	FastDiv fdx, fdy;  // division by nx and ny (vertex index decode)
};

struct MgLevel {
	int nx = 0, ny = 0, nz = 0;
	long long n = 0;
	long long sy = 0, sz = 0;   // pitches (sz = nx*ny, also in 2D where nz == 1)
	float* A = nullptr;          // S arrays of n floats (levels >= 1), symmetric upper half
	uint16_t* Ah = nullptr;      // the same coefficients in BF16 (mixed-precision mode only)
	float* x = nullptr;          // pre-smoothed iterate (level 0: before the up leg)
	float* xp = nullptr;         // post-smoothed iterate (levels 1..L-2)
	float* b = nullptr;
	float* r = nullptr;
	uint8_t* type = nullptr;     // 0 inactive, 1 active, 2 active with the canonical stencil
	Canon canon{};
	bool hasCanon = false;
};

class MgHierarchy {
public:
	~MgHierarchy();
	// builds levels, coarse operators and the coarsest-level inverse from the
	// level-0 matrix code
	void build(Scene* s, const uint8_t* code, int precision);
	// z = V-cycle(r) from z = 0; kernels skip their work once st->done is set.
	// Returns the number of partial sums of fl(z*r) the last kernel wrote into
	// s->dpart (stride 3), or 0 if the caller has to compute dot(z, r) itself.
	int vcycle(Scene* s, const uint8_t* code, const float* r, float* z, const CgState* st);
	// parameters of every V-cycle step (for batched work lists); tuned = the
	// level-0 plane chunk of single solves, else the batched heuristic
	struct Steps;
	Steps steps(const Dims& d, const uint8_t* code, const float* r, float* z, const CgState* st, double* part,
	            bool tuned, int forceChunk = 0) const;
	bool dim3d() const { return is3d; }
	int levels() const { return (int)L.size(); }
	int precision() const { return prec; }

private:
	bool is3d = true;
	int S = 14;
	int prec = 0;                // 0: FP32 coefficients, 1: BF16 smoother/residual coefficients
	std::vector<MgLevel> L;      // L[0] holds only r (x = z, b = rhs on the fly)
	double* G = nullptr;         // dense generalized inverse of the coarsest operator
	int* cidx = nullptr;         // active coarsest vertices
	int nc = 0;
	float* bc = nullptr;         // gathered coarsest rhs (double via G)
	std::vector<void*> allocs;
	void* alloc(size_t bytes);
	void build_coarsest(Scene* s, const uint8_t* code0);
	void classify(Scene* s, int l);
};

}  // namespace mfx
