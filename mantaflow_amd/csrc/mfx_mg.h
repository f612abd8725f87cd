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
};

struct MgLevel {
	int nx = 0, ny = 0, nz = 0;
	long long n = 0;
	long long sy = 0, sz = 0;   // pitches (sz = nx*ny, also in 2D where nz == 1)
	float* A = nullptr;          // S arrays of n floats (levels >= 1), symmetric upper half
	float* x = nullptr;          // pre-smoothed iterate (level 0: before the up leg)
	float* xp = nullptr;         // post-smoothed iterate (levels 1..L-2)
	float* b = nullptr;
	float* r = nullptr;
	uint8_t* type = nullptr;     // 1 active, 0 inactive
};

class MgHierarchy {
public:
	~MgHierarchy();
	// builds levels, coarse operators and the coarsest-level inverse from the
	// level-0 matrix code
	void build(Scene* s, const uint8_t* code, int precision);
	void setTrivial(const uint8_t*) {}
	// z = V-cycle(r) from z = 0; kernels skip their work once st->done is set.
	// Returns the number of partial sums of fl(z*r) the last kernel wrote into
	// s->dpart (stride 3), or 0 if the caller has to compute dot(z, r) itself.
	int vcycle(Scene* s, const uint8_t* code, const float* r, float* z, const CgState* st);
	int levels() const { return (int)L.size(); }

private:
	bool is3d = true;
	int S = 14;
	std::vector<MgLevel> L;      // L[0] holds only r (x = z, b = rhs on the fly)
	double* G = nullptr;         // dense generalized inverse of the coarsest operator
	int* cidx = nullptr;         // active coarsest vertices
	int nc = 0;
	float* bc = nullptr;         // gathered coarsest rhs (double via G)
	std::vector<void*> allocs;
	void* alloc(size_t bytes);
	void build_coarsest(Scene* s, const uint8_t* code0);
};

}  // namespace mfx
