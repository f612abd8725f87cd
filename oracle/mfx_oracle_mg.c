/*
 * CPU oracle, part 2: restatement of the reference's multigrid preconditioner
 * GridMg (multigrid.cpp, after Dick et al. 2015).  TEST INFRASTRUCTURE ONLY.
 *
 * Follows multigrid.cpp step by step: level sizes (:220-262), precomputed
 * level-1 coarsening paths (:273-305), setA with trivial-equation scaling
 * (:323-415), the sequential bucket-min-heap coarse-vertex selection
 * (NKMinHeap :59-203, genCoarseGrid :520-578), Galerkin operators (:580-666),
 * multi-colour Gauss-Seidel (:668-737), residual (:739-791), FP64 Jacobi-PCG on
 * the coarsest level (:796-902), restriction/interpolation (:904-959) and the
 * V-cycle (:448-504).  Stencil storage is the reference's symmetric one:
 * 4 (3 in 2D) entries per vertex on level 0, 14 (5) on coarser levels.
 *
 * The level-1 paths are sorted with a stable insertion sort on the
 * reference's key (stencil index, then fine offset); std::sort leaves the
 * order of equal keys unspecified, which can change float rounding of the
 * level-1 operator in the last bit when trivial-scaled rows are involved.
 */
#include "mfx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for num_threads(og_get_threads()) if(og_get_threads() > 1) schedule(static)")
#else
#define OMP_FOR
#endif

#define MAXL 32

typedef struct { int x, y, z; } v3i;

static inline v3i mk(int x, int y, int z) { v3i v = {x, y, z}; return v; }
static inline int dot139(v3i v) { return v.x + 3 * v.y + 9 * v.z; }
static inline int vget(v3i v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

enum { vtInactive = 0, vtActive = 1, vtActiveTrivial = 2, vtRemoved = 3, vtZero = 4, vtFree = 5 };

typedef struct {
	v3i U, W, N;
	int sc, sf;
	og_real rw, iw;
	int inUStencil;
} cpath;

struct og_mg {
	int is3d, dim, S, S0;
	v3i smin, smax;
	int nl;
	v3i size[MAXL], pitch[MAXL];
	long n[MAXL];
	og_real *A[MAXL], *x[MAXL], *b[MAXL], *r[MAXL];
	signed char* type[MAXL];
	double *cg1, *cg2, *cg3, *cg4;
	cpath* paths;
	int npaths;
	og_real coarsestAcc, trivialScale;
	int nPre, nPost, isASet;
};

static inline v3i vecIdx(const og_mg* m, long v, int l) {
	long sx = m->size[l].x, sxy = (long)m->size[l].x * m->size[l].y;
	return mk((int)(v % sx), (int)((v % sxy) / sx), (int)(v / sxy));
}
static inline long linIdx(const og_mg* m, v3i V, int l) {
	return (long)V.x + (long)V.y * m->pitch[l].y + (long)V.z * m->pitch[l].z;
}
static inline int inGrid(const og_mg* m, v3i V, int l) {
	return V.x >= 0 && V.y >= 0 && V.z >= 0 && V.x < m->size[l].x && V.y < m->size[l].y &&
	       V.z < m->size[l].z;
}

/* GridMg::GridMg (multigrid.cpp:220-319) */
og_mg* og_mg_create(og_dims d) {
	og_mg* m = (og_mg*)calloc(1, sizeof(og_mg));
	m->is3d = d.nz > 1;
	m->dim = m->is3d ? 3 : 2;
	m->S = m->is3d ? 14 : 5;
	m->S0 = m->is3d ? 4 : 3;
	m->smin = mk(-1, -1, m->is3d ? -1 : 0);
	m->smax = mk(1, 1, m->is3d ? 1 : 0);
	m->coarsestAcc = (og_real)1E-8;
	m->trivialScale = (og_real)1E-6;
	m->nPre = m->nPost = 1;
	m->size[0] = mk(d.nx, d.ny, d.nz);
	long n = (long)d.nx * d.ny * d.nz;
	m->nl = 1;
	for (int l = 1; l <= 100 && l < MAXL; l++) {
		v3i p = m->size[l - 1];
		if (p.x <= 5 && p.y <= 5 && p.z <= 5) break;
		if (n <= 1000) break;
		m->size[l] = mk((p.x + 2) / 2, (p.y + 2) / 2, (p.z + 2) / 2);
		n = (long)m->size[l].x * m->size[l].y * m->size[l].z;
		m->nl++;
	}
	for (int l = 0; l < m->nl; l++) {
		v3i s = m->size[l];
		m->pitch[l] = mk(1, s.x, s.x * s.y);
		m->n[l] = (long)s.x * s.y * s.z;
		int ss = l == 0 ? m->S0 : m->S;
		m->A[l] = (og_real*)calloc((size_t)(m->n[l] * ss), sizeof(og_real));
		m->x[l] = (og_real*)calloc((size_t)m->n[l], sizeof(og_real));
		m->b[l] = (og_real*)calloc((size_t)m->n[l], sizeof(og_real));
		m->r[l] = (og_real*)calloc((size_t)m->n[l], sizeof(og_real));
		m->type[l] = (signed char*)calloc((size_t)m->n[l], 1);
	}
	long nc = m->n[m->nl - 1];
	m->cg1 = (double*)calloc((size_t)nc, sizeof(double));
	m->cg2 = (double*)calloc((size_t)nc, sizeof(double));
	m->cg3 = (double*)calloc((size_t)nc, sizeof(double));
	m->cg4 = (double*)calloc((size_t)nc, sizeof(double));

	/* precomputed coarsening paths (V) <-R- (U) <-A- (W) <-I- (N) */
	v3i p7[7] = {mk(0, 0, 0), mk(-1, 0, 0), mk(1, 0, 0), mk(0, -1, 0), mk(0, 1, 0), mk(0, 0, -1), mk(0, 0, 1)};
	v3i V = mk(1, 1, 1);
	int cap = 4096;
	m->paths = (cpath*)malloc(sizeof(cpath) * (size_t)cap);
	m->npaths = 0;
	for (int uz = 2 * V.z + m->smin.z; uz <= 2 * V.z + m->smax.z; uz++)
		for (int uy = 2 * V.y + m->smin.y; uy <= 2 * V.y + m->smax.y; uy++)
			for (int ux = 2 * V.x + m->smin.x; ux <= 2 * V.x + m->smax.x; ux++) {
				v3i U = mk(ux, uy, uz);
				for (int i = 0; i < 1 + 2 * m->dim; i++) {
					v3i W = mk(U.x + p7[i].x, U.y + p7[i].y, U.z + p7[i].z);
					for (int nz = W.z / 2; nz <= (W.z + 1) / 2; nz++)
						for (int ny = W.y / 2; ny <= (W.y + 1) / 2; ny++)
							for (int nx = W.x / 2; nx <= (W.x + 1) / 2; nx++) {
								v3i N = mk(nx, ny, nz);
								int s = dot139(N);
								if (s >= 13) {
									cpath p;
									p.N = mk(N.x - 1, N.y - 1, N.z - 1);
									p.U = mk(U.x - 2 * V.x, U.y - 2 * V.y, U.z - 2 * V.z);
									p.W = mk(W.x - 2 * V.x, W.y - 2 * V.y, W.z - 2 * V.z);
									p.sc = s - 13;
									p.sf = (i + 1) / 2;
									p.inUStencil = (i % 2 == 0);
									p.rw = (og_real)1 / (og_real)(1 << ((U.x % 2) + (U.y % 2) + (U.z % 2)));
									p.iw = (og_real)1 / (og_real)(1 << ((W.x % 2) + (W.y % 2) + (W.z % 2)));
									if (m->npaths < cap) m->paths[m->npaths++] = p;
								}
							}
				}
			}
	/* stable sort by (sc, dot(U+1,(1,3,9))) */
	for (int a = 1; a < m->npaths; a++) {
		cpath key = m->paths[a];
		int kk = key.sc * 100 + dot139(mk(key.U.x + 1, key.U.y + 1, key.U.z + 1));
		int b = a - 1;
		while (b >= 0) {
			cpath* q = &m->paths[b];
			int kq = q->sc * 100 + dot139(mk(q->U.x + 1, q->U.y + 1, q->U.z + 1));
			if (kq <= kk) break;
			m->paths[b + 1] = m->paths[b];
			b--;
		}
		m->paths[b + 1] = key;
	}
	return m;
}

void og_mg_destroy(og_mg* m) {
	if (!m) return;
	for (int l = 0; l < m->nl; l++) {
		free(m->A[l]); free(m->x[l]); free(m->b[l]); free(m->r[l]); free(m->type[l]);
	}
	free(m->cg1); free(m->cg2); free(m->cg3); free(m->cg4);
	free(m->paths);
	free(m);
}

int og_mg_num_levels(const og_mg* m) { return m->nl; }
void og_mg_level_size(const og_mg* m, int l, int* sx, int* sy, int* sz) {
	*sx = m->size[l].x; *sy = m->size[l].y; *sz = m->size[l].z;
}
int og_mg_level_types(const og_mg* m, int l, int8_t* out) {
	int cnt = 0;
	for (long v = 0; v < m->n[l]; v++) {
		out[v] = (int8_t)m->type[l][v];
		if (m->type[l][v] != vtInactive) cnt++;
	}
	return cnt;
}
void og_mg_set_coarsest_accuracy(og_mg* m, og_real acc) { m->coarsestAcc = acc; }
/* stencil storage of level l: n*S0 (level 0) or n*S (coarse) values */
int og_mg_level_matrix(const og_mg* m, int l, og_real* out) {
	const int ss = l == 0 ? m->S0 : m->S;
	memcpy(out, m->A[l], sizeof(og_real) * (size_t)(m->n[l] * ss));
	return ss;
}

/* ---- NKMinHeap (multigrid.cpp:59-203) ---- */
typedef struct { int key, prev, next; } hent;
typedef struct { int N, K, size, minKey; hent* e; } nkheap;

static void heap_init(nkheap* h, int N, int K) {
	h->N = N; h->K = K; h->size = 0; h->minKey = -1;
	h->e = (hent*)malloc(sizeof(hent) * (size_t)(N + K));
	for (int i = 0; i < N + K; i++) { h->e[i].key = -1; h->e[i].prev = -1; h->e[i].next = -1; }
}
static void heap_set(nkheap* h, int ID, int key) {
	const int kid = h->K + ID;
	if (h->e[kid].key == key) return;
	if (h->e[kid].key != -1) {
		int pred = h->e[kid].prev, succ = h->e[kid].next;
		h->e[pred].next = succ;
		if (succ != -1) h->e[succ].prev = pred;
		int removedKey = h->e[kid].key;
		if (removedKey == h->minKey) {
			if (h->size == 1) h->minKey = -1;
			else
				for (; h->minKey < h->K; h->minKey++)
					if (h->e[h->minKey].next != -1) break;
		}
		h->size--;
	}
	h->e[kid].key = key;
	if (key == -1) { h->e[kid].next = h->e[kid].prev = -1; return; }
	h->size++;
	if (h->minKey == -1) h->minKey = key;
	else h->minKey = imin(h->minKey, key);
	int tmp = h->e[key].next;
	h->e[key].next = kid;
	h->e[kid].prev = key;
	h->e[kid].next = tmp;
	if (tmp != -1) h->e[tmp].prev = kid;
}
static int heap_pop(nkheap* h) {
	const int kid = h->e[h->minKey].next;
	const int ID = kid - h->K;
	int pred = h->e[kid].prev, succ = h->e[kid].next;
	h->e[pred].next = succ;
	if (succ != -1) h->e[succ].prev = pred;
	h->e[kid].key = -1; h->e[kid].prev = -1; h->e[kid].next = -1;
	h->size--;
	if (h->size == 0) h->minKey = -1;
	else
		for (; h->minKey < h->K; h->minKey++)
			if (h->e[h->minKey].next != -1) break;
	return ID;
}

/* genCoarseGrid (multigrid.cpp:520-578) + knActivateCoarseVertices (:508-517) */
static void gen_coarse_grid(og_mg* m, int l) {
	for (long i = 0; i < m->n[l]; i++) m->type[l][i] = vtFree;
	nkheap h;
	heap_init(&h, (int)m->n[l - 1], m->is3d ? 9 : 5);
	for (long v = 0; v < m->n[l - 1]; v++) {
		if (m->type[l - 1][v] != vtInactive) {
			v3i V = vecIdx(m, v, l - 1);
			int fiv = 1 << ((V.x % 2) + (V.y % 2) + (V.z % 2));
			heap_set(&h, (int)v, fiv);
		}
	}
	while (h.size > 0) {
		int v = heap_pop(&h);
		v3i V = vecIdx(m, v, l - 1);
		int vdone = 0;
		for (int iz = V.z / 2; iz <= (V.z + 1) / 2; iz++)
			for (int iy = V.y / 2; iy <= (V.y + 1) / 2; iy++)
				for (int ix = V.x / 2; ix <= (V.x + 1) / 2; ix++) {
					v3i I = mk(ix, iy, iz);
					long i = linIdx(m, I, l);
					if (m->type[l][i] == vtFree) {
						if (vdone) m->type[l][i] = vtRemoved;
						else { m->type[l][i] = vtZero; vdone = 1; }
						v3i lo = mk(imax(0, 2 * I.x - 1), imax(0, 2 * I.y - 1), imax(0, 2 * I.z - 1));
						v3i hi = mk(imin(m->size[l - 1].x - 1, 2 * I.x + 1), imin(m->size[l - 1].y - 1, 2 * I.y + 1),
						            imin(m->size[l - 1].z - 1, 2 * I.z + 1));
						for (int rz = lo.z; rz <= hi.z; rz++)
							for (int ry = lo.y; ry <= hi.y; ry++)
								for (int rx = lo.x; rx <= hi.x; rx++) {
									long r = linIdx(m, mk(rx, ry, rz), l - 1);
									int key = h.e[h.K + r].key;
									if (key > 1) heap_set(&h, (int)r, key - 1);
									else if (key > -1) heap_set(&h, (int)r, -1);
								}
					}
				}
	}
	free(h.e);
	for (long i = 0; i < m->n[l]; i++) {
		if (m->type[l][i] == vtFree) m->type[l][i] = vtRemoved;
		if (m->type[l][i] == vtZero) m->type[l][i] = vtActive;
		if (m->type[l][i] == vtRemoved) m->type[l][i] = vtInactive;
	}
}

/* knGenCoarseGridOperator (multigrid.cpp:580-666) */
static void gen_coarse_op(og_mg* m, int l) {
	const int S = m->S, S0 = m->S0;
	OMP_FOR
	for (long idx = 0; idx < m->n[l]; idx++) {
		if (m->type[l][idx] == vtInactive) continue;
		og_real* A = m->A[l] + idx * S;
		for (int i = 0; i < S; i++) A[i] = 0;
		v3i V = vecIdx(m, idx, l);
		if (l == 1) {
			for (int p = 0; p < m->npaths; p++) {
				const cpath* it = &m->paths[p];
				v3i N = mk(V.x + it->N.x, V.y + it->N.y, V.z + it->N.z);
				if (!inGrid(m, N, l) || m->type[l][linIdx(m, N, l)] == vtInactive) continue;
				v3i U = mk(2 * V.x + it->U.x, 2 * V.y + it->U.y, 2 * V.z + it->U.z);
				if (!inGrid(m, U, l - 1)) continue;
				long u = linIdx(m, U, l - 1);
				if (m->type[l - 1][u] == vtInactive) continue;
				v3i W = mk(2 * V.x + it->W.x, 2 * V.y + it->W.y, 2 * V.z + it->W.z);
				if (!inGrid(m, W, l - 1)) continue;
				long w = linIdx(m, W, l - 1);
				if (m->type[l - 1][w] == vtInactive) continue;
				if (it->inUStencil) A[it->sc] += it->rw * m->A[l - 1][u * S0 + it->sf] * it->iw;
				else A[it->sc] += it->rw * m->A[l - 1][w * S0 + it->sf] * it->iw;
			}
		} else {
			v3i ulo = mk(imax(0, 2 * V.x - 1), imax(0, 2 * V.y - 1), imax(0, 2 * V.z - 1));
			v3i uhi = mk(imin(m->size[l - 1].x - 1, 2 * V.x + 1), imin(m->size[l - 1].y - 1, 2 * V.y + 1),
			             imin(m->size[l - 1].z - 1, 2 * V.z + 1));
			for (int uz = ulo.z; uz <= uhi.z; uz++)
				for (int uy = ulo.y; uy <= uhi.y; uy++)
					for (int ux = ulo.x; ux <= uhi.x; ux++) {
						v3i U = mk(ux, uy, uz);
						long u = linIdx(m, U, l - 1);
						if (m->type[l - 1][u] == vtInactive) continue;
						og_real rw = (og_real)1 / (og_real)(1 << ((U.x % 2) + (U.y % 2) + (U.z % 2)));
						v3i nlo = mk((U.x - 1) / 2, (U.y - 1) / 2, (U.z - 1) / 2);
						v3i nhi = mk(imin(m->size[l].x - 1, (U.x + 2) / 2), imin(m->size[l].y - 1, (U.y + 2) / 2),
						             imin(m->size[l].z - 1, (U.z + 2) / 2));
						for (int nz = nlo.z; nz <= nhi.z; nz++)
							for (int ny = nlo.y; ny <= nhi.y; ny++)
								for (int nx = nlo.x; nx <= nhi.x; nx++) {
									v3i N = mk(nx, ny, nz);
									long nn = linIdx(m, N, l);
									if (m->type[l][nn] == vtInactive) continue;
									v3i SC = mk(N.x - V.x + m->smax.x, N.y - V.y + m->smax.y, N.z - V.z + m->smax.z);
									int sc = SC.x + 3 * SC.y + 9 * SC.z;
									if (sc < S - 1) continue;
									v3i wlo = mk(imax(0, imax(U.x - 1, 2 * N.x - 1)), imax(0, imax(U.y - 1, 2 * N.y - 1)),
									             imax(0, imax(U.z - 1, 2 * N.z - 1)));
									v3i whi = mk(imin(m->size[l - 1].x - 1, imin(U.x + 1, 2 * N.x + 1)),
									             imin(m->size[l - 1].y - 1, imin(U.y + 1, 2 * N.y + 1)),
									             imin(m->size[l - 1].z - 1, imin(U.z + 1, 2 * N.z + 1)));
									for (int wz = wlo.z; wz <= whi.z; wz++)
										for (int wy = wlo.y; wy <= whi.y; wy++)
											for (int wx = wlo.x; wx <= whi.x; wx++) {
												v3i W = mk(wx, wy, wz);
												long w = linIdx(m, W, l - 1);
												if (m->type[l - 1][w] == vtInactive) continue;
												v3i SF = mk(W.x - U.x + m->smax.x, W.y - U.y + m->smax.y, W.z - U.z + m->smax.z);
												int sf = SF.x + 3 * SF.y + 9 * SF.z;
												og_real iw = (og_real)1 / (og_real)(1 << ((W.x % 2) + (W.y % 2) + (W.z % 2)));
												if (sf < S) A[sc - S + 1] += rw * m->A[l - 1][w * S + S - 1 - sf] * iw;
												else A[sc - S + 1] += rw * m->A[l - 1][u * S + sf - S + 1] * iw;
											}
								}
					}
		}
	}
}

/* GridMg::setA (multigrid.cpp:386-415) with knCopyA, knActivateVertices and
 * analyzeStencil (:323-384) */
void og_mg_set_a(og_mg* m, const og_real* pA0, const og_real* pAi, const og_real* pAj,
                 const og_real* pAk) {
	const int S0 = m->S0;
	og_real* A = m->A[0];
	for (long i = 0; i < m->n[0]; i++) {
		A[i * S0 + 0] = pA0[i];
		A[i * S0 + 1] = pAi[i];
		A[i * S0 + 2] = pAj[i];
		if (m->is3d) A[i * S0 + 3] = pAk[i];
	}
	for (long v = 0; v < m->n[0]; v++) {
		m->type[0][v] = vtInactive;
		if (A[v * S0] != (og_real)0) {
			m->type[0][v] = vtActive;
			v3i V = vecIdx(m, v, 0);
			og_real a[7];
			a[0] = A[v * S0 + 0];
			a[1] = A[v * S0 + 1];
			a[2] = A[v * S0 + 2];
			a[3] = m->is3d ? A[v * S0 + 3] : (og_real)0;
			a[4] = V.x != 0 ? A[(v - m->pitch[0].x) * S0 + 1] : (og_real)0;
			a[5] = V.y != 0 ? A[(v - m->pitch[0].y) * S0 + 2] : (og_real)0;
			a[6] = V.z != 0 && m->is3d ? A[(v - m->pitch[0].z) * S0 + 3] : (og_real)0;
			int trivial = a[0] == (og_real)1 && a[1] == 0 && a[2] == 0 && a[3] == 0 && a[4] == 0 &&
			              a[5] == 0 && a[6] == 0;
			if (trivial) {
				m->type[0][v] = vtActiveTrivial;
				A[v * S0] *= m->trivialScale;
			}
		}
	}
	for (int l = 1; l < m->nl; l++) {
		gen_coarse_grid(m, l);
		gen_coarse_op(m, l);
	}
	m->isASet = 1;
}

/* knSmoothColor / smoothGS (multigrid.cpp:668-737) */
static void smooth_gs(og_mg* m, int l, int reversed) {
	static const int a[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
	                            {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
	int colors[8][4], ncol[8], nc;
	if (m->is3d) {
		if (l == 0) {
			nc = 2;
			int c0[4] = {0, 3, 5, 6}, c1[4] = {1, 2, 4, 7};
			memcpy(colors[0], c0, sizeof c0); memcpy(colors[1], c1, sizeof c1);
			ncol[0] = ncol[1] = 4;
		} else {
			nc = 8;
			for (int c = 0; c < 8; c++) { colors[c][0] = c; ncol[c] = 1; }
		}
	} else {
		if (l == 0) {
			nc = 2;
			colors[0][0] = 0; colors[0][1] = 3; colors[1][0] = 1; colors[1][1] = 2;
			ncol[0] = ncol[1] = 2;
		} else {
			nc = 4;
			for (int c = 0; c < 4; c++) { colors[c][0] = c; ncol[c] = 1; }
		}
	}
	v3i bs = mk((m->size[l].x + 1) / 2, (m->size[l].y + 1) / 2, (m->size[l].z + 1) / 2);
	long nblocks = (long)bs.x * bs.y * bs.z;
	const int S = m->S, S0 = m->S0;
	og_real* x = m->x[l];
	const og_real* bb = m->b[l];
	const og_real* A = m->A[l];
	for (int c = 0; c < nc; c++) {
		int color = reversed ? nc - 1 - c : c;
		OMP_FOR
		for (long idx = 0; idx < nblocks; idx++) {
			v3i off = mk((int)(idx % bs.x), (int)((idx % ((long)bs.x * bs.y)) / bs.x), (int)(idx / ((long)bs.x * bs.y)));
			for (int o = 0; o < ncol[color]; o++) {
				const int* co = a[colors[color][o]];
				v3i V = mk(off.x * 2 + co[0], off.y * 2 + co[1], off.z * 2 + co[2]);
				if (!inGrid(m, V, l)) continue;
				long v = linIdx(m, V, l);
				if (m->type[l][v] == vtInactive) continue;
				og_real sum = bb[v];
				if (l == 0) {
					for (int dd = 0; dd < m->dim; dd++) {
						long pd = vget(m->pitch[0], dd);
						if (vget(V, dd) > 0) { long nn = v - pd; sum -= A[nn * S0 + dd + 1] * x[nn]; }
						if (vget(V, dd) < vget(m->size[0], dd) - 1) { long nn = v + pd; sum -= A[v * S0 + dd + 1] * x[nn]; }
					}
					x[v] = sum / A[v * S0 + 0];
				} else {
					int s = 0;
					for (int sz = m->smin.z; sz <= m->smax.z; sz++)
						for (int sy = m->smin.y; sy <= m->smax.y; sy++)
							for (int sx = m->smin.x; sx <= m->smax.x; sx++, s++) {
								if (s == S - 1) continue;
								v3i N = mk(V.x + sx, V.y + sy, V.z + sz);
								if (!inGrid(m, N, l)) continue;
								long nn = linIdx(m, N, l);
								if (m->type[l][nn] == vtInactive) continue;
								if (s < S) sum -= A[nn * S + S - 1 - s] * x[nn];
								else sum -= A[v * S + s - S + 1] * x[nn];
							}
					x[v] = sum / A[v * S + 0];
				}
			}
		}
	}
}

/* knCalcResidual (multigrid.cpp:739-777) */
static void calc_residual(og_mg* m, int l) {
	const int S = m->S, S0 = m->S0;
	const og_real* A = m->A[l];
	const og_real* x = m->x[l];
	OMP_FOR
	for (long idx = 0; idx < m->n[l]; idx++) {
		if (m->type[l][idx] == vtInactive) continue;
		v3i V = vecIdx(m, idx, l);
		og_real sum = m->b[l][idx];
		if (l == 0) {
			for (int dd = 0; dd < m->dim; dd++) {
				long pd = vget(m->pitch[0], dd);
				if (vget(V, dd) > 0) { long nn = idx - pd; sum -= A[nn * S0 + dd + 1] * x[nn]; }
				if (vget(V, dd) < vget(m->size[0], dd) - 1) { long nn = idx + pd; sum -= A[idx * S0 + dd + 1] * x[nn]; }
			}
			sum -= A[idx * S0 + 0] * x[idx];
		} else {
			int s = 0;
			for (int sz = m->smin.z; sz <= m->smax.z; sz++)
				for (int sy = m->smin.y; sy <= m->smax.y; sy++)
					for (int sx = m->smin.x; sx <= m->smax.x; sx++, s++) {
						v3i N = mk(V.x + sx, V.y + sy, V.z + sz);
						if (!inGrid(m, N, l)) continue;
						long nn = linIdx(m, N, l);
						if (m->type[l][nn] == vtInactive) continue;
						if (s < S) sum -= A[nn * S + S - 1 - s] * x[nn];
						else sum -= A[idx * S + s - S + 1] * x[nn];
					}
		}
		m->r[l][idx] = sum;
	}
}

/* coarsest-level stencil in FP64 (solveCG's applyAStencil, multigrid.cpp:800-829) */
static double apply_stencil_d(const og_mg* m, long v, int l, const double* vec) {
	const int S = m->S, S0 = m->S0;
	const og_real* A = m->A[l];
	v3i V = vecIdx(m, v, l);
	double sum = 0;
	if (l == 0) {
		for (int dd = 0; dd < m->dim; dd++) {
			long pd = vget(m->pitch[0], dd);
			if (vget(V, dd) > 0) { long nn = v - pd; sum += A[nn * S0 + dd + 1] * vec[nn]; }
			if (vget(V, dd) < vget(m->size[0], dd) - 1) { long nn = v + pd; sum += A[v * S0 + dd + 1] * vec[nn]; }
		}
		sum += A[v * S0 + 0] * vec[v];
	} else {
		int s = 0;
		for (int sz = m->smin.z; sz <= m->smax.z; sz++)
			for (int sy = m->smin.y; sy <= m->smax.y; sy++)
				for (int sx = m->smin.x; sx <= m->smax.x; sx++, s++) {
					v3i N = mk(V.x + sx, V.y + sy, V.z + sz);
					if (!inGrid(m, N, l)) continue;
					long nn = linIdx(m, N, l);
					if (m->type[l][nn] == vtInactive) continue;
					if (s < S) sum += A[nn * S + S - 1 - s] * vec[nn];
					else sum += A[v * S + s - S + 1] * vec[nn];
				}
	}
	return sum;
}

/* GridMg::solveCG (multigrid.cpp:796-902): FP64 Jacobi-PCG */
static void solve_cg(og_mg* m, int l) {
	double *z = m->cg1, *p = m->cg2, *x = m->cg3, *r = m->cg4;
	const long n = m->n[l];
	const int ss = l == 0 ? m->S0 : m->S;
	double alphaTop = 0, initialResidual = 0;
	for (long v = 0; v < n; v++) x[v] = m->x[l][v];
	for (long v = 0; v < n; v++) {
		if (m->type[l][v] == vtInactive) continue;
		r[v] = m->b[l][v] - apply_stencil_d(m, v, l, x);
		z[v] = r[v] / m->A[l][v * ss + 0];
		initialResidual += r[v] * r[v];
		p[v] = z[v];
		alphaTop += r[v] * z[v];
	}
	initialResidual = sqrt(initialResidual);
	const int maxIter = 10000;
	double residual = -1;
	for (int iter = 0; iter < maxIter && initialResidual > 1E-12; iter++) {
		double alphaBot = 0;
		for (long v = 0; v < n; v++) {
			if (m->type[l][v] == vtInactive) continue;
			z[v] = apply_stencil_d(m, v, l, p);
			alphaBot += p[v] * z[v];
		}
		double alpha = alphaTop / alphaBot;
		double alphaTopNew = 0;
		residual = 0;
		for (long v = 0; v < n; v++) {
			if (m->type[l][v] == vtInactive) continue;
			x[v] += alpha * p[v];
			r[v] -= alpha * z[v];
			residual += r[v] * r[v];
			z[v] = r[v] / m->A[l][v * ss + 0];
			alphaTopNew += r[v] * z[v];
		}
		residual = sqrt(residual);
		if (residual / initialResidual < m->coarsestAcc) break;
		double beta = alphaTopNew / alphaTop;
		alphaTop = alphaTopNew;
		for (long v = 0; v < n; v++) p[v] = z[v] + beta * p[v];
	}
	for (long v = 0; v < n; v++) m->x[l][v] = (og_real)x[v];
}

/* knRestrict (multigrid.cpp:904-929) */
static void restrict_l(og_mg* m, int ld) {
	const int ls = ld - 1;
	const og_real* src = m->r[ls];
	OMP_FOR
	for (long idx = 0; idx < m->n[ld]; idx++) {
		if (m->type[ld][idx] == vtInactive) continue;
		v3i V = vecIdx(m, idx, ld);
		og_real sum = 0;
		v3i lo = mk(imax(0, 2 * V.x - 1), imax(0, 2 * V.y - 1), imax(0, 2 * V.z - 1));
		v3i hi = mk(imin(m->size[ls].x - 1, 2 * V.x + 1), imin(m->size[ls].y - 1, 2 * V.y + 1),
		            imin(m->size[ls].z - 1, 2 * V.z + 1));
		for (int rz = lo.z; rz <= hi.z; rz++)
			for (int ry = lo.y; ry <= hi.y; ry++)
				for (int rx = lo.x; rx <= hi.x; rx++) {
					long r = linIdx(m, mk(rx, ry, rz), ls);
					if (m->type[ls][r] == vtInactive) continue;
					og_real rw = (og_real)1 / (og_real)(1 << ((rx % 2) + (ry % 2) + (rz % 2)));
					sum += rw * src[r];
				}
		m->b[ld][idx] = sum;
	}
}

/* knInterpolate (multigrid.cpp:936-959): dst = r[ld] */
static void interpolate_l(og_mg* m, int ld) {
	const int ls = ld + 1;
	const og_real* src = m->x[ls];
	OMP_FOR
	for (long idx = 0; idx < m->n[ld]; idx++) {
		if (m->type[ld][idx] == vtInactive) continue;
		v3i V = vecIdx(m, idx, ld);
		og_real sum = 0;
		for (int iz = V.z / 2; iz <= (V.z + 1) / 2; iz++)
			for (int iy = V.y / 2; iy <= (V.y + 1) / 2; iy++)
				for (int ix = V.x / 2; ix <= (V.x + 1) / 2; ix++) {
					long i = linIdx(m, mk(ix, iy, iz), ls);
					if (m->type[ls][i] != vtInactive) sum += src[i];
				}
		og_real iw = (og_real)1 / (og_real)(1 << ((V.x % 2) + (V.y % 2) + (V.z % 2)));
		m->r[ld][idx] = iw * sum;
	}
}

/* GridMg::setRhs + doVCycle (multigrid.cpp:425-504) */
og_real og_mg_vcycle(og_mg* m, const og_real* rhs, og_real* dst) {
	for (long i = 0; i < m->n[0]; i++) {
		m->b[0][i] = rhs[i];
		if (m->type[0][i] == vtActiveTrivial) m->b[0][i] *= m->trivialScale;
	}
	const int maxLevel = m->nl - 1;
	memset(m->x[0], 0, sizeof(og_real) * (size_t)m->n[0]);
	for (int l = 0; l < maxLevel; l++) {
		for (int i = 0; i < m->nPre; i++) smooth_gs(m, l, 0);
		calc_residual(m, l);
		restrict_l(m, l + 1);
		memset(m->x[l + 1], 0, sizeof(og_real) * (size_t)m->n[l + 1]);
	}
	solve_cg(m, maxLevel);
	for (int l = maxLevel - 1; l >= 0; l--) {
		interpolate_l(m, l);
		for (long i = 0; i < m->n[l]; i++) m->x[l][i] += m->r[l][i];
		for (int i = 0; i < m->nPost; i++) smooth_gs(m, l, 1);
	}
	calc_residual(m, 0);
	og_real res = 0;
	for (long i = 0; i < m->n[0]; i++)
		if (m->type[0][i] != vtInactive) res += m->r[0][i] * m->r[0][i];
	memcpy(dst, m->x[0], sizeof(og_real) * (size_t)m->n[0]);
	return (og_real)sqrt(res);
}
