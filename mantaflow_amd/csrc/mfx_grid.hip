// Scenes, grids, element-wise grid operations, reductions, flag-grid setup
// and shape rasterisation of libmfx.
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <set>

#include "mfx_internal.h"
#include "mfx_mg.h"

namespace mfx {

static thread_local std::string g_err;
void set_last_error(const std::string& s) { g_err = s; }

#define API_BEGIN try {
#define API_END                                         \
	}                                                   \
	catch (const std::exception& e) {                   \
		::mfx::set_last_error(e.what());                \
		return -1;                                      \
	}                                                   \
	catch (...) {                                       \
		::mfx::set_last_error("unknown error");         \
		return -1;                                      \
	}                                                   \
	return 0;

// ---------------------------------------------------------------------------
// scene
// ---------------------------------------------------------------------------
Scene::~Scene() {
	if (stream) hipStreamSynchronize(stream);
	mgStatic.reset();
	mgMic.reset();
	if (dhash) hipFree(dhash);
	if (hhash) hipHostFree(hhash);
	for (void* p : owned) hipFree(p);
	if (dpart) hipFree(dpart);
	if (fpart) hipFree(fpart);
	if (dscal) hipFree(dscal);
	if (code) hipFree(code);
	if (hscal) hipHostFree(hscal);
	if (ev) hipEventDestroy(ev);
	if (evState) hipEventDestroy(evState);
	if (arena) hipFree(arena);
	if (hstage) hipHostFree(hstage);
	if (bexec) hipGraphExecDestroy(bexec);
	if (stream) hipStreamDestroy(stream);
}

float* Scene::getReal() {
	if (!freeReal.empty()) {
		float* p = freeReal.back();
		freeReal.pop_back();
		return p;
	}
	float* p = nullptr;
	MFX_HIP(hipMalloc(&p, sizeof(float) * (size_t)d.n));
	owned.push_back(p);
	return p;
}
void Scene::putReal(float* p) { freeReal.push_back(p); }
float* Scene::getVec() {
	if (!freeVec.empty()) {
		float* p = freeVec.back();
		freeVec.pop_back();
		return p;
	}
	float* p = nullptr;
	MFX_HIP(hipMalloc(&p, sizeof(float) * 3 * cstride));
	owned.push_back(p);
	return p;
}
void Scene::putVec(float* p) { freeVec.push_back(p); }

void Scene::ensurePartials(size_t nd, size_t nf) {
	if (nd > dpartCap) {
		if (dpart) MFX_HIP(hipFree(dpart));
		MFX_HIP(hipMalloc(&dpart, sizeof(double) * nd));
		dpartCap = nd;
	}
	if (nf > fpartCap) {
		if (fpart) MFX_HIP(hipFree(fpart));
		MFX_HIP(hipMalloc(&fpart, sizeof(float) * nf));
		fpartCap = nf;
	}
}

void sync(Scene* s) { MFX_HIP(hipStreamSynchronize(s->stream)); }

// ---------------------------------------------------------------------------
// layout kernels
// ---------------------------------------------------------------------------
__global__ void k_aos_to_soa(const float* __restrict__ aos, float* __restrict__ soa, long long n,
                             long long cs) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x) {
		soa[i] = aos[3 * i];
		soa[cs + i] = aos[3 * i + 1];
		soa[2 * cs + i] = aos[3 * i + 2];
	}
}
__global__ void k_soa_to_aos(const float* __restrict__ soa, float* __restrict__ aos, long long n,
                             long long cs) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x) {
		aos[3 * i] = soa[i];
		aos[3 * i + 1] = soa[cs + i];
		aos[3 * i + 2] = soa[2 * cs + i];
	}
}
__global__ void k_pack_flags(const int* __restrict__ f, uint8_t* __restrict__ f8, long long n) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x)
		f8[i] = (uint8_t)(f[i] & 0x7f);
}

static std::atomic<uint64_t> g_flag_gen{0};
void flags_touched(Grid* g) { g->gen = ++g_flag_gen; }

void flags_repack(Grid* g) {
	Scene* s = g->s;
	flags_touched(g);
	g->ogen = g->gen;
	hipLaunchKernelGGL(k_pack_flags, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream,
	                   (const int*)g->data, g->f8, s->d.n);
}

Grid* make_grid(Scene* s, int type) {
	Grid* g = new Grid();
	g->s = s;
	g->type = type;
	size_t bytes;
	if (type & kTypeVec3) bytes = sizeof(float) * 3 * s->cstride;
	else bytes = sizeof(float) * (size_t)s->d.n;  // float and int32 are both 4 bytes
	try {
		MFX_HIP(hipMalloc(&g->data, bytes));
		MFX_HIP(hipMemsetAsync(g->data, 0, bytes, s->stream));
		if (type & kTypeFlags) {
			MFX_HIP(hipMalloc(&g->f8, (size_t)s->d.n));
			MFX_HIP(hipMemsetAsync(g->f8, 0, (size_t)s->d.n, s->stream));
		}
	} catch (...) {
		if (g->data) hipFree(g->data);
		delete g;
		throw;
	}
	s->grids.push_back(g);
	return g;
}

void free_grid(Grid* g) {
	Scene* s = g->s;
	hipStreamSynchronize(s->stream);
	if (g->data) hipFree(g->data);
	if (g->f8) hipFree(g->f8);
	if (g->olist) hipFree(g->olist);
	if (g->ocount) hipFree(g->ocount);
	for (size_t i = 0; i < s->grids.size(); i++)
		if (s->grids[i] == g) {
			s->grids.erase(s->grids.begin() + i);
			break;
		}
	delete g;
}

// ---------------------------------------------------------------------------
// element-wise operations
// ---------------------------------------------------------------------------
enum { EW_SET = 0, EW_ADDC, EW_MULC, EW_ADD, EW_SUB, EW_MUL, EW_ADDSCALED, EW_CLAMP, EW_COPY };

template <typename T>
__global__ void k_elementwise(T* __restrict__ a, const T* __restrict__ b, long long n, int op, T v,
                              T v2) {
	// set and copy do not read the old values (b is null for a set)
	if (op == EW_SET) {
		for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
		     i += (long long)gridDim.x * blockDim.x)
			a[i] = v;
		return;
	}
	if (op == EW_COPY) {
		for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
		     i += (long long)gridDim.x * blockDim.x)
			a[i] = b[i];
		return;
	}
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x) {
		T x = a[i];
		switch (op) {
		case EW_ADDC: x += v; break;
		case EW_MULC: x *= v; break;
		case EW_ADD: x += b[i]; break;
		case EW_SUB: x -= b[i]; break;
		case EW_MUL: x *= b[i]; break;
		case EW_ADDSCALED: x += v * b[i]; break;  // gridScaledAdd: me += factor * other
		case EW_CLAMP: x = x < v ? v : (x > v2 ? v2 : x); break;
		}
		a[i] = x;
	}
}

// Four floats per element: the same per-value operation as k_elementwise,
// for float grids whose size is a multiple of 4 and whose planes are 16-byte
// aligned (every pooled grid is).
__device__ inline float ew_op(float x, float bv, int op, float v, float v2) {
	switch (op) {
	case EW_ADDC: return x + v;
	case EW_MULC: return x * v;
	case EW_ADD: return x + bv;
	case EW_SUB: return x - bv;
	case EW_MUL: return x * bv;
	case EW_ADDSCALED: return x + v * bv;
	case EW_CLAMP: return x < v ? v : (x > v2 ? v2 : x);
	}
	return x;
}
__global__ void k_elementwise4(float4* __restrict__ a, const float4* __restrict__ b, long long n4, int op, float v,
                               float v2) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
	     i += (long long)gridDim.x * blockDim.x) {
		if (op == EW_SET) {
			a[i] = make_float4(v, v, v, v);
		} else if (op == EW_COPY) {
			a[i] = b[i];
		} else {
			const float4 x = a[i];
			const float4 y = b ? b[i] : make_float4(0.f, 0.f, 0.f, 0.f);
			a[i] = make_float4(ew_op(x.x, y.x, op, v, v2), ew_op(x.y, y.y, op, v, v2), ew_op(x.z, y.z, op, v, v2),
			                   ew_op(x.w, y.w, op, v, v2));
		}
	}
}
static void launch_elementwise_f(Scene* s, float* a, const float* b, long long n, int op, float v, float v2) {
	if (n % 4 == 0 && ((uintptr_t)a | (uintptr_t)b) % 16 == 0)
		hipLaunchKernelGGL(k_elementwise4, dim3(flat_blocks(n / 4)), dim3(256), 0, s->stream, (float4*)a,
		                   (const float4*)b, n / 4, op, v, v2);
	else
		hipLaunchKernelGGL(k_elementwise<float>, dim3(flat_blocks(n)), dim3(256), 0, s->stream, a, b, n, op, v, v2);
}

// g = g * a, then g = g + b (two roundings, as the two calls), one pass
__global__ void k_mult_add4(float4* __restrict__ g, const float4* __restrict__ a, const float4* __restrict__ b,
                            long long n4) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
	     i += (long long)gridDim.x * blockDim.x) {
		const float4 x = g[i], y = a[i], z = b[i];
		g[i] = make_float4(x.x * y.x + z.x, x.y * y.y + z.y, x.z * y.z + z.z, x.w * y.w + z.w);
	}
}
__global__ void k_mult_add(float* __restrict__ g, const float* __restrict__ a, const float* __restrict__ b,
                           long long n) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
		g[i] = g[i] * a[i] + b[i];
}

static void elementwise(Grid* g, const Grid* other, int op, const float* v, const float* v2) {
	Scene* s = g->s;
	long long n = s->d.n;
	unsigned nb = flat_blocks(n);
	if (other) {
		MFX_CHECK(other->s->d.nx == s->d.nx && other->s->d.ny == s->d.ny && other->s->d.nz == s->d.nz,
		          "different grid resolutions");
		MFX_CHECK((other->type & (kTypeVec3 | kTypeInt | kTypeReal)) == (g->type & (kTypeVec3 | kTypeInt | kTypeReal)),
		          "grid types do not match");
	}
	if (g->isVec()) {
		for (int c = 0; c < 3; c++) {
			float a = v ? v[c] : 0.f, b = v2 ? v2[c] : 0.f;
			launch_elementwise_f(s, g->comp(c), other ? (const float*)other->comp(c) : nullptr, n, op, a, b);
		}
	} else if (g->isInt()) {
		int a = v ? (int)v[0] : 0, b = v2 ? (int)v2[0] : 0;
		hipLaunchKernelGGL(k_elementwise<int>, dim3(nb), dim3(256), 0, s->stream, g->ints(),
		                   other ? (const int*)other->ints() : nullptr, n, op, a, b);
		if (g->isFlags()) flags_repack(g);
	} else {
		float a = v ? v[0] : 0.f, b = v2 ? v2[0] : 0.f;
		launch_elementwise_f(s, g->real(), other ? (const float*)other->real() : nullptr, n, op, a, b);
	}
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// mode 0: min/max of floats (NaN ignored like "if (v < m) m = v") -> fpart[2*b], fpart[2*b+1]
// mode 1: sum |x| (double), mode 2: sum x^2 (double), mode 3: sum x (double)
// bnd restricts to the interior (FOR_IJKT_BND) for modes 1/2
__global__ void k_reduce_real(const float* __restrict__ a, Dims d, int mode, int bnd,
                              double* __restrict__ dp, float* __restrict__ fp) {
	__shared__ double sh[16];
	__shared__ float shf[32];
	float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
	double acc = 0.0;
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < d.n;
	     i += (long long)gridDim.x * blockDim.x) {
		float v = a[i];
		if (mode == 0) {
			if (v < mn) mn = v;
			if (v > mx) mx = v;
		} else {
			if (bnd > 0) {
				long long ii = i % d.nx, jj = (i / d.nx) % d.ny, kk = i / ((long long)d.nx * d.ny);
				if (ii < bnd || ii >= d.nx - bnd || jj < bnd || jj >= d.ny - bnd) continue;
				if (d.is3d && (kk < bnd || kk >= d.nz - bnd)) continue;
			}
			if (mode == 1) acc += (double)fabsf(v);
			else if (mode == 2) acc += (double)v * (double)v;
			else acc += (double)v;
		}
	}
	if (mode == 0) {
		mn = wave_min(mn);
		mx = wave_max(mx);
		int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
		if (lane == 0) {
			shf[w] = mn;
			shf[16 + w] = mx;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
				mn = fminf(mn, shf[q]);
				mx = fmaxf(mx, shf[16 + q]);
			}
			mn = fminf(shf[0], mn);
			mx = fmaxf(shf[16], mx);
			fp[2 * blockIdx.x] = mn;
			fp[2 * blockIdx.x + 1] = mx;
		}
	} else {
		double r = block_sum(acc, sh);
		if (threadIdx.x == 0) dp[blockIdx.x] = r;
	}
}

static void reduce_real_host(Scene* s, const float* a, int mode, int bnd, double* outSum, float* outMin,
                             float* outMax) {
	unsigned nb = flat_blocks(s->d.n, 256, 4096);
	s->ensurePartials(nb, 2 * nb);
	hipLaunchKernelGGL(k_reduce_real, dim3(nb), dim3(256), 0, s->stream, a, s->d, mode, bnd, s->dpart,
	                   s->fpart);
	if (mode == 0) {
		std::vector<float> h(2 * nb);
		MFX_HIP(hipMemcpyAsync(h.data(), s->fpart, sizeof(float) * 2 * nb, hipMemcpyDeviceToHost, s->stream));
		sync(s);
		float mn = std::numeric_limits<float>::max(), mx = -std::numeric_limits<float>::max();
		for (unsigned b = 0; b < nb; b++) {
			if (h[2 * b] < mn) mn = h[2 * b];
			if (h[2 * b + 1] > mx) mx = h[2 * b + 1];
		}
		*outMin = mn;
		*outMax = mx;
	} else {
		std::vector<double> h(nb);
		MFX_HIP(hipMemcpyAsync(h.data(), s->dpart, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
		sync(s);
		double acc = 0;
		for (unsigned b = 0; b < nb; b++) acc += h[b];
		*outSum = acc;
	}
}

double reduce_sum_double(Scene* s, const float* a, int mode, long long n) {
	(void)n;
	double r = 0;
	float mn, mx;
	reduce_real_host(s, a, mode, 0, &r, &mn, &mx);
	return r;
}

// vec3: normSquare per cell (min/max), norm (L1), normSquare (L2)
__global__ void k_vec_norm2(const float* __restrict__ x, const float* __restrict__ y,
                            const float* __restrict__ z, float* __restrict__ out, long long n) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x)
		out[i] = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
}
// norm() with the reference's epsilon handling (vectorbase.h:380-384)
__global__ void k_vec_norm(const float* __restrict__ x, const float* __restrict__ y,
                           const float* __restrict__ z, float* __restrict__ out, long long n) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x) {
		float l = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
		float r;
		if (l <= 1e-6f * 1e-6f) r = 0.f;
		else r = (fabs((double)l - 1.) < 1e-6f * 1e-6f) ? 1.f : sqrtf(l);
		out[i] = r;
	}
}
__global__ void k_int_to_float(const int* __restrict__ a, float* __restrict__ o, long long n) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x)
		o[i] = (float)a[i];
}
__global__ void k_int_minmax(const int* __restrict__ a, long long n, int* __restrict__ out) {
	__shared__ int smn[4], smx[4];
	int mn = 2147483647, mx = -2147483647;
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x) {
		int v = a[i];
		mn = v < mn ? v : mn;
		mx = v > mx ? v : mx;
	}
	for (int o = 32; o > 0; o >>= 1) {
		int a2 = __shfl_xor(mn, o, 64), b2 = __shfl_xor(mx, o, 64);
		mn = a2 < mn ? a2 : mn;
		mx = b2 > mx ? b2 : mx;
	}
	if ((threadIdx.x & 63) == 0) {
		smn[threadIdx.x >> 6] = mn;
		smx[threadIdx.x >> 6] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int q = 1; q < 4; q++) {
			mn = smn[q] < mn ? smn[q] : mn;
			mx = smx[q] > mx ? smx[q] : mx;
		}
		mn = smn[0] < mn ? smn[0] : mn;
		mx = smx[0] > mx ? smx[0] : mx;
		out[2 * blockIdx.x] = mn;
		out[2 * blockIdx.x + 1] = mx;
	}
}

static double grid_reduce(const Grid* g, int op, int bnd) {
	Scene* s = g->s;
	if (g->isInt()) {
		if (op == MFX_RED_L1 || op == MFX_RED_L2 || op == MFX_RED_SUM) {
			float* t = s->getReal();
			hipLaunchKernelGGL(k_int_to_float, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream,
			                   (const int*)g->data, t, s->d.n);
			double r = 0;
			float a, b;
			reduce_real_host(s, t, op == MFX_RED_L1 ? 1 : (op == MFX_RED_L2 ? 2 : 3), bnd, &r, &a, &b);
			s->putReal(t);
			return op == MFX_RED_L2 ? (double)(float)std::sqrt(r) : (double)(float)r;
		}
		unsigned nb = flat_blocks(s->d.n, 256, 4096);
		int* buf;
		MFX_HIP(hipMalloc(&buf, sizeof(int) * 2 * nb));
		hipLaunchKernelGGL(k_int_minmax, dim3(nb), dim3(256), 0, s->stream, (const int*)g->data, s->d.n, buf);
		std::vector<int> h(2 * nb);
		MFX_HIP(hipMemcpyAsync(h.data(), buf, sizeof(int) * 2 * nb, hipMemcpyDeviceToHost, s->stream));
		sync(s);
		hipFree(buf);
		int mn = 2147483647, mx = -2147483647;
		for (unsigned b = 0; b < nb; b++) {
			mn = h[2 * b] < mn ? h[2 * b] : mn;
			mx = h[2 * b + 1] > mx ? h[2 * b + 1] : mx;
		}
		if (op == MFX_RED_MAX) return (float)mx;
		if (op == MFX_RED_MIN) return (float)mn;
		return std::max(std::fabs((float)mn), std::fabs((float)mx));
	}
	if (g->isVec()) {
		float* t = s->getReal();
		unsigned nb = flat_blocks(s->d.n);
		double r = 0;
		float mn, mx;
		if (op == MFX_RED_L1) {
			hipLaunchKernelGGL(k_vec_norm, dim3(nb), dim3(256), 0, s->stream, g->comp(0), g->comp(1),
			                   g->comp(2), t, s->d.n);
			reduce_real_host(s, t, 1, bnd, &r, &mn, &mx);
			s->putReal(t);
			return (float)r;
		}
		hipLaunchKernelGGL(k_vec_norm2, dim3(nb), dim3(256), 0, s->stream, g->comp(0), g->comp(1),
		                   g->comp(2), t, s->d.n);
		if (op == MFX_RED_L2 || op == MFX_RED_SUM) {
			reduce_real_host(s, t, op == MFX_RED_L2 ? 3 : 3, bnd, &r, &mn, &mx);
			s->putReal(t);
			return op == MFX_RED_L2 ? (double)(float)std::sqrt(r) : r;
		}
		reduce_real_host(s, t, 0, 0, &r, &mn, &mx);
		s->putReal(t);
		if (op == MFX_RED_MIN) return std::sqrt(mn);
		return std::sqrt(mx);  // getMax and getMaxAbs: sqrt(max normSquare)
	}
	double r = 0;
	float mn, mx;
	switch (op) {
	case MFX_RED_MAX: reduce_real_host(s, g->real(), 0, 0, &r, &mn, &mx); return mx;
	case MFX_RED_MIN: reduce_real_host(s, g->real(), 0, 0, &r, &mn, &mx); return mn;
	case MFX_RED_MAXABS:
		reduce_real_host(s, g->real(), 0, 0, &r, &mn, &mx);
		return std::max(std::fabs(mn), std::fabs(mx));
	case MFX_RED_L1: reduce_real_host(s, g->real(), 1, bnd, &r, &mn, &mx); return (float)r;
	case MFX_RED_L2: reduce_real_host(s, g->real(), 2, bnd, &r, &mn, &mx); return (float)std::sqrt(r);
	case MFX_RED_SUM: reduce_real_host(s, g->real(), 3, bnd, &r, &mn, &mx); return r;
	}
	throw Error("unknown reduction");
}

// ---------------------------------------------------------------------------
// boundary setters (grid.cpp:585-648)
// ---------------------------------------------------------------------------
__global__ void k_set_bound(float* a, float* b, float* c, Dims d, float v0, float v1, float v2, int w,
                            int ncomp, int mode, int normalOnly) {
	for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < d.n;
	     idx += (long long)gridDim.x * blockDim.x) {
		int i = (int)(idx % d.nx), j = (int)((idx / d.nx) % d.ny), k = (int)(idx / ((long long)d.nx * d.ny));
		if (mode == 0) {  // knSetBoundary
			bool bnd = (i <= w || i >= d.nx - 1 - w || j <= w || j >= d.ny - 1 - w ||
			            (d.is3d && (k <= w || k >= d.nz - 1 - w)));
			if (bnd) {
				a[idx] = v0;
				if (ncomp == 3) {
					b[idx] = v1;
					c[idx] = v2;
				}
			}
		} else if (mode == 1) {  // knSetBoundaryNeumann
			bool set = false;
			int si = i, sj = j, sk = k;
			if (i <= w) { si = w + 1; set = true; }
			if (i >= d.nx - 1 - w) { si = d.nx - 1 - w - 1; set = true; }
			if (j <= w) { sj = w + 1; set = true; }
			if (j >= d.ny - 1 - w) { sj = d.ny - 1 - w - 1; set = true; }
			if (d.is3d) {
				if (k <= w) { sk = w + 1; set = true; }
				if (k >= d.nz - 1 - w) { sk = d.nz - 1 - w - 1; set = true; }
			}
			(void)set;
		} else {  // knSetBoundaryMAC / knSetBoundaryMACNorm
			if (!normalOnly) {
				if (i <= w || i >= d.nx - w || j <= w - 1 || j >= d.ny - 1 - w ||
				    (d.is3d && (k <= w - 1 || k >= d.nz - 1 - w)))
					a[idx] = v0;
				if (i <= w - 1 || i >= d.nx - 1 - w || j <= w || j >= d.ny - w ||
				    (d.is3d && (k <= w - 1 || k >= d.nz - 1 - w)))
					b[idx] = v1;
				if (i <= w - 1 || i >= d.nx - 1 - w || j <= w - 1 || j >= d.ny - 1 - w ||
				    (d.is3d && (k <= w || k >= d.nz - w)))
					c[idx] = v2;
			} else {
				if (i <= w || i >= d.nx - w) a[idx] = v0;
				if (j <= w || j >= d.ny - w) b[idx] = v1;
				if (d.is3d && (k <= w || k >= d.nz - w)) c[idx] = v2;
			}
		}
	}
}

// Neumann copy: reads interior, writes boundary (disjoint sets, no race)
__global__ void k_set_bound_neumann(float* a, Dims d, int w) {
	for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < d.n;
	     idx += (long long)gridDim.x * blockDim.x) {
		int i = (int)(idx % d.nx), j = (int)((idx / d.nx) % d.ny), k = (int)(idx / ((long long)d.nx * d.ny));
		bool set = false;
		int si = i, sj = j, sk = k;
		if (i <= w) { si = w + 1; set = true; }
		if (i >= d.nx - 1 - w) { si = d.nx - 1 - w - 1; set = true; }
		if (j <= w) { sj = w + 1; set = true; }
		if (j >= d.ny - 1 - w) { sj = d.ny - 1 - w - 1; set = true; }
		if (d.is3d) {
			if (k <= w) { sk = w + 1; set = true; }
			if (k >= d.nz - 1 - w) { sk = d.nz - 1 - w - 1; set = true; }
		}
		if (set) a[idx] = a[d.idx(si, sj, sk)];
	}
}

// ---------------------------------------------------------------------------
// flags
// ---------------------------------------------------------------------------
__global__ void k_init_boundaries(int* f, Dims d, int w, int t0, int t1, int t2, int t3, int t4, int t5) {
	for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < d.n;
	     idx += (long long)gridDim.x * blockDim.x) {
		int i = (int)(idx % d.nx), j = (int)((idx / d.nx) % d.ny), k = (int)(idx / ((long long)d.nx * d.ny));
		int v = kEmpty;  // setConst(TypeEmpty)
		if (i <= w) v = t0;
		if (i >= d.nx - 1 - w) v = t1;
		if (j <= w) v = t2;
		if (j >= d.ny - 1 - w) v = t3;
		if (d.is3d) {
			if (k <= w) v = t4;
			if (k >= d.nz - 1 - w) v = t5;
		}
		f[idx] = v;
	}
}

__global__ void k_fill_grid(int* f, long long n, int type) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x) {
		int v = f[i];
		if ((v & kObstacle) == 0 && (v & kInflow) == 0 && (v & kOutflow) == 0 && (v & kOpen) == 0)
			f[i] = (v & ~(kEmpty | kFluid)) | type;
	}
}

__global__ void k_set_open_bound(int* f, Dims d, int bw, int lox, int loy, int loz, int upx, int upy,
                                 int upz, int type) {
	for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < d.n;
	     idx += (long long)gridDim.x * blockDim.x) {
		int i = (int)(idx % d.nx), j = (int)((idx / d.nx) % d.ny), k = (int)(idx / ((long long)d.nx * d.ny));
		bool loX = lox && i <= bw, loY = loy && j <= bw;
		bool upX = upx && i >= d.nx - bw - 1, upY = upy && j >= d.ny - bw - 1;
		bool innerI = i > bw && i < d.nx - bw - 1, innerJ = j > bw && j < d.ny - bw - 1;
		bool obs = f[idx] & kObstacle;
		if ((!d.is3d) && (loX || upX || loY || upY)) {
			if ((loX || upX || innerI) && (loY || upY || innerJ) && obs) f[idx] = type;
		} else {
			bool loZ = loz && k <= bw, upZ = upz && k >= d.nz - bw - 1;
			bool innerK = k > bw && k < d.nz - bw - 1;
			if (loX || upX || loY || upY || loZ || upZ) {
				if ((loX || upX || innerI) && (loY || upY || innerJ) && (loZ || upZ || innerK) && obs)
					f[idx] = type;
			}
		}
	}
}

// resetOutflow over the outflow cells listed by outflow_scan (the uint8
// mirror f8 holds the outflow bit, so the scan reads 1 B per cell and the
// int32 flags are touched only at outflow cells)
__global__ void k_reset_outflow(int* f, uint8_t* f8, float* phi, float* real, const unsigned* __restrict__ list,
                                const unsigned* __restrict__ count) {
	const unsigned m = *count;
	for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x) {
		const unsigned i = list[q];
		const int v = (f[i] | kEmpty) & ~kFluid;
		f[i] = v;
		f8[i] = (uint8_t)(v & 0x7f);
		if (phi) phi[i] = 0.5f;
		if (real) real[i] = 0.f;
	}
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

// 16 flag bytes per thread (one 16-byte load where the run is whole); a
// thread with outflow cells reserves its slots with one atomic
__global__ void k_outflow_scan(const uint8_t* __restrict__ f8, long long n, unsigned* __restrict__ list,
                               unsigned* __restrict__ count) {
	for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t * 16 < n;
	     t += (long long)gridDim.x * blockDim.x) {
		const long long b0 = t * 16;
		uint8_t fl[16];
		if (b0 + 16 <= n) {
			const uint4 w = *(const uint4*)(f8 + b0);
			const unsigned ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
			for (int e = 0; e < 16; e++) fl[e] = (uint8_t)(ws[e >> 2] >> (8 * (e & 3)));
		} else {
#pragma unroll
			for (int e = 0; e < 16; e++) fl[e] = b0 + e < n ? f8[b0 + e] : (uint8_t)0;
		}
		unsigned m = 0;
#pragma unroll
		for (int e = 0; e < 16; e++) m += (fl[e] & kOutflow) ? 1u : 0u;
		if (m == 0) continue;
		unsigned at = atomicAdd(count, m);
#pragma unroll
		for (int e = 0; e < 16; e++)
			if (fl[e] & kOutflow) list[at++] = (unsigned)(b0 + e);
	}
}

const unsigned* outflow_list(Grid* g, const unsigned** count) {
	Scene* s = g->s;
	MFX_CHECK(g->isFlags() && g->f8, "outflow cells: needs a FlagGrid");
	MFX_CHECK(s->d.n < (1LL << 32), "grid too large for the outflow cell list");
	if (!g->olist) {
		MFX_HIP(hipMalloc((void**)&g->olist, sizeof(unsigned) * (size_t)s->d.n));
		MFX_HIP(hipMalloc((void**)&g->ocount, 256));
		g->olistGen = 0;
	}
	if (g->olistGen != g->ogen) {
		hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s->stream, g->ocount);
		const long long runs = (s->d.n + 15) / 16;
		hipLaunchKernelGGL(k_outflow_scan, dim3(flat_blocks(runs)), dim3(256), 0, s->stream, (const uint8_t*)g->f8,
		                   s->d.n, g->olist, g->ocount);
		g->olistGen = g->ogen;
	}
	*count = g->ocount;
	return g->olist;
}

__global__ void k_count_cells(const int* f, Dims d, int flag, int bnd, float* mask,
                              double* __restrict__ dp) {
	__shared__ double sh[16];
	double cnt = 0;
	for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < d.n;
	     idx += (long long)gridDim.x * blockDim.x) {
		int i = (int)(idx % d.nx), j = (int)((idx / d.nx) % d.ny), k = (int)(idx / ((long long)d.nx * d.ny));
		if (mask) mask[idx] = 0.f;
		if (bnd > 0) {  // isInBounds(p, bnd)
			bool in = (i >= bnd && j >= bnd && i < d.nx - bnd && j < d.ny - bnd);
			if (d.is3d) in = in && (k >= bnd && k < d.nz - bnd);
			else in = in && (k == 0);
			if (!in) continue;
		}
		if (f[idx] & flag) {
			cnt += 1;
			if (mask) mask[idx] = 1.f;
		}
	}
	double r = block_sum(cnt, sh);
	if (threadIdx.x == 0) dp[blockIdx.x] = r;
}

// ---------------------------------------------------------------------------
// shapes: isInside of Box / Sphere / Cylinder (shapes.cpp:151-154,240,324-329)
// ---------------------------------------------------------------------------
struct ShapeDesc {
	int kind;
	float a0, a1, a2, b0, b1, b2, r, zlen;
};
__device__ inline bool shape_inside(const ShapeDesc& s, float x, float y, float z) {
	if (s.kind == MFX_SHAPE_BOX) {
		return (x >= s.a0 && y >= s.a1 && z >= s.a2 && x <= s.b0 && y <= s.b1 && z <= s.b2);
	} else if (s.kind == MFX_SHAPE_SPHERE) {
		float dx = (x - s.a0) / s.b0, dy = (y - s.a1) / s.b1, dz = (z - s.a2) / s.b2;
		return dx * dx + dy * dy + dz * dz <= s.r * s.r;
	} else {
		float px = x - s.a0, py = y - s.a1, pz = z - s.a2;
		float zz = px * s.b0 + py * s.b1 + pz * s.b2;
		if (fabsf(zz) > s.zlen) return false;
		float r2 = (px * px + py * py + pz * pz) - zz * zz;
		return r2 < s.r * s.r;
	}
}

__global__ void k_apply_shape(ShapeDesc sh, Dims d, int mode, float* g0, float* g1, float* g2, int* gi,
                              float v0, float v1, float v2, const uint8_t* respect) {
	for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < d.n;
	     idx += (long long)gridDim.x * blockDim.x) {
		int i = (int)(idx % d.nx), j = (int)((idx / d.nx) % d.ny), k = (int)(idx / ((long long)d.nx * d.ny));
		if (respect && (respect[idx] & kObstacle)) continue;
		// cell centre (i+0.5 etc. evaluated in double, exact in float)
		float ci = (float)((double)i + 0.5), cj = (float)((double)j + 0.5), ck = (float)((double)k + 0.5);
		if (mode == 3) {  // MAC: face positions (ApplyShapeToMACGrid)
			if (shape_inside(sh, (float)i, cj, ck)) g0[idx] = v0;
			if (shape_inside(sh, ci, (float)j, ck)) g1[idx] = v1;
			if (shape_inside(sh, ci, cj, (float)k)) g2[idx] = v2;
			continue;
		}
		if (!shape_inside(sh, ci, cj, ck)) continue;
		if (mode == 0) gi[idx] = (int)v0;
		else if (mode == 1) g0[idx] = v0;
		else {
			g0[idx] = v0;
			g1[idx] = v1;
			g2[idx] = v2;
		}
	}
}

// component copies
__global__ void k_copy(float* __restrict__ dst, const float* __restrict__ src, long long n) {
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x)
		dst[i] = src[i];
}

// max diff kernels (double accumulation of |a-b|)
__global__ void k_maxdiff(const float* a0, const float* a1, const float* a2, const float* b0,
                          const float* b1, const float* b2, const int* ia, const int* ib, long long n,
                          int mode, double* __restrict__ dp) {
	__shared__ double sh[4];
	double m = 0;
	for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
	     i += (long long)gridDim.x * blockDim.x) {
		double v;
		if (mode == 0) v = (double)fabsf(a0[i] - b0[i]);
		else if (mode == 1) v = fabs((double)ia[i] - (double)ib[i]);
		else
			v = fabs((double)a0[i] - (double)b0[i]) + fabs((double)a1[i] - (double)b1[i]) +
			    fabs((double)a2[i] - (double)b2[i]);
		m = v > m ? v : m;
	}
	for (int o = 32; o > 0; o >>= 1) {
		double t = __shfl_xor(m, o, 64);
		m = t > m ? t : m;
	}
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int q = 1; q < 4; q++) m = sh[q] > m ? sh[q] : m;
		m = sh[0] > m ? sh[0] : m;
		dp[blockIdx.x] = m;
	}
}

static std::mutex g_scenes_mu;
static std::set<Scene*> g_scenes;
std::vector<Scene*> live_scenes() {
	std::lock_guard<std::mutex> g(g_scenes_mu);
	return std::vector<Scene*>(g_scenes.begin(), g_scenes.end());
}

}  // namespace mfx

using namespace mfx;

struct mfx_scene : public Scene {};
struct mfx_grid : public Grid {};

static inline Scene* S(mfx_scene* s) { return static_cast<Scene*>(s); }
static inline Grid* G(const mfx_grid* g) { return const_cast<Grid*>(static_cast<const Grid*>(g)); }

#define NEED(p, what) MFX_CHECK((p) != nullptr, std::string(what) + " must not be null")

extern "C" {

const char* mfx_last_error(void) { return g_err.c_str(); }

int mfx_device_count(int* out) {
	API_BEGIN
	int c = 0;
	if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
	*out = c;
	API_END
}

int mfx_device_memory(size_t* freeBytes, size_t* totalBytes) {
	API_BEGIN
	size_t f = 0, t = 0;
	MFX_HIP(hipMemGetInfo(&f, &t));
	if (freeBytes) *freeBytes = f;
	if (totalBytes) *totalBytes = t;
	API_END
}

int mfx_scene_create(int nx, int ny, int nz, int dim, mfx_scene** out) {
	API_BEGIN
	MFX_CHECK(dim == 2 || dim == 3, "Only 2D and 3D solvers allowed.");
	MFX_CHECK(dim != 2 || nz == 1, "Trying to create 2D solver with size.z != 1");
	MFX_CHECK(nx >= 3 && ny >= 3 && nz >= 1, "grid too small");
	MFX_CHECK((long long)nx * ny * nz < (1LL << 31), "grid too large (2^31 cells max)");
	Scene* s = new mfx_scene();
	s->d.nx = nx;
	s->d.ny = ny;
	s->d.nz = nz;
	s->d.is3d = dim == 3 ? 1 : 0;
	s->dim = dim;
	s->d.n = (long long)nx * ny * nz;
	s->d.sz = dim == 3 ? (long long)nx * ny : 0;
	s->cstride = (size_t)((s->d.n + 63) / 64 * 64);
	try {
		MFX_HIP(hipGetDevice(&s->device));
		MFX_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
		MFX_HIP(hipMalloc(&s->dscal, 4096));
		MFX_HIP(hipHostMalloc(&s->hscal, 4096, hipHostMallocDefault));
	} catch (...) {
		delete s;
		throw;
	}
	{
		std::lock_guard<std::mutex> g(g_scenes_mu);
		g_scenes.insert(s);
	}
	*out = static_cast<mfx_scene*>(s);
	API_END
}

int mfx_scene_destroy(mfx_scene* s) {
	API_BEGIN
	if (!s) return 0;
	Scene* sc = S(s);
	{
		std::lock_guard<std::mutex> g(g_scenes_mu);
		g_scenes.erase(sc);
	}
	while (!sc->grids.empty()) free_grid(sc->grids.back());
	delete static_cast<mfx_scene*>(s);
	API_END
}

int mfx_scene_set_dt(mfx_scene* s, float dt) {
	API_BEGIN
	NEED(s, "scene");
	S(s)->dt = dt;
	API_END
}
int mfx_scene_get_dt(const mfx_scene* s, float* dt) {
	API_BEGIN
	NEED(s, "scene");
	*dt = static_cast<const Scene*>(s)->dt;
	API_END
}
int mfx_scene_sync(mfx_scene* s) {
	API_BEGIN
	NEED(s, "scene");
	sync(S(s));
	API_END
}
int mfx_scene_memory(const mfx_scene* s, size_t* bytes) {
	API_BEGIN
	NEED(s, "scene");
	const Scene* sc = static_cast<const Scene*>(s);
	size_t b = 0;
	for (Grid* g : sc->grids) b += (g->isVec() ? 3 * sc->cstride * 4 : (size_t)sc->d.n * 4) + (g->f8 ? sc->d.n : 0);
	b += (sc->freeReal.size()) * (size_t)sc->d.n * 4;
	b += (sc->freeVec.size()) * 3 * sc->cstride * 4;
	*bytes = b;
	API_END
}

int mfx_grid_create(mfx_scene* s, int type, mfx_grid** out) {
	API_BEGIN
	NEED(s, "scene");
	int t = type;
	MFX_CHECK(t == MFX_GRID_REAL || t == MFX_GRID_INT || t == MFX_GRID_VEC3 || t == MFX_GRID_MAC ||
	              t == MFX_GRID_LEVELSET || t == MFX_GRID_FLAGS,
	          "Grid Type is not supported (only Real, Int, Vec3, MAC, Levelset, Flags)");
	*out = static_cast<mfx_grid*>(make_grid(S(s), t));
	API_END
}

int mfx_grid_destroy(mfx_grid* g) {
	API_BEGIN
	if (g) free_grid(G(g));
	API_END
}

int mfx_grid_type(const mfx_grid* g, int* type) {
	API_BEGIN
	NEED(g, "grid");
	*type = G(g)->type;
	API_END
}

int mfx_grid_upload(mfx_grid* gg, const void* host, size_t bytes) {
	API_BEGIN
	NEED(gg, "grid");
	NEED(host, "host buffer");
	Grid* g = G(gg);
	Scene* s = g->s;
	size_t need = (size_t)s->d.n * 4 * (g->isVec() ? 3 : 1);
	MFX_CHECK(bytes == need, "upload size mismatch: expected " + std::to_string(need) + " bytes, got " +
	                             std::to_string(bytes));
	if (g->isVec()) {
		float* stage = s->getVec();
		MFX_HIP(hipMemcpyAsync(stage, host, bytes, hipMemcpyHostToDevice, s->stream));
		hipLaunchKernelGGL(k_aos_to_soa, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, stage,
		                   (float*)g->data, s->d.n, (long long)s->cstride);
		sync(s);
		s->putVec(stage);
	} else {
		MFX_HIP(hipMemcpyAsync(g->data, host, bytes, hipMemcpyHostToDevice, s->stream));
		if (g->isFlags()) flags_repack(g);
		sync(s);
	}
	API_END
}

int mfx_grid_download(const mfx_grid* gg, void* host, size_t bytes) {
	API_BEGIN
	NEED(gg, "grid");
	NEED(host, "host buffer");
	Grid* g = G(gg);
	Scene* s = g->s;
	size_t need = (size_t)s->d.n * 4 * (g->isVec() ? 3 : 1);
	MFX_CHECK(bytes == need, "download size mismatch: expected " + std::to_string(need) + " bytes, got " +
	                             std::to_string(bytes));
	if (g->isVec()) {
		float* stage = s->getVec();
		hipLaunchKernelGGL(k_soa_to_aos, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream,
		                   (const float*)g->data, stage, s->d.n, (long long)s->cstride);
		MFX_HIP(hipMemcpyAsync(host, stage, bytes, hipMemcpyDeviceToHost, s->stream));
		sync(s);
		s->putVec(stage);
	} else {
		MFX_HIP(hipMemcpyAsync(host, g->data, bytes, hipMemcpyDeviceToHost, s->stream));
		sync(s);
	}
	API_END
}

int mfx_grid_clear(mfx_grid* gg) {
	API_BEGIN
	NEED(gg, "grid");
	float z[3] = {0, 0, 0};
	elementwise(G(gg), nullptr, EW_SET, z, nullptr);
	API_END
}

int mfx_grid_copy(mfx_grid* dst, const mfx_grid* src) {
	API_BEGIN
	NEED(dst, "grid");
	NEED(src, "grid");
	Grid *a = G(dst), *b = G(src);
	MFX_CHECK(a->s->d.nx == b->s->d.nx && a->s->d.ny == b->s->d.ny && a->s->d.nz == b->s->d.nz,
	          "different grid resolutions");
	MFX_CHECK(a->isVec() == b->isVec(), "copyFrom: grid types do not match");
	Scene* s = a->s;
	if (a->s != b->s) MFX_HIP(hipStreamSynchronize(b->s->stream));
	size_t bytes = a->isVec() ? 3 * s->cstride * 4 : (size_t)s->d.n * 4;
	MFX_HIP(hipMemcpyAsync(a->data, b->data, bytes, hipMemcpyDeviceToDevice, s->stream));
	if (a->isFlags()) flags_repack(a);
	API_END
}

int mfx_grid_swap(mfx_grid* x, mfx_grid* y) {
	API_BEGIN
	NEED(x, "grid");
	NEED(y, "grid");
	Grid *a = G(x), *b = G(y);
	MFX_CHECK(a->s->d.nx == b->s->d.nx && a->s->d.ny == b->s->d.ny && a->s->d.nz == b->s->d.nz,
	          "Grid::swap(): Grid dimensions mismatch.");
	MFX_CHECK(!a->external && !b->external, "Grid::swap(): Cannot swap if one grid stores externalData.");
	MFX_CHECK(a->isVec() == b->isVec() && a->isFlags() == b->isFlags(), "Grid::swap(): grid types differ.");
	std::swap(a->data, b->data);
	std::swap(a->f8, b->f8);
	if (a->isFlags()) {
		flags_touched(a);
		flags_touched(b);
		// the outflow cell lists follow the data they were built from
		std::swap(a->olist, b->olist);
		std::swap(a->ocount, b->ocount);
		std::swap(a->olistGen, b->olistGen);
		std::swap(a->ogen, b->ogen);
	}
	API_END
}

int mfx_grid_set_const(mfx_grid* g, const float* v) {
	API_BEGIN
	NEED(g, "grid");
	elementwise(G(g), nullptr, EW_SET, v, nullptr);
	API_END
}
int mfx_grid_add_const(mfx_grid* g, const float* v) {
	API_BEGIN
	NEED(g, "grid");
	elementwise(G(g), nullptr, EW_ADDC, v, nullptr);
	API_END
}
int mfx_grid_mult_const(mfx_grid* g, const float* v) {
	API_BEGIN
	NEED(g, "grid");
	elementwise(G(g), nullptr, EW_MULC, v, nullptr);
	API_END
}
int mfx_grid_add(mfx_grid* g, const mfx_grid* a) {
	API_BEGIN
	NEED(g, "grid");
	NEED(a, "grid");
	elementwise(G(g), G(a), EW_ADD, nullptr, nullptr);
	API_END
}
int mfx_grid_sub(mfx_grid* g, const mfx_grid* a) {
	API_BEGIN
	NEED(g, "grid");
	NEED(a, "grid");
	elementwise(G(g), G(a), EW_SUB, nullptr, nullptr);
	API_END
}
int mfx_grid_mult(mfx_grid* g, const mfx_grid* a) {
	API_BEGIN
	NEED(g, "grid");
	NEED(a, "grid");
	elementwise(G(g), G(a), EW_MUL, nullptr, nullptr);
	API_END
}
int mfx_grid_mult_add(mfx_grid* g, const mfx_grid* a, const mfx_grid* b) {
	API_BEGIN
	NEED(g, "grid");
	NEED(a, "grid");
	NEED(b, "grid");
	Grid *x = G(g), *y = G(a), *z = G(b);
	MFX_CHECK(x->isReal() && y->isReal() && z->isReal(), "mult+add: needs Real grids");
	MFX_CHECK(z != x, "mult+add: the added grid must not be the result grid");
	for (const Grid* o : {y, z})
		MFX_CHECK(o->s->d.nx == x->s->d.nx && o->s->d.ny == x->s->d.ny && o->s->d.nz == x->s->d.nz,
		          "different grid resolutions");
	Scene* s = x->s;
	const long long n = s->d.n;
	float* gp = x->real();
	const float *ap = y->real(), *bp = z->real();
	if (n % 4 == 0 && ((uintptr_t)gp | (uintptr_t)ap | (uintptr_t)bp) % 16 == 0)
		hipLaunchKernelGGL(k_mult_add4, dim3(flat_blocks(n / 4)), dim3(256), 0, s->stream, (float4*)gp,
		                   (const float4*)ap, (const float4*)bp, n / 4);
	else
		hipLaunchKernelGGL(k_mult_add, dim3(flat_blocks(n)), dim3(256), 0, s->stream, gp, ap, bp, n);
	API_END
}
int mfx_grid_add_scaled(mfx_grid* g, const mfx_grid* a, const float* factor) {
	API_BEGIN
	NEED(g, "grid");
	NEED(a, "grid");
	elementwise(G(g), G(a), EW_ADDSCALED, factor, nullptr);
	API_END
}
int mfx_grid_clamp(mfx_grid* g, float vmin, float vmax) {
	API_BEGIN
	NEED(g, "grid");
	float a[3] = {vmin, vmin, vmin}, b[3] = {vmax, vmax, vmax};
	elementwise(G(g), nullptr, EW_CLAMP, a, b);
	API_END
}

int mfx_grid_reduce(const mfx_grid* g, int op, int bnd, double* out) {
	API_BEGIN
	NEED(g, "grid");
	*out = grid_reduce(G(g), op, bnd);
	API_END
}

int mfx_grid_set_bound(mfx_grid* gg, const float* value, int w) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	Scene* s = g->s;
	MFX_CHECK(!g->isInt(), "setBound: int grids not supported");
	int nc = g->isVec() ? 3 : 1;
	hipLaunchKernelGGL(k_set_bound, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, g->comp(0),
	                   nc == 3 ? g->comp(1) : nullptr, nc == 3 ? g->comp(2) : nullptr, s->d, value[0],
	                   nc == 3 ? value[1] : 0.f, nc == 3 ? value[2] : 0.f, w, nc, 0, 0);
	API_END
}
int mfx_grid_set_bound_neumann(mfx_grid* gg, int w) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	Scene* s = g->s;
	MFX_CHECK(!g->isInt(), "setBoundNeumann: int grids not supported");
	int nc = g->isVec() ? 3 : 1;
	for (int c = 0; c < nc; c++)
		hipLaunchKernelGGL(k_set_bound_neumann, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream,
		                   g->comp(c), s->d, w);
	API_END
}
int mfx_grid_set_bound_mac(mfx_grid* gg, const float* value, int w, int normalOnly) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	Scene* s = g->s;
	MFX_CHECK(g->isVec(), "setBoundMAC needs a MAC grid");
	hipLaunchKernelGGL(k_set_bound, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, g->comp(0),
	                   g->comp(1), g->comp(2), s->d, value[0], value[1], value[2], w, 3, 2, normalOnly);
	API_END
}

int mfx_grid_max_diff(const mfx_grid* a, const mfx_grid* b, double* out) {
	API_BEGIN
	NEED(a, "grid");
	NEED(b, "grid");
	Grid *x = G(a), *y = G(b);
	Scene* s = x->s;
	MFX_CHECK(x->isVec() == y->isVec() && x->isInt() == y->isInt(), "gridMaxDiff: grid types differ");
	if (x->s != y->s) MFX_HIP(hipStreamSynchronize(y->s->stream));
	unsigned nb = flat_blocks(s->d.n, 256, 4096);
	s->ensurePartials(nb, 0);
	int mode = x->isInt() ? 1 : (x->isVec() ? 2 : 0);
	hipLaunchKernelGGL(k_maxdiff, dim3(nb), dim3(256), 0, s->stream, x->comp(0), x->isVec() ? x->comp(1) : nullptr,
	                   x->isVec() ? x->comp(2) : nullptr, y->comp(0), y->isVec() ? y->comp(1) : nullptr,
	                   y->isVec() ? y->comp(2) : nullptr, (const int*)x->data, (const int*)y->data, s->d.n,
	                   mode, s->dpart);
	std::vector<double> h(nb);
	MFX_HIP(hipMemcpyAsync(h.data(), s->dpart, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
	sync(s);
	double m = 0;
	for (double v : h) m = v > m ? v : m;
	*out = (double)(float)m;
	API_END
}

int mfx_get_component(const mfx_grid* vec, mfx_grid* real, int c) {
	API_BEGIN
	NEED(vec, "grid");
	NEED(real, "grid");
	MFX_CHECK(c >= 0 && c < 3, "component out of range");
	Grid *v = G(vec), *r = G(real);
	MFX_CHECK(v->isVec() && r->isReal(), "getComponent: needs a Vec3 source and a Real target");
	hipLaunchKernelGGL(k_copy, dim3(flat_blocks(v->s->d.n)), dim3(256), 0, v->s->stream, r->real(),
	                   (const float*)v->comp(c), v->s->d.n);
	API_END
}
int mfx_set_component(const mfx_grid* real, mfx_grid* vec, int c) {
	API_BEGIN
	NEED(vec, "grid");
	NEED(real, "grid");
	MFX_CHECK(c >= 0 && c < 3, "component out of range");
	Grid *v = G(vec), *r = G(real);
	MFX_CHECK(v->isVec() && r->isReal(), "setComponent: needs a Real source and a Vec3 target");
	hipLaunchKernelGGL(k_copy, dim3(flat_blocks(v->s->d.n)), dim3(256), 0, v->s->stream, v->comp(c),
	                   (const float*)r->real(), v->s->d.n);
	API_END
}
int mfx_copy_mac_to_vec3(const mfx_grid* mac, mfx_grid* vec) {
	API_BEGIN
	NEED(mac, "grid");
	NEED(vec, "grid");
	Grid *m = G(mac), *v = G(vec);
	MFX_CHECK(m->isVec() && v->isVec(), "copyMacToVec3: needs Vec3 grids");
	for (int c = 0; c < 3; c++)
		hipLaunchKernelGGL(k_copy, dim3(flat_blocks(m->s->d.n)), dim3(256), 0, m->s->stream, v->comp(c),
		                   (const float*)m->comp(c), m->s->d.n);
	API_END
}

int mfx_flags_init_domain(mfx_grid* gg, int boundaryWidth, const char* wallIn, const char* openIn,
                          const char* inflowIn, const char* outflowIn) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	MFX_CHECK(g->isFlags(), "initDomain needs a FlagGrid");
	Scene* s = g->s;
	// type selection of FlagGrid::initDomain (grid.cpp:744-826)
	std::string wall = std::string(wallIn ? wallIn : "") + "      ";
	std::string open = std::string(openIn ? openIn : "") + "      ";
	std::string inflow = std::string(inflowIn ? inflowIn : "") + "      ";
	std::string outflow = std::string(outflowIn ? outflowIn : "") + "      ";
	int types[6] = {0, 0, 0, 0, 0, 0};
	bool set[6] = {false, false, false, false, false, false};
	const char dirs[6] = {'x', 'X', 'y', 'Y', 'z', 'Z'};
	for (int i = 0; i < 6; ++i)
		for (int q = 0; q < 6; ++q) {
			if (q >= 4 && !s->d.is3d) continue;
			if (set[q]) continue;
			char c = dirs[q];
			if (open[i] == c) { types[q] = kOpen; set[q] = true; }
			else if (inflow[i] == c) { types[q] = kInflow; set[q] = true; }
			else if (outflow[i] == c) { types[q] = kOutflow; set[q] = true; }
			else if (wall[i] == c) { types[q] = kObstacle; set[q] = true; }
		}
	hipLaunchKernelGGL(k_init_boundaries, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, g->ints(), s->d,
	                   boundaryWidth, types[0], types[1], types[2], types[3], types[4], types[5]);
	flags_repack(g);
	API_END
}

int mfx_flags_fill_grid(mfx_grid* gg, int type) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	MFX_CHECK(g->isFlags(), "fillGrid needs a FlagGrid");
	Scene* s = g->s;
	hipLaunchKernelGGL(k_fill_grid, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, g->ints(), s->d.n, type);
	flags_repack(g);
	API_END
}

int mfx_flags_count_cells(const mfx_grid* gg, int flag, int bnd, mfx_grid* mask, long long* out) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	Scene* s = g->s;
	MFX_CHECK(!mask || G(mask)->isReal(), "countCells: mask must be a Real grid");
	unsigned nb = flat_blocks(s->d.n, 256, 4096);
	s->ensurePartials(nb, 0);
	hipLaunchKernelGGL(k_count_cells, dim3(nb), dim3(256), 0, s->stream, (const int*)g->data, s->d, flag, bnd,
	                   mask ? G(mask)->real() : nullptr, s->dpart);
	std::vector<double> h(nb);
	MFX_HIP(hipMemcpyAsync(h.data(), s->dpart, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
	sync(s);
	double c = 0;
	for (double v : h) c += v;
	*out = (long long)c;
	API_END
}

int mfx_set_open_bound(mfx_grid* gg, int bWidth, const char* openBound, int type) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	MFX_CHECK(g->isFlags(), "setOpenBound needs a FlagGrid");
	if (!openBound || !openBound[0]) return 0;
	int lo[3] = {0, 0, 0}, up[3] = {0, 0, 0};
	for (const char* p = openBound; *p; p++) {
		switch (*p) {
		case 'x': lo[0] = 1; break;
		case 'y': lo[1] = 1; break;
		case 'z': lo[2] = 1; break;
		case 'X': up[0] = 1; break;
		case 'Y': up[1] = 1; break;
		case 'Z': up[2] = 1; break;
		default: throw Error("invalid character in boundary description string. Only [xyzXYZ] allowed.");
		}
	}
	Scene* s = g->s;
	hipLaunchKernelGGL(k_set_open_bound, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, g->ints(), s->d,
	                   bWidth, lo[0], lo[1], lo[2], up[0], up[1], up[2], type);
	flags_repack(g);
	API_END
}

int mfx_reset_outflow(mfx_grid* gg, mfx_grid* phi, mfx_grid* real) {
	API_BEGIN
	NEED(gg, "grid");
	Grid* g = G(gg);
	MFX_CHECK(g->isFlags(), "resetOutflow needs a FlagGrid");
	MFX_CHECK(!phi || G(phi)->isReal(), "resetOutflow: phi must be a Real grid");
	MFX_CHECK(!real || G(real)->isReal(), "resetOutflow: real must be a Real grid");
	Scene* s = g->s;
	flags_touched(g);
	// (the outflow flags themselves are not changed here: the list stays valid)
	const unsigned* count = nullptr;
	const unsigned* list = outflow_list(g, &count);
	hipLaunchKernelGGL(k_reset_outflow, dim3(kOutflowListBlocks), dim3(256), 0, s->stream, g->ints(), g->f8,
	                   phi ? G(phi)->real() : nullptr, real ? G(real)->real() : nullptr, list, count);
	API_END
}

int mfx_apply_shape(mfx_grid* gg, int kind, const float* a, const float* b, float r, const float* value,
                    const mfx_grid* respectFlags) {
	API_BEGIN
	NEED(gg, "grid");
	NEED(a, "shape parameters");
	NEED(value, "value");
	Grid* g = G(gg);
	Scene* s = g->s;
	MFX_CHECK(kind == MFX_SHAPE_BOX || kind == MFX_SHAPE_SPHERE || kind == MFX_SHAPE_CYLINDER, "unknown shape");
	ShapeDesc sh;
	sh.kind = kind;
	sh.a0 = a[0];
	sh.a1 = a[1];
	sh.a2 = a[2];
	sh.b0 = b ? b[0] : 1.f;
	sh.b1 = b ? b[1] : 1.f;
	sh.b2 = b ? b[2] : 1.f;
	sh.r = r;
	sh.zlen = 0;
	if (kind == MFX_SHAPE_CYLINDER) {
		// Cylinder ctor: mZDir = z; mZ = normalize(mZDir) (vectorbase.h:416-430)
		float l = sh.b0 * sh.b0 + sh.b1 * sh.b1 + sh.b2 * sh.b2;
		float nrm;
		if (std::fabs(l - 1.) < 1e-6f * 1e-6f) nrm = 1.f;
		else if (l > 1e-6f * 1e-6f) {
			nrm = std::sqrt(l);
			float sc = (float)(1. / nrm);
			sh.b0 *= sc;
			sh.b1 *= sc;
			sh.b2 *= sc;
		} else {
			sh.b0 = sh.b1 = sh.b2 = 0;
			nrm = 0.f;
		}
		sh.zlen = nrm;
	}
	int mode;
	if (g->isInt()) mode = 0;
	else if (g->isMAC()) mode = 3;
	else if (g->isVec()) mode = 2;
	else mode = 1;
	const uint8_t* resp = respectFlags ? G(respectFlags)->f8 : nullptr;
	hipLaunchKernelGGL(k_apply_shape, dim3(flat_blocks(s->d.n)), dim3(256), 0, s->stream, sh, s->d, mode,
	                   g->isInt() ? nullptr : g->comp(0), g->isVec() ? g->comp(1) : nullptr,
	                   g->isVec() ? g->comp(2) : nullptr, g->isInt() ? g->ints() : nullptr, value[0],
	                   g->isVec() ? value[1] : 0.f, g->isVec() ? value[2] : 0.f, resp);
	if (g->isFlags()) flags_repack(g);
	API_END
}

}  // extern "C"
