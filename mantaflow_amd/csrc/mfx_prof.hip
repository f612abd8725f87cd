// Optional per-kernel timing with HIP events recorded on the launching
// stream.  Off by default; enabled through mfx_kernel_timing(1).  Used by the
// benchmark to time the roofline kernel inside its timed region.
#include <atomic>
#include <mutex>

#include "mfx_internal.h"
#include "mfx_prof.h"

namespace mfx {

static std::atomic<int> g_on{0};
static std::mutex g_mu;
struct KStat {
	std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
	double totalMs = 0;
	long long count = 0;
};
static std::map<std::string, KStat> g_stats;
static std::vector<hipEvent_t> g_pool;

bool kt_on() { return g_on.load(std::memory_order_relaxed) != 0; }

static hipEvent_t ev_get() {
	if (!g_pool.empty()) {
		hipEvent_t e = g_pool.back();
		g_pool.pop_back();
		return e;
	}
	hipEvent_t e;
	MFX_HIP(hipEventCreate(&e));
	return e;
}

KtScope::KtScope(const char* n, hipStream_t s) : name(n), stream(s) {
	if (!kt_on()) return;
	std::lock_guard<std::mutex> lk(g_mu);
	a = ev_get();
	b = ev_get();
	MFX_HIP(hipEventRecord(a, stream));
	active = true;
}
KtScope::~KtScope() {
	if (!active) return;
	std::lock_guard<std::mutex> lk(g_mu);
	hipEventRecord(b, stream);
	g_stats[name].pending.push_back({a, b});
}

static void drain(KStat& k) {
	for (auto& p : k.pending) {
		hipEventSynchronize(p.second);
		float ms = 0;
		hipEventElapsedTime(&ms, p.first, p.second);
		k.totalMs += ms;
		k.count++;
		g_pool.push_back(p.first);
		g_pool.push_back(p.second);
	}
	k.pending.clear();
}

}  // namespace mfx

using namespace mfx;

extern "C" {

int mfx_kernel_timing(int enable) {
	g_on.store(enable ? 1 : 0);
	return 0;
}

int mfx_kernel_timing_reset(void) {
	std::lock_guard<std::mutex> lk(g_mu);
	for (auto& kv : g_stats) {
		drain(kv.second);
		kv.second.totalMs = 0;
		kv.second.count = 0;
	}
	return 0;
}

int mfx_kernel_timing_get(const char* name, double* totalMs, long long* count) {
	std::lock_guard<std::mutex> lk(g_mu);
	auto it = g_stats.find(name ? name : "");
	if (it == g_stats.end()) {
		*totalMs = 0;
		*count = 0;
		return 0;
	}
	drain(it->second);
	*totalMs = it->second.totalMs;
	*count = it->second.count;
	return 0;
}

}  // extern "C"
