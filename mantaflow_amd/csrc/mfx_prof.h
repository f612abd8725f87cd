// Scoped HIP-event timing of one kernel launch (active only when enabled).
#pragma once
#include <hip/hip_runtime.h>

namespace mfx {
bool kt_on();
struct KtScope {
	const char* name;
	hipStream_t stream;
	hipEvent_t a{}, b{};
	bool active = false;
	KtScope(const char* n, hipStream_t s);
	~KtScope();
};
}  // namespace mfx
