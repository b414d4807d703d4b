// Measures the device's rate of independent random 8-byte reads from a
// buffer far larger than the caches: each read fetches its own memory line,
// which is what bounds the index lookups of the query-major join (a context
// table slot, a filter word, a record run per query key).  Diagnostic only.
#include <hip/hip_runtime.h>

#include "kssdhip.h"
#include "kssdhip_internal.h"

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

// mask = words - 1 (a power of two): every index stays inside the buffer
__global__ __launch_bounds__(256) void k_random_reads(const unsigned long long *__restrict__ a, uint64_t mask,
						      uint64_t reads, unsigned long long *sink)
{
	unsigned long long acc = 0;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < reads; e += stride) {
		const uint64_t h = mix32((uint32_t)e) | ((uint64_t)mix32((uint32_t)e ^ 0x9e3779b9u) << 32);
		acc += a[h & mask];
	}
	if (acc == 0x9e3779b97f4a7c15ull) // practically never: keeps the loads live
		atomicAdd(sink, 1ull);
}

__global__ void k_fill_words(unsigned long long *p, uint64_t n)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
		p[i] = i;
}

} // namespace

extern "C" int kssdhip_probe_random_reads(size_t bytes, uint64_t reads, double *reads_per_s)
{
	if (!reads_per_s || reads == 0 || bytes < 4096)
		return kssdhip_fail(KSSDHIP_EINVAL, "probe_random_reads: bad arguments");
	*reads_per_s = 0;
	uint64_t words = 1;
	while (words * 2 * 8 <= bytes)
		words *= 2;
	unsigned long long *buf = nullptr, *sink = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hipStream_t s = nullptr;
	int rc = 0;
	float best = 0;
	if (hipMalloc(&buf, words * 8) != hipSuccess || hipMalloc(&sink, 8) != hipSuccess) {
		rc = kssdhip_fail(KSSDHIP_ENOMEM, "probe_random_reads: %zu bytes", (size_t)(words * 8));
		goto out;
	}
	if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
	    hipEventCreate(&e1) != hipSuccess) {
		rc = kssdhip_fail(KSSDHIP_EHIP, "probe_random_reads: stream or events");
		goto out;
	}
	hipLaunchKernelGGL(k_fill_words, dim3(8192), dim3(256), 0, s, buf, words);
	for (int rep = 0; rep < 3; ++rep) {
		(void)hipEventRecord(e0, s);
		hipLaunchKernelGGL(k_random_reads, dim3(8192), dim3(256), 0, s, buf, words - 1, reads, sink);
		(void)hipEventRecord(e1, s);
		float ms = 0;
		if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "probe_random_reads: %s", hipGetErrorString(hipGetLastError()));
			goto out;
		}
		if (rep > 0 && (best == 0 || ms < best))
			best = ms;
	}
	*reads_per_s = (double)reads / (best * 1e-3);
out:
	if (e0)
		(void)hipEventDestroy(e0);
	if (e1)
		(void)hipEventDestroy(e1);
	if (s)
		(void)hipStreamDestroy(s);
	if (buf)
		(void)hipFree(buf);
	if (sink)
		(void)hipFree(sink);
	return rc;
}
