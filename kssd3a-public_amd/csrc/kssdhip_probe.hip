// Measures the device's rate of independent random 128-byte line reads from
// a buffer far larger than the caches, in the form the query-major join
// reads its record table: eight lanes read one line together (16 bytes
// each, nontemporal), so one wave load instruction reads eight whole lines.
// Random line fetches bound the join's lookups (one home line per query
// context).  Diagnostic only.
#include <hip/hip_runtime.h>

#include "kssdhip.h"
#include "kssdhip_internal.h"

namespace {

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

// mask = buffer lines - 1 (a power of two): every line read stays inside
// the buffer.  Each wave reads lines e .. e + 63 of the sequence per step
// (`reads` a multiple of 64), group g of eight lanes taking e + 8k + g in
// load k, all eight loads in flight together.
__global__ __launch_bounds__(256) void k_random_lines(const u64x2 *__restrict__ a, uint64_t mask, uint64_t reads,
						      unsigned long long *sink)
{
	const uint32_t g = (threadIdx.x & 63u) >> 3, piece = threadIdx.x & 7u;
	unsigned long long acc = 0;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u); e < reads; e += stride) {
		u64x2 v[8];
#pragma unroll
		for (uint32_t k = 0; k < 8; ++k) {
			const uint64_t x = e + 8 * k + g;
			const uint64_t h = mix32((uint32_t)x) | ((uint64_t)mix32((uint32_t)x ^ 0x9e3779b9u) << 32);
			v[k] = __builtin_nontemporal_load(a + (h & mask) * 8 + piece);
		}
#pragma unroll
		for (uint32_t k = 0; k < 8; ++k)
			acc += v[k].x ^ v[k].y;
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
	uint64_t lines = 1;
	while (lines * 2 * 128 <= bytes)
		lines *= 2;
	reads = (reads + 63) / 64 * 64;
	unsigned long long *buf = nullptr, *sink = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hipStream_t s = nullptr;
	int rc = 0;
	float best = 0;
	if (hipMalloc(&buf, lines * 128) != hipSuccess || hipMalloc(&sink, 8) != hipSuccess) {
		rc = kssdhip_fail(KSSDHIP_ENOMEM, "probe_random_reads: %zu bytes", (size_t)(lines * 128));
		goto out;
	}
	if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
	    hipEventCreate(&e1) != hipSuccess) {
		rc = kssdhip_fail(KSSDHIP_EHIP, "probe_random_reads: stream or events");
		goto out;
	}
	hipLaunchKernelGGL(k_fill_words, dim3(8192), dim3(256), 0, s, buf, lines * 16);
	for (int rep = 0; rep < 3; ++rep) {
		(void)hipEventRecord(e0, s);
		hipLaunchKernelGGL(k_random_lines, dim3(8192), dim3(256), 0, s, (const u64x2 *)buf, lines - 1, reads, sink);
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
