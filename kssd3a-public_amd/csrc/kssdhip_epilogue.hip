// FP64 epilogue: counts -> distance / ANI per cell, fused after the join.
#include <hip/hip_runtime.h>

#include "kssd_models.h"
#include "kssdhip.h"
#include "kssdhip_internal.h"
#include "kssdhip_epi.h"

using namespace kssd;

namespace {

struct EpiArgs {
	const uint2 *ctx; // {num_ctx, num_mut2_ctx}
	const uint2 *obj; // {diff_obj, diff_obj_section}
	const uint32_t *qry_ctx;
	const uint32_t *ref_ctx;
	int nq, nr;
	kssdhip_epilogue_opts_t o;
	double *out;
};

__global__ __launch_bounds__(256) void k_epilogue(EpiArgs a)
{
	const uint64_t cells = (uint64_t)a.nq * (uint64_t)a.nr;
	const double exc = a.o.values_ani ? 1.0 - a.o.exception : a.o.exception;
	const int s = epi_effective_metric(a.o);
	for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < cells;
	     c += (uint64_t)gridDim.x * blockDim.x) {
		const int q = (int)(c / (uint64_t)a.nr);
		const int r = (int)(c - (uint64_t)q * (uint64_t)a.nr);
		if (a.o.lower && r >= a.o.qry_gid_offset + q)
			continue;
		const uint32_t qc = a.qry_ctx[q], rc = a.ref_ctx[r];
		if (qc == 0 || rc == 0) {
			a.out[c] = exc;
			continue;
		}
		const uint2 cm = a.ctx[c], os = a.obj[c];
		a.out[c] = epi_cell_value(a.o, s, cm.x, cm.y, os.x, os.y, qc, rc);
	}
}

__global__ void k_zero_count(uint32_t *count)
{
	if (blockIdx.x == 0 && threadIdx.x == 0)
		*count = 0;
}

// Cells with value >= min_value, compacted one wave at a time: a ballot of
// the wave's passing cells, one counter atomic per wave, each lane writing at
// its rank among them (stored only below cap).
__global__ __launch_bounds__(256) void k_select_hits(const double *__restrict__ v, uint64_t cells, uint32_t nr,
						      double min_value, uint32_t cap, kssdhip_hit_t *hits,
						      uint32_t *count)
{
	const uint32_t lane = threadIdx.x & 63;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < cells; base += stride) {
		const uint64_t c = base + lane;
		const double x = c < cells ? v[c] : 0.0;
		const bool hit = c < cells && x >= min_value;
		const uint64_t m = __ballot(hit);
		if (!m)
			continue;
		uint32_t at = 0;
		if (lane == 0)
			at = atomicAdd(count, (uint32_t)__popcll(m));
		at = __shfl(at, 0);
		if (hit) {
			const uint32_t k = at + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
			if (k < cap)
				hits[k] = kssdhip_hit_t{(uint32_t)(c / nr), (uint32_t)(c % nr), x};
		}
	}
}

} // namespace

extern "C" int kssdhip_select_hits_dev(const double *d_values, int nq, int nr, double min_value, uint32_t cap,
				       kssdhip_hit_t *d_hits, uint32_t *d_count, void *stream)
{
	if (nq < 0 || nr < 0 || !d_count || (cap && !d_hits))
		return kssdhip_fail(KSSDHIP_EINVAL, "select_hits: bad arguments");
	const uint64_t cells = (uint64_t)nq * (uint64_t)nr;
	if (cells && !d_values)
		return kssdhip_fail(KSSDHIP_EINVAL, "select_hits: null values");
	hipStream_t s = (hipStream_t)stream;
	hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, s, d_count);
	if (cells) {
		uint64_t g = (cells + 255) / 256;
		if (g > 4096)
			g = 4096;
		hipLaunchKernelGGL(k_select_hits, dim3((unsigned)g), dim3(256), 0, s, d_values, cells, (uint32_t)nr,
				   min_value, cap, d_hits, d_count);
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "select_hits launch: %s", hipGetErrorString(e));
	return 0;
}

extern "C" int kssdhip_epilogue_dev(const kssd_ctx_mut2_t *d_ctx, const kssd_obj_section_t *d_obj,
				    int nq, int nr, const uint32_t *d_qry_ctx, const uint32_t *d_ref_ctx,
				    const kssdhip_epilogue_opts_t *opts, double *d_out, void *stream)
{
	if (!opts || nq < 0 || nr < 0)
		return kssdhip_fail(KSSDHIP_EINVAL, "epilogue: bad arguments");
	if (!epi_metric_supported(*opts))
		return kssdhip_fail(KSSDHIP_EINVAL, "epilogue: metric %d needs the calibration model (use 3..9)",
				    opts->metric);
	const uint64_t cells = (uint64_t)nq * (uint64_t)nr;
	if (!cells)
		return 0;
	if (!d_ctx || !d_obj || !d_qry_ctx || !d_ref_ctx || !d_out)
		return kssdhip_fail(KSSDHIP_EINVAL, "epilogue: null device pointer");
	EpiArgs a;
	a.ctx = (const uint2 *)d_ctx;
	a.obj = (const uint2 *)d_obj;
	a.qry_ctx = d_qry_ctx;
	a.ref_ctx = d_ref_ctx;
	a.nq = nq;
	a.nr = nr;
	a.o = *opts;
	a.out = d_out;
	uint64_t g = (cells + 255) / 256;
	if (g > (1u << 20))
		g = 1u << 20;
	hipLaunchKernelGGL(k_epilogue, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, a);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "epilogue launch: %s", hipGetErrorString(e));
	return 0;
}
