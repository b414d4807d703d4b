// Partitioned merge-join: the batched counting path.
//
// The resident index is cut into P partitions of about `chunk` records whose
// boundaries never split the records of one context.  For a block of queries
// a split table S[q][p] (first key of query q whose context is >= the first
// context of partition p) is computed once; then one workgroup per partition
// stages the partition's records in LDS with coalesced loads and, one thread
// per query, merges that query's keys of the partition against the staged
// records: binary search in LDS, walk of the context's per-genome runs,
// minimum per-base mismatch over the object cross product, and per-thread
// accumulation of the pair counts, flushed with one 64-bit atomic per
// (partition, query, genome) cell.  Consecutive partitions go to the same XCD
// so the split-table and query-key lines they share stay in that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_internal.h"

using namespace kssd;

namespace {

constexpr uint32_t LDS_SLACK = 64; // partitions may exceed `chunk` by a short context run
constexpr int NUM_XCD = 8;

__global__ void k_part_cand(const uint64_t *__restrict__ ctxgid, uint64_t n, uint64_t chunk,
			    uint64_t ncand, uint32_t *cand)
{
	for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < ncand;
	     j += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t e = j * chunk;
		if (e >= n) {
			cand[j] = (uint32_t)n;
			continue;
		}
		// first record of the context that record e belongs to
		const uint64_t target = (ctxgid[e] >> GID_NBITS) << GID_NBITS;
		uint64_t lo = 0, hi = e;
		while (lo < hi) {
			const uint64_t mid = lo + ((hi - lo) >> 1);
			if (ctxgid[mid] < target)
				lo = mid + 1;
			else
				hi = mid;
		}
		cand[j] = (uint32_t)lo;
	}
}

__global__ void k_part_ctx(const uint64_t *__restrict__ ctxgid, const uint32_t *__restrict__ pstart,
			   uint32_t P, uint64_t *pctx)
{
	for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= P; p += gridDim.x * blockDim.x)
		pctx[p] = p < P ? (ctxgid[pstart[p]] >> GID_NBITS) : ~0ull;
}

// S[q*(P+1)+p] = first key index (relative to the block) of query q whose
// context is >= pctx[p]; pctx[P] = ~0 gives the end of q.
__global__ __launch_bounds__(256) void k_split(const uint64_t *__restrict__ keys,
					       const uint64_t *__restrict__ skidx, int nq,
					       unsigned obj_bits, const uint64_t *__restrict__ pctx,
					       uint32_t P, uint32_t *__restrict__ S)
{
	const uint64_t width = (uint64_t)P + 1;
	const uint64_t total = (uint64_t)nq * width;
	for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
	     t += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t q = t / width, p = t - q * width;
		uint64_t lo = skidx[q], hi = skidx[q + 1];
		const uint64_t c = pctx[p];
		while (lo < hi) {
			const uint64_t mid = lo + ((hi - lo) >> 1);
			if ((keys[mid] >> obj_bits) < c)
				lo = mid + 1;
			else
				hi = mid;
		}
		S[t] = (uint32_t)lo;
	}
}

struct PJArgs {
	JoinArgs j;
	const uint32_t *pstart;
	const uint32_t *S;
	uint32_t P;
	uint32_t cap; // records that fit the LDS stage
};

// Per-thread accumulation of consecutive hits on one (query, genome) cell.
struct CellAcc {
	uint32_t gid = ~0u;
	unsigned long long c = 0, o = 0;
	__device__ __forceinline__ void add(uint32_t g, int m, unsigned long long *ctx,
					    unsigned long long *obj, uint64_t row)
	{
		if (g != gid) {
			flush(ctx, obj, row);
			gid = g;
		}
		c += 1ull | ((unsigned long long)(m > 1) << 32);
		if (m)
			o += 1ull | ((unsigned long long)m << 32);
	}
	__device__ __forceinline__ void flush(unsigned long long *ctx, unsigned long long *obj, uint64_t row)
	{
		if (gid != ~0u) {
			atomicAdd(&ctx[row + gid], c);
			if (o)
				atomicAdd(&obj[row + gid], o);
		}
		gid = ~0u;
		c = o = 0;
	}
};

// Merge one partition against all queries of the block.  REC/OBJ point at the
// partition's records (LDS stage or global memory).
template <typename RecPtr, typename ObjPtr>
__device__ __forceinline__ void merge_partition(const PJArgs &a, uint32_t p, RecPtr rec, ObjPtr robj,
						uint32_t cnt)
{
	const JoinArgs &j = a.j;
	const uint64_t objmask = j.obj_bits >= 64 ? ~0ull : ((1ull << j.obj_bits) - 1ull);
	const bool lower = (j.flags & KSSDHIP_LOWER) != 0;
	const bool unique = !lower && (j.flags & KSSDHIP_UNIQUE) != 0;
	const bool ignore = (j.flags & KSSDHIP_IGNORECONFLICT) != 0;
	const uint64_t width = (uint64_t)a.P + 1;
	for (int q = threadIdx.x; q < j.nq; q += blockDim.x) {
		const uint64_t srow = (uint64_t)q * width + p;
		uint32_t i = a.S[srow];
		const uint32_t hi = a.S[srow + 1];
		if (i >= hi)
			continue;
		const int qgid = j.qry_gid_offset + q;
		const uint64_t row = (uint64_t)q * (uint64_t)j.ref_gnum;
		CellAcc acc, uacc;
		uint64_t key = j.keys[i];
		while (i < hi) {
			const uint64_t qctx = key >> j.obj_bits;
			uint32_t s_end = i + 1;
			uint64_t next = 0;
			while (s_end < hi) {
				next = j.keys[s_end];
				if ((next >> j.obj_bits) != qctx)
					break;
				++s_end;
			}
			// lower_bound of the context among the partition's records
			const uint64_t target = qctx << GID_NBITS;
			uint32_t lo = 0, h2 = cnt;
			while (lo < h2) {
				const uint32_t mid = (lo + h2) >> 1;
				if (rec[mid] < target)
					lo = mid + 1;
				else
					h2 = mid;
			}
			uint32_t d = lo;
			int owners = 0, sole_diff = 0;
			uint32_t sole_gid = 0;
			bool sole_allowed = false;
			while (d < cnt) {
				const uint64_t cg = rec[d];
				if ((cg >> GID_NBITS) != qctx)
					break;
				const uint32_t gid = (uint32_t)(cg & GID_MASK);
				const uint32_t rb = d;
				do {
					++d;
				} while (d < cnt && rec[d] == cg);
				const uint32_t re = d;
				if (lower && (int)gid >= qgid) {
					if ((int)gid > qgid)
						break;
					continue;
				}
				if (unique) {
					++owners;
					sole_gid = gid;
				}
				if (ignore && re - rb > 1)
					continue;
				int m = MIN_INIT;
				for (uint32_t x = i; x < s_end && m; ++x) {
					const uint32_t oq = (uint32_t)((x == i ? key : j.keys[x]) & objmask);
					for (uint32_t r = rb; r < re; ++r) {
						const uint32_t diff = oq ^ robj[r];
						if (diff == 0) {
							m = 0;
							break;
						}
						const int c = dna_popcount_dev(diff);
						m = c < m ? c : m;
					}
				}
				acc.add(gid, m, j.ctx, j.objc, row);
				if (unique) {
					sole_diff = m;
					sole_allowed = true;
				}
			}
			if (unique && owners == 1 && sole_allowed)
				uacc.add(sole_gid, sole_diff, j.uctx, j.uobj, row);
			i = s_end;
			key = next;
		}
		acc.flush(j.ctx, j.objc, row);
		if (unique)
			uacc.flush(j.uctx, j.uobj, row);
	}
}

__global__ void k_pjoin(PJArgs a)
{
	extern __shared__ uint64_t lds_rec[]; // [cap] ctxgid, then [cap] obj (u32)
	uint32_t *lds_obj = (uint32_t *)(lds_rec + a.cap);
	// XCD-aware order: workgroup b runs on XCD b % 8; give each XCD a
	// contiguous run of partitions.
	const uint32_t G = gridDim.x, b = blockIdx.x;
	const uint32_t x = b % NUM_XCD, i = b / NUM_XCD, base = G / NUM_XCD, rem = G % NUM_XCD;
	const uint32_t p = x * base + (x < rem ? x : rem) + i;
	if (p >= a.P)
		return;
	const uint32_t rb = a.pstart[p], re = a.pstart[p + 1], cnt = re - rb;
	if (cnt <= a.cap) {
		for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
			lds_rec[k] = a.j.ctxgid[rb + k];
			lds_obj[k] = a.j.robj[rb + k];
		}
		__syncthreads();
		merge_partition(a, p, (const uint64_t *)lds_rec, (const uint32_t *)lds_obj, cnt);
	} else {
		// a context run longer than the stage: merge straight from HBM
		merge_partition(a, p, a.j.ctxgid + rb, a.j.robj + rb, cnt);
	}
}

unsigned grid_cap(uint64_t n, unsigned block)
{
	uint64_t g = (n + block - 1) / block;
	if (g < 1)
		g = 1;
	if (g > (1u << 20))
		g = 1u << 20;
	return (unsigned)g;
}

} // namespace

void kssd_free_partitions(kssdhip_index *ix)
{
	if (ix->d_pstart)
		(void)hipFree(ix->d_pstart);
	if (ix->d_pctx)
		(void)hipFree(ix->d_pctx);
	ix->d_pstart = nullptr;
	ix->d_pctx = nullptr;
	ix->P = 0;
	ix->max_part = 0;
}

int kssd_build_partitions(kssdhip_index *ix, int chunk)
{
	if (chunk < 64 || chunk > 12288)
		return kssdhip_fail(KSSDHIP_EINVAL, "chunk_records %d out of range 64..12288", chunk);
	kssd_free_partitions(ix);
	ix->chunk = chunk;
	if (ix->n == 0)
		return 0;
	const uint64_t ncand = (ix->n + chunk - 1) / chunk;
	uint32_t *d_cand = nullptr;
	std::vector<uint32_t> cand(ncand);
	if (hipMalloc(&d_cand, ncand * 4) != hipSuccess)
		return kssdhip_fail(KSSDHIP_ENOMEM, "partition candidates: device allocation failed");
	hipLaunchKernelGGL(k_part_cand, dim3(grid_cap(ncand, 256)), dim3(256), 0, ix->stream, ix->d_ctxgid,
			   ix->n, (uint64_t)chunk, ncand, d_cand);
	hipError_t e = hipMemcpyAsync(cand.data(), d_cand, ncand * 4, hipMemcpyDeviceToHost, ix->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(ix->stream);
	(void)hipFree(d_cand);
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "partition candidates: %s", hipGetErrorString(e));
	std::vector<uint32_t> pstart;
	pstart.reserve(ncand + 1);
	for (uint64_t j = 0; j < ncand; ++j)
		if (pstart.empty() || cand[j] != pstart.back())
			pstart.push_back(cand[j]);
	if (pstart.empty() || pstart[0] != 0)
		pstart.insert(pstart.begin(), 0u);
	if (pstart.back() != (uint32_t)ix->n)
		pstart.push_back((uint32_t)ix->n);
	const uint32_t P = (uint32_t)pstart.size() - 1;
	uint32_t mx = 0;
	for (uint32_t p = 0; p < P; ++p)
		mx = std::max(mx, pstart[p + 1] - pstart[p]);
	if (hipMalloc(&ix->d_pstart, (P + 1) * 4) != hipSuccess ||
	    hipMalloc(&ix->d_pctx, (P + 1) * 8) != hipSuccess) {
		kssd_free_partitions(ix);
		return kssdhip_fail(KSSDHIP_ENOMEM, "partition table: device allocation failed");
	}
	e = hipMemcpyAsync(ix->d_pstart, pstart.data(), (P + 1) * 4, hipMemcpyHostToDevice, ix->stream);
	if (e == hipSuccess) {
		hipLaunchKernelGGL(k_part_ctx, dim3(grid_cap(P + 1, 256)), dim3(256), 0, ix->stream,
				   ix->d_ctxgid, ix->d_pstart, P, ix->d_pctx);
		e = hipGetLastError();
	}
	if (e == hipSuccess)
		e = hipStreamSynchronize(ix->stream);
	if (e != hipSuccess) {
		kssd_free_partitions(ix);
		return kssdhip_fail(KSSDHIP_EHIP, "partition table: %s", hipGetErrorString(e));
	}
	ix->P = P;
	ix->max_part = mx;
	return 0;
}

int kssd_launch_pjoin(kssdhip_index *ix, const JoinArgs &ja, hipStream_t s, bool *used)
{
	*used = false;
	if (ix->P == 0 || ja.total >= 0xffffffffull)
		return 0;
	const bool want = ix->mode == 2 ||
			  (ix->mode == 0 && (double)ja.total >= ix->auto_min_keys * (double)ix->P);
	if (!want)
		return 0;
	const uint64_t width = (uint64_t)ix->P + 1;
	const uint64_t sbytes = (uint64_t)ja.nq * width * 4;
	int rc = ix->split.reserve(sbytes);
	if (rc)
		return rc;
	uint32_t *S = (uint32_t *)ix->split.ptr;
	hipLaunchKernelGGL(k_split, dim3(grid_cap((uint64_t)ja.nq * width, 256)), dim3(256), 0, s, ja.keys,
			   ja.skidx, ja.nq, ja.obj_bits, (const uint64_t *)ix->d_pctx, ix->P, S);
	PJArgs a;
	a.j = ja;
	a.pstart = ix->d_pstart;
	a.S = S;
	a.P = ix->P;
	a.cap = (uint32_t)ix->chunk + LDS_SLACK;
	const size_t lds = (size_t)a.cap * 12;
	static thread_local size_t lds_set = 0;
	if (lds > 65536 && lds_set < lds) {
		hipError_t e = hipFuncSetAttribute((const void *)k_pjoin,
						   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
		if (e != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "pjoin: LDS %zu bytes: %s", lds, hipGetErrorString(e));
		lds_set = lds;
	}
	hipLaunchKernelGGL(k_pjoin, dim3(ix->P), dim3(ix->block), lds, s, a);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "pjoin launch: %s", hipGetErrorString(e));
	*used = true;
	return 0;
}

extern "C" int kssdhip_index_set_tiling(kssdhip_index *ix, const kssdhip_tiling_t *t)
{
	if (!ix || !t)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: null argument");
	if (t->mode < 0 || t->mode > 2)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: mode %d", t->mode);
	if (t->block_threads < 64 || t->block_threads > 1024 || t->block_threads % 64)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: block_threads %d", t->block_threads);
	if (t->auto_min_keys < 0)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: auto_min_keys < 0");
	if (hipSetDevice(ix->device) != hipSuccess || hipStreamSynchronize(ix->stream) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "set_tiling: device not ready");
	if (t->chunk_records != ix->chunk) {
		int rc = kssd_build_partitions(ix, t->chunk_records);
		if (rc)
			return rc;
	}
	ix->mode = t->mode;
	ix->block = t->block_threads;
	ix->auto_min_keys = t->auto_min_keys;
	return 0;
}

extern "C" int kssdhip_index_get_tiling(const kssdhip_index *ix, kssdhip_tiling_t *t)
{
	if (!ix || !t)
		return kssdhip_fail(KSSDHIP_EINVAL, "get_tiling: null argument");
	t->mode = ix->mode;
	t->chunk_records = ix->chunk;
	t->block_threads = ix->block;
	t->auto_min_keys = ix->auto_min_keys;
	return 0;
}
