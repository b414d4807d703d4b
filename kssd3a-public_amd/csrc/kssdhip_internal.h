// Shared between the translation units of libkssdhip (not part of the ABI).
#ifndef KSSDHIP_INTERNAL_H
#define KSSDHIP_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace kssd {
constexpr int GID_NBITS = 20;                         // sketch_rearrange.h:7
constexpr uint64_t GID_MASK = (1ull << GID_NBITS) - 1ull;
constexpr int MIN_INIT = 12;                          // NUM_CODENS + 1, command_ani.h:535
constexpr size_t INDEX_PAD = 8;                       // device index arrays hold n + 8 records
constexpr uint64_t KSSD_SCAN_BLOCK = 4096;            // elements per workgroup of kssd_scan_u32_excl

// dna_popcount.h:45-49: one bit per differing 2-bit base lane
__device__ __forceinline__ int dna_popcount_dev(uint32_t x)
{
	x |= x >> 1;
	x &= 0x55555555u;
	return __popc(x);
}

// Copies between pageable host memory and the device, staged through pinned
// buffers by several host threads when large (kssdhip_copy.hip). copy_h2d
// returns once the source may be reused (the copy is queued on s); copy_d2h
// returns with dst filled.
hipError_t copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);

struct JoinArgs {
	const uint64_t *ctxgid;
	const uint32_t *robj;
	const uint32_t *bucket;
	uint64_t n;
	int shift;
	uint64_t ctx_max;
	int ref_gnum;
	const uint64_t *keys;
	const uint64_t *skidx;
	int nq;
	uint64_t total;
	int qry_gid_offset;
	unsigned obj_bits;
	unsigned flags;
	// vote counting (the sparse-edge nomination of pairwise_indexed_self_scan,
	// pairwise_graph.c:1200-1235): only every vote_step-th key of a query,
	// and only contexts with at most vote_max_freq index records (0: all)
	uint32_t vote_max_freq;
	uint32_t vote_step;
	unsigned long long *ctx;
	unsigned long long *objc;
	unsigned long long *uctx;
	unsigned long long *uobj;
};
// Largest q in [0, nq) with skidx[q] <= e (skidx[nq] > e must hold): the
// query that owns key e of a block.
__device__ __forceinline__ int owner_query(const uint64_t *__restrict__ skidx, int nq, uint64_t e)
{
	int lo = 0, hi = nq;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (skidx[mid] <= e)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

// Workspace of the one-launch plain pair kernel: per-block partial sums and
// the count of finished blocks (left at zero by every launch).
constexpr unsigned PAIR_MAX_BLOCKS = 1024;
struct PairWorkspace {
	unsigned done;
	unsigned pad[3];
	uint4 partial[PAIR_MAX_BLOCKS];
};
} // namespace kssd

// grow-only device buffer
struct DevBuf {
	void *ptr = nullptr;
	size_t bytes = 0;
	// bytes past a user's fixed header that the user keeps zero at rest
	// (reset when the buffer is reallocated)
	size_t zero_bytes = 0;
	int reserve(size_t want);
	void release();
};

// Resident index handle (opaque in the C ABI).
struct kssdhip_index {
	int device = 0;
	hipStream_t stream = nullptr;
	uint64_t n = 0;
	int ref_gnum = 0;
	unsigned obj_bits = 0;
	uint64_t *d_ctxgid = nullptr;
	uint32_t *d_obj = nullptr;
	uint32_t *d_bucket = nullptr;
	int k = 0, shift = 0;
	uint64_t ctx_max = 0;
	// partitions of the sorted index for the partitioned merge-join:
	// partition p = records [pstart[p], pstart[p+1]), pctx[p] = its first
	// context, pctx[P] = ~0; boundaries never split a context's records.
	uint32_t P = 0;
	uint32_t max_part = 0;
	uint32_t *d_pstart = nullptr;
	uint64_t *d_pctx = nullptr;
	// plut[b] = first p with pctx[p] >= b << 32 (b = 0..4096): the window of
	// partition starts holding a 44-bit context's top 12 bits
	uint32_t *d_plut = nullptr;
	// tiling (kssdhip_tiling_t)
	int mode = 0;
	int chunk = 0;
	int block = 256;
	double auto_min_keys = 0;
	int group = 1;
	int qthreads = 0; // query-major workgroup size (0: default)
	int qslots = 0;   // query-major LDS slots (0: default)
	// query-major join (kssdhip_qjoin.hip): the record table, br_lines
	// 128-byte lines (2^br_bits home lines, then br_ovf overflow lines) over
	// contexts of br_w bits; nullptr when the index cannot have one
	unsigned long long *d_brec = nullptr;
	uint64_t br_lines = 0, br_ovf = 0;
	int br_w = 0, br_bits = 0;
	int cus = 0; // compute units of the device
	int qj_per_cu[2] = {0, 0}; // query-major workgroups per CU (plain, unique counters) at the tiling
	// grow-only scratch: host-pointer entry points, the split table, the
	// query-major work counter and spill rows, select outputs
	DevBuf keys, skidx, ctx, obj, uctx, uobj, idx, split, qws, rows, qmeta, qctxb, rctxb, sdense;
};

// partitioned merge-join (kssdhip_pjoin.hip)
int kssd_build_partitions(kssdhip_index *ix, int chunk);
void kssd_free_partitions(kssdhip_index *ix);
// true when the partitioned path was launched for this call; `split` holds
// the call's split table (callers on different streams pass their own)
int kssd_launch_pjoin(kssdhip_index *ix, const kssd::JoinArgs &a, hipStream_t s, DevBuf *split, bool *used);
// kssdhip_count_block_dev with explicit workspaces (split table; query-major
// work counter and spill rows, nullptr: that kernel is not used)
int kssd_count_dev(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx, int nq, uint64_t total,
		   int qry_gid_offset, uint32_t flags, kssd_ctx_mut2_t *d_ctx, kssd_obj_section_t *d_obj,
		   kssd_ctx_mut2_t *d_uctx, kssd_obj_section_t *d_uobj, hipStream_t s, DevBuf *split, DevBuf *qws,
		   uint32_t vote_max_freq = 0, uint32_t vote_step = 1);

// exclusive scan of a[0, n) in place (kssdhip_build.hip); sums: at least
// ceil(n / KSSD_SCAN_BLOCK) words of scratch
void kssd_scan_u32_excl(uint32_t *a, uint64_t n, uint32_t *sums, hipStream_t s);

// query-major join (kssdhip_qjoin.hip)
int kssd_qjoin_build(kssdhip_index *ix);
// occupancy of the query-major kernels at the handle's tiling (index open,
// set_tiling), so that launches make no runtime queries
int kssd_qjoin_prepare(kssdhip_index *ix);
void kssd_qjoin_free(kssdhip_index *ix);
// dense rows of a block by the query-major kernel; `ws` is the call's
// workspace (work counter and spill rows)
int kssd_launch_qjoin_dense(kssdhip_index *ix, const kssd::JoinArgs &a, hipStream_t s, DevBuf *ws);
// kssdhip_select_block_dev with the caller's workspaces (a server runs one
// set per slot, so concurrent batches never share them): `ws` the
// query-major join's, `split` and `dense` the dense-row path's that small
// blocks take (nullptr: always the query-major join)
int kssd_select_dev_ws(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx, int nq, uint64_t total,
		       uint32_t flags, const uint32_t *d_qry_ctx, const uint32_t *d_ref_ctx,
		       const kssdhip_select_opts_t *opts, kssdhip_ani_row_t *d_rows, uint32_t cap,
		       uint32_t *d_row_count, uint32_t *d_q_first, uint32_t *d_q_count, hipStream_t s, DevBuf *ws,
		       DevBuf *split, DevBuf *dense);
// true when a block of this shape should take the query-major kernel in
// auto mode
bool kssd_qjoin_auto(const kssdhip_index *ix, int nq, uint64_t total);
// looks up the address of the no-allocation scratch array on `device`
// (current on the calling thread) once per device
int kssd_noalloc_setup(int device);
// autotuner (kssdhip_tune.hip): a new handle's shape when nothing is cached,
// and the cached shape parameters for its index shape (1 when found)
void kssd_default_tiling(const kssdhip_index *ix, kssdhip_tiling_t *t);
int kssd_cached_index_tiling(const kssdhip_index *ix, kssdhip_tiling_t *t);
// fails with KSSDHIP_EINVAL when `want` bytes would grow `b` while `s` is
// being captured into a graph, else reserves them
int kssd_reserve_for_stream(DevBuf *b, size_t want, hipStream_t s, const char *what);

// read-only view of an index handle for other translation units
struct IndexView {
	const uint64_t *ctxgid;
	const uint32_t *obj;
	const uint32_t *bucket;
	uint64_t n;
	int ref_gnum;
	unsigned obj_bits;
	int shift;
	int k;
	uint64_t ctx_max;
	int device;
};
struct kssdhip_index;
int kssdhip_index_info(const kssdhip_index *ix, IndexView *v);
int kssdhip_fail(int code, const char *fmt, ...);

#endif
