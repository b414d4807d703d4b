/*
 * kssdhip -- MI355X (gfx950) sketch-comparison kernels behind a plain C ABI.
 *
 * This is the drop-in boundary for the KSSD3A `ani` / `matrix` counting path.
 * Every type below is byte-compatible with the reference's in-memory types,
 * so reference-style host code can hand its own arrays straight in:
 *
 *   kssd_ctxgidobj_t    == ctxgidobj_t    sketch_rearrange.h:19 (packed, 12 B)
 *   kssd_ctx_mut2_t     == ctx_mut2_t     command_ani.h:120-124
 *   kssd_obj_section_t  == obj_section_t  command_ani.h:114-118
 *   kssd_ani_features_t == ani_features_t model_ani.h:215-222
 *
 * Conventions
 *   - Return value: 0 on success, a negative KSSDHIP_E* code on failure;
 *     kssdhip_last_error() then returns a message (per calling thread).
 *     The reference exits the process instead (err()/errx()); the host
 *     front end of this package maps a nonzero return to that behaviour.
 *   - Functions without a _dev suffix take HOST pointers and are synchronous.
 *     Functions with a _dev suffix take DEVICE pointers and a hipStream_t
 *     passed as void*; they are asynchronous on that stream.
 *   - Outputs are OVERWRITTEN (the library zeroes them); the reference
 *     accumulates into arrays the caller zeroed, which gives the same result.
 *   - One kssdhip_index may be used by one host thread at a time; distinct
 *     handles are independent.
 */
#ifndef KSSDHIP_H
#define KSSDHIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct __attribute__((packed)) {
	uint64_t ctxgid; /* ctx << 20 | gid */
	uint32_t obj;
} kssd_ctxgidobj_t;

typedef struct {
	uint32_t num_ctx;
	uint32_t num_mut2_ctx;
} kssd_ctx_mut2_t;

typedef struct {
	uint32_t diff_obj;
	uint32_t diff_obj_section;
} kssd_obj_section_t;

typedef struct {
	uint32_t XnY_ctx;
	uint32_t X_ctx;
	uint32_t N_diff_obj_section;
	uint32_t N_mut2_ctx;
	uint32_t N_diff_obj;
} kssd_ani_features_t;

/* byte compatibility with the reference's types (sketch_rearrange.h:19,
 * command_ani.h:114-124, model_ani.h:215-222) */
#ifdef __cplusplus
#define KSSDHIP_STATIC_ASSERT static_assert
#else
#define KSSDHIP_STATIC_ASSERT _Static_assert
#endif
KSSDHIP_STATIC_ASSERT(sizeof(kssd_ctxgidobj_t) == 12, "ctxgidobj_t is 12 packed bytes");
KSSDHIP_STATIC_ASSERT(sizeof(kssd_ctx_mut2_t) == 8, "ctx_mut2_t is 2 x u32");
KSSDHIP_STATIC_ASSERT(sizeof(kssd_obj_section_t) == 8, "obj_section_t is 2 x u32");
KSSDHIP_STATIC_ASSERT(sizeof(kssd_ani_features_t) == 20, "ani_features_t is 5 x u32");

typedef struct kssdhip_index kssdhip_index;

/* count flags */
#define KSSDHIP_LOWER 0x1u          /* only gid < qry_gid (command_ani.h:685-689) */
#define KSSDHIP_UNIQUE 0x2u         /* also fill unique-owner counters (command_ani.h:564-574) */
#define KSSDHIP_IGNORECONFLICT 0x4u /* skip reference runs with >1 object (command_ani.h:532) */
#define KSSDHIP_ACCUMULATE 0x8u     /* _dev only: add into the outputs instead of
                                       overwriting them (the reference's own
                                       accumulate-into-zeroed contract) */

/* error codes */
#define KSSDHIP_EINVAL (-1)
#define KSSDHIP_ENOMEM (-2)
#define KSSDHIP_EHIP (-3)
#define KSSDHIP_ENODEV (-4)
#define KSSDHIP_EAGAIN (-5) /* serving: queue full / not done yet */
#define KSSDHIP_ENOSPC (-6) /* an output did not fit its buffer; the call
                               reports the size it needs */

/* Message for the last failure on this thread ("" if none). */
const char *kssdhip_last_error(void);

/* Select the HIP device used by handles created afterwards on this thread.
 * No reference counterpart (the reference has no device). */
int kssdhip_init(int device);

/*
 * Upload a sorted context index (the contents of sortedcomb_ctxgid64obj32)
 * to HBM and build its bucket table there.
 * Replaces: reading the index (command_ani.c:508-516, :1550-1575) and
 * kssd_build_fenceposts_ctxgid (command_ani.c:3656-3678), which the
 * reference redoes on every block (command_ani.h:476, :650).
 * records must be sorted by (ctxgid, obj) (kssdlib_sort.c:145-148);
 * n < 2^32; ref_gnum <= 2^20; obj_bits in 1..32.
 */
int kssdhip_index_open(const kssd_ctxgidobj_t *records, size_t n, int ref_gnum,
		       unsigned obj_bits, kssdhip_index **out);
int kssdhip_index_close(kssdhip_index *index);

/*
 * Build the sorted context index of a sketch on the current device: the
 * records {ctx << 20 | gid, obj} of every key sorted by (ctxgid, obj)
 * (the bytes of sortedcomb_ctxgid64obj32) and, when unique_bits is not
 * NULL, its unique-marker bitset ((n + 7) / 8 bytes, LSB first: bit i set
 * iff record i's object occurs once among its context's records).
 * Replaces: gen_inverted_index4comblco (command_sketch.c:916-938:
 * ctxobj64_2ctxgidobj, sketch_rearrange.c:65-78, and ctxgidobj_sort_array,
 * kssdlib_sort.c:215-219) and write_unique_marker_bits_from_sorted
 * (command_sketch.c:855-914).
 * keys/skidx: comblco and comblco.index (skidx[0] == 0, n = skidx[nsamples]
 * < 2^32 - 1); each sample's keys ascending, as comblco stores them
 * (KSSDHIP_EINVAL otherwise); nsamples < 2^20; obj_bits in 1..32 with every
 * context (key >> obj_bits) below 2^44.  records: n entries.
 */
int kssdhip_build_index(const uint64_t *keys, const uint64_t *skidx, int nsamples, unsigned obj_bits,
			kssd_ctxgidobj_t *records, uint8_t *unique_bits);
/* Device bytes held by the handle (index arrays, bucket table, scratch). */
size_t kssdhip_index_device_bytes(const kssdhip_index *index);

/*
 * Kernel shape of the counting path for a handle.  No reference counterpart
 * (the reference's block loop has no device tiling); the project's
 * autotuner sets it.
 *   mode 0 auto: the query-major join whenever the index has a record
 *          table (indexes of fewer than 2^20 genomes); else, per call, the
 *          lookup kernel for small blocks (fewer than auto_min_keys keys
 *          per index partition) and the partitioned merge-join otherwise;
 *        1 lookup: one thread per query key, bucket-table lookup into HBM,
 *          counters updated with global atomics;
 *        2 partitioned: one workgroup per index partition of about
 *          chunk_records records staged in LDS; queries are split at the
 *          partition boundaries and merged against the staged records;
 *        3 query-major: persistent workgroups take whole queries from a
 *          device work queue (largest first) or, when the block has fewer
 *          queries than twice the resident workgroups, key-range items
 *          (rounds of halving sizes, then 1,024 keys); each query context
 *          is looked up in the index's record table (one 128-byte home
 *          line), the query's counters are summed in an LDS table and
 *          written once per query (items add theirs into the outputs; no
 *          per-match counter atomics in HBM).
 * chunk_records (partitions are rebuilt when it changes) must leave the
 * staged partition within the 160 KB LDS: 64 <= chunk_records <= 12288.
 * group_partitions (1..64): consecutive partitions one workgroup walks;
 * above 1 the workgroup keeps one LDS slot per query that sums the hits on
 * the first genome flushed for it over all the group's partitions and adds
 * it to the outputs once (hits on other genomes go to the outputs
 * directly), which needs chunk_records <= 6144.
 * query_threads (mode 3; 0 = default 512): threads per query-major
 * workgroup, 64..512, multiple of 64.  query_slots (mode 3): 0 = a direct
 * per-genome LDS row when it fits (ref_gnum x 16 bytes, 32 with unique
 * counters, up to 64 KiB), else 512 hashed slots; otherwise the slots of a
 * hashed LDS table per query, a power of two in 256..8192, also where the
 * direct row would fit (a smaller table lets more workgroups share a CU); a
 * query that touches more genomes than the slots spills the rest to a
 * per-workgroup row in HBM.
 */
typedef struct {
	int mode;
	int chunk_records;
	int block_threads; /* 64..1024, multiple of 64 */
	double auto_min_keys;
	int group_partitions;
	int query_threads;
	int query_slots;
} kssdhip_tiling_t;
int kssdhip_index_set_tiling(kssdhip_index *index, const kssdhip_tiling_t *tiling);
int kssdhip_index_get_tiling(const kssdhip_index *index, kssdhip_tiling_t *tiling);

/*
 * Autotuner (no reference counterpart; the reference's knobs are its fixed
 * block sizes, command_ani.c:403, :1517-1529, command_matrix.c:689-699).
 * kssdhip_index_autotune times every candidate kernel shape (the lookup
 * kernel; the partitioned merge-join over chunk sizes, workgroup sizes and
 * partition groups; the query-major join over workgroup sizes and LDS
 * slots) on a device-resident block (relative offsets, as
 * kssdhip_count_block_dev; the outputs go to scratch the call allocates),
 * applies the fastest to the handle and records it in a JSON cache keyed
 * by the device and the block's shape (index records, query keys, queries
 * and references, each rounded down to a power of two, and the flags).
 * With force == 0 a cached entry for the shape is applied without timing.
 * cache_path NULL: $KSSDHIP_TUNE_CACHE, else ~/.cache/kssdhip/tuning.json
 * (written atomically under an advisory lock).  Synchronous.
 * kssdhip_index_autotune_host: the same for a host block with ABSOLUTE
 * offsets (as kssdhip_count_block).
 * kssdhip_index_tiling_from_cache: applies the cached shape for a block of
 * this shape if there is one (returns 1), else leaves the handle as it is
 * (returns 0); no timing.
 * kssdhip_index_open starts a handle from the cached shape parameters of
 * the largest block tuned for an index of its shape on its device (kernel
 * still chosen per call), else from partition groups fitted to the compute
 * units; KSSDHIP_NO_TUNE_CACHE set skips the cache.
 */
typedef struct {
	kssdhip_tiling_t tiling; /* applied to the handle */
	double ms;               /* the block's counting time with it */
	int from_cache;          /* 1: taken from the cache, nothing timed */
	int candidates;          /* shapes timed */
	int cached;              /* 1: the result was written to the cache */
} kssdhip_tune_result_t;
int kssdhip_index_autotune(kssdhip_index *index, const uint64_t *d_keys, const uint64_t *d_skidx, int section_gnum,
			   uint64_t total_keys, uint32_t flags, const char *cache_path, int force,
			   kssdhip_tune_result_t *result);
int kssdhip_index_autotune_host(kssdhip_index *index, const uint64_t *section_sk, const uint64_t *section_skidx,
				int section_gnum, uint32_t flags, const char *cache_path, int force,
				kssdhip_tune_result_t *result);
int kssdhip_index_tiling_from_cache(kssdhip_index *index, int section_gnum, uint64_t total_keys, uint32_t flags,
				    const char *cache_path, kssdhip_tune_result_t *result);

/*
 * Count one block of queries against the index.
 * Replaces the counting part of
 *   count_ctx_obj_frm_comb_sketch_section_with_unique (command_ani.h:463-575)
 *   count_ctx_obj_frm_comb_sketch_section_lower       (command_ani.h:635-721)
 * Arguments mirror the reference: section_sk are the block's query keys
 * (each query sorted ascending), section_skidx are section_gnum+1 ABSOLUTE
 * offsets into the query comblco (the block starts at section_skidx[0]),
 * qry_gid_offset is the gid of the block's first query (LOWER only).
 * ctx/obj (and uctx/uobj with KSSDHIP_UNIQUE) are dense row-major
 * [section_gnum][ref_gnum] arrays (MCTX/MOBJ, command_ani.h:435-436).
 * Ranking, filtering and printing stay with the caller (command_ani.h:577-621).
 */
int kssdhip_count_block(kssdhip_index *index, const uint64_t *section_sk,
			const uint64_t *section_skidx, int section_gnum,
			int qry_gid_offset, uint32_t flags, kssd_ctx_mut2_t *ctx,
			kssd_obj_section_t *obj, kssd_ctx_mut2_t *unique_ctx,
			kssd_obj_section_t *unique_obj);

/*
 * Device-resident form of kssdhip_count_block for batched and serving use:
 * d_keys/d_skidx/d_ctx/... are device pointers; d_skidx holds section_gnum+1
 * offsets RELATIVE to d_keys (d_skidx[0] = 0, d_skidx[section_gnum] =
 * total_keys).  Asynchronous on `stream` (a hipStream_t).
 * No reference counterpart beyond kssdhip_count_block.
 */
int kssdhip_count_block_dev(kssdhip_index *index, const uint64_t *d_keys,
			    const uint64_t *d_skidx, int section_gnum,
			    uint64_t total_keys, int qry_gid_offset, uint32_t flags,
			    kssd_ctx_mut2_t *d_ctx, kssd_obj_section_t *d_obj,
			    kssd_ctx_mut2_t *d_unique_ctx,
			    kssd_obj_section_t *d_unique_obj, void *stream);

/*
 * Votes of the sparse-edge nomination (matrix --format edges/clusters on an
 * indexed sketch): for every query of the block and every genome gid below
 * the query's own (qry_gid_offset + q), the number of the query's sampled
 * keys (every sample_step-th, from its first) whose context the genome
 * carries, counting only contexts with at most max_ctx_freq index records
 * (0: all).  votes: nq * ref_gnum u32, row-major.
 * Replaces: the vote loop of pairwise_indexed_self_scan
 * (pairwise_graph.c:1200-1235).  Same key and offset conventions as
 * kssdhip_count_block; sample_step >= 1.
 */
int kssdhip_count_votes(kssdhip_index *index, const uint64_t *section_sk, const uint64_t *section_skidx,
			int nq, int qry_gid_offset, uint32_t max_ctx_freq, uint32_t sample_step, uint32_t *votes);

/*
 * Leftmost index record whose context equals each query key's context, or
 * SIZE_MAX.  Replaces kssd_find_first_occurrences_fenceposts
 * (command_ani.c:3703-3738).  a: n sorted keys (host); idx: n results (host).
 */
int kssdhip_find_first(kssdhip_index *index, const uint64_t *a, size_t n,
		       uint64_t *idx);

/*
 * Plain kernel: the features of ONE (query, reference) pair, one launch.
 * Replaces get_ani_features_ctx_min_over_conflicts_both_filtered
 * (command_ani.h:303-346).  a (n keys) and b (m keys) are host arrays, each
 * sorted ascending; X_ctx of the result is 0, as in the reference.
 */
int kssdhip_pair_features(const uint64_t *a, size_t n, const uint64_t *b,
			  size_t m, unsigned obj_bits, int ignore_b_conflict,
			  kssd_ani_features_t *out);

/*
 * Device-resident form of the plain kernel: exactly one kernel launch per
 * call.  d_a/d_b are device arrays, d_out a device kssd_ani_features_t that
 * is overwritten.  d_workspace is KSSDHIP_PAIR_WORKSPACE_BYTES of device
 * memory prepared once with kssdhip_pair_workspace_init_dev; every call
 * leaves it ready for the next one.  Calls sharing a workspace must be
 * ordered (same stream).  Asynchronous on `stream`.
 */
#define KSSDHIP_PAIR_WORKSPACE_BYTES 16400u
int kssdhip_pair_workspace_init_dev(void *d_workspace, void *stream);
int kssdhip_pair_features_dev(const uint64_t *d_a, size_t n, const uint64_t *d_b,
			      size_t m, unsigned obj_bits, int ignore_b_conflict,
			      kssd_ani_features_t *d_out, void *d_workspace, void *stream);

/*
 * FP64 epilogue on the device: one value per (query, reference) cell of a
 * dense count block, as the matrix outputs compute it.
 * Replaces, for the metrics that need no calibration model:
 *   ani_matrix_pair_value        command_ani.c:1458-1496
 *   ani_indexed_self_pair_metric command_ani.c:1627-1655
 *   the matrix branch of ani_block_print_internal command_ani.c:3479-3515
 * via make_selected_output_row (command_ani.c:1349-1379).
 * metric is abs(-s): 3 CtxMoE, 4 Naive, 5 MashD, 6 AafD, 7 MashD_if_far,
 * 8 AafD_if_far, 9 p_dist.  values_ani != 0 writes 1 - distance (s < 0).
 * float_af != 0 rounds the align fractions to float first, as the
 * mem_eff block printer does (command_ani.c:3489-3490).
 * Cells with qry_ctx == 0 or ref_ctx == 0 get the exception value
 * (e, or 1 - e for ANI).  lower != 0 evaluates only rgid < qry_gid_offset+q
 * and writes nothing elsewhere.
 */
typedef struct {
	int metric;
	int values_ani;
	int unassembled;  /* --qraw without --unified-metric: metrics 1..4 use Naive */
	int float_af;
	int lower;
	int qry_gid_offset;
	int drfold;       /* ani_model_drfold */
	unsigned obj_bits; /* Bitslen.obj */
	unsigned ctx_bits; /* Bitslen.ctx; Mash/AAF use k = ctx_bits / 2 */
	double afcut;
	double exception;  /* -e */
} kssdhip_epilogue_opts_t;

int kssdhip_epilogue_dev(const kssd_ctx_mut2_t *d_ctx, const kssd_obj_section_t *d_obj,
			 int nq, int nr, const uint32_t *d_qry_ctx, const uint32_t *d_ref_ctx,
			 const kssdhip_epilogue_opts_t *opts, double *d_out, void *stream);

/*
 * Cells of a dense value block (nq x nr doubles, e.g. kssdhip_epilogue_dev's
 * ANI) with value >= min_value, as (query, ref, value) in no particular
 * order; at most `cap` are stored and *d_count receives how many there are
 * (it may exceed cap).  The call zeroes *d_count with a kernel, so it can be
 * captured in a graph.  Replaces, for device-resident blocks, the anicut
 * test of the reference's row selection (`selected_ani >= anicut`,
 * command_ani.c:3517, :2197); ranking and output stay on the host.
 */
typedef struct {
	uint32_t query, ref;
	double value;
} kssdhip_hit_t;
int kssdhip_select_hits_dev(const double *d_values, int nq, int nr, double min_value, uint32_t cap,
			    kssdhip_hit_t *d_hits, uint32_t *d_count, void *stream);

/*
 * Count a block of queries against the index AND apply the index path's
 * pass gate on the device, returning only the (query, reference) pairs
 * that pass it.  Replaces count_ctx_obj_frm_comb_sketch_section_with_unique
 * with fmt == 0 (command_ani.h:463-621): the counting and the per-pair gate
 * of its ranking loop (command_ani.h:585-606):
 *   shared contexts >= ctxcut;
 *   qry_ctx[q] > 0, ref_ctx[r] > 0 and max(shared/qry_ctx, shared/ref_ctx)
 *     >= afcut (ani_report_af_pass, command_ani.h:70-78);
 *   rank_ani >= anicut, where rank_ani is the float-rounded crude ANI
 *     (float)(1 - (float)diff_obj / num_ctx / 7) or, with unique_rank, the
 *     unique-context naive ANI (command_ani.h:597-604).
 * The ordering and top-N cut of the passing pairs (command_ani.h:607-621)
 * stay with the caller, who gets every passing pair with its counts.
 * ctxcut must be >= 1 (with 0 every pair passes the first gate; use
 * kssdhip_count_block).  With unique_rank the device keeps pairs whose
 * unique-context ANI is within 1e-9 below anicut as well, so a caller that
 * re-evaluates the rule on the host in FP64 sees every pair it would keep.
 * value: the FP64 value of each kept pair as kssdhip_epilogue_dev computes
 * it for a cell (its lower/qry_gid_offset fields are ignored).
 * flags: KSSDHIP_UNIQUE (needed for unique_rank), KSSDHIP_IGNORECONFLICT.
 */
typedef struct {
	uint32_t ctxcut;
	float afcut;  /* float, as ani_opt_t holds it (command_ani.h:38) */
	float anicut; /* command_ani.h:40 */
	int unique_rank;
	kssdhip_epilogue_opts_t value;
} kssdhip_select_opts_t;

typedef struct {
	uint32_t query; /* block-relative query index */
	uint32_t ref;   /* reference gid */
	kssd_ctx_mut2_t ctx;
	kssd_obj_section_t obj;
	kssd_ctx_mut2_t unique_ctx;    /* zero without KSSDHIP_UNIQUE */
	kssd_obj_section_t unique_obj;
	double rank_ani; /* the gate's ranking value (see above) */
	double value;    /* opts->value metric of the pair */
} kssdhip_ani_row_t;
KSSDHIP_STATIC_ASSERT(sizeof(kssdhip_ani_row_t) == 56, "kssdhip_ani_row_t is 56 bytes");

/*
 * Host-pointer form: section_sk / section_skidx as kssdhip_count_block
 * (absolute offsets), qry_ctx: section_gnum context totals of the block's
 * queries, ref_ctx: ref_gnum totals of the references.  rows receives the
 * passing pairs ordered by (query, ref); *n_rows is set to their number.
 * When it exceeds cap nothing beyond cap is stored and KSSDHIP_ENOSPC is
 * returned, so the caller can grow rows and call again.
 */
int kssdhip_select_block(kssdhip_index *index, const uint64_t *section_sk, const uint64_t *section_skidx,
			 int section_gnum, uint32_t flags, const uint32_t *qry_ctx, const uint32_t *ref_ctx,
			 const kssdhip_select_opts_t *opts, kssdhip_ani_row_t *rows, size_t cap, size_t *n_rows);

/*
 * Device-resident form: relative offsets as kssdhip_count_block_dev.  The
 * passing pairs of query q are d_rows[d_q_first[q] .. + d_q_count[q]) in no
 * particular order; *d_row_count receives the total, which may exceed cap
 * (rows past cap are not stored).  The call zeroes its counters with
 * kernels and, once the handle's workspace has grown to the block's size,
 * allocates nothing, so it can be captured in a HIP graph (capturing a call
 * that would grow the workspace fails with KSSDHIP_EINVAL).  A graph holds
 * the workspace's address: after a call on the same handle that grows it
 * (a larger block, another tiling) the graph must not be replayed again.
 */
int kssdhip_select_block_dev(kssdhip_index *index, const uint64_t *d_keys, const uint64_t *d_skidx, int section_gnum,
			     uint64_t total_keys, uint32_t flags, const uint32_t *d_qry_ctx, const uint32_t *d_ref_ctx,
			     const kssdhip_select_opts_t *opts, kssdhip_ani_row_t *d_rows, uint32_t cap,
			     uint32_t *d_row_count, uint32_t *d_q_first, uint32_t *d_q_count, void *stream);

/*
 * kssdhip_select_block_dev with no allocation at all: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results; the work queue, work order and spill rows live in
 * one file-scope device array of KSSDHIP_NOALLOC_SCRATCH_BYTES per device
 * (allocated when the library's device code loads and held until the
 * process ends) instead of the handle's grow-only workspace.  Limits:
 * section_gnum <= KSSDHIP_NOALLOC_MAX_QUERIES, and with reference sets too
 * large for per-genome LDS rows (ref_gnum * 16 bytes, or 32 with
 * KSSDHIP_UNIQUE, above 64 KiB) the kernel runs as many workgroups as
 * spill rows of ref_gnum cells fit in the array's spill part, at least
 * one; larger requests fail with KSSDHIP_EINVAL before anything is
 * launched.  Every call on a device shares that device's array and clears
 * what it needs zeroed with kernels; calls that overlap on one device (any
 * handles or streams) race and must be serialised by the caller.
 */
#define KSSDHIP_NOALLOC_SCRATCH_BYTES (64u << 20)
#define KSSDHIP_NOALLOC_MAX_QUERIES (1u << 20)
int kssdhip_select_block_dev_noalloc(kssdhip_index *index, const uint64_t *d_keys, const uint64_t *d_skidx,
				     int section_gnum, uint64_t total_keys, uint32_t flags, const uint32_t *d_qry_ctx,
				     const uint32_t *d_ref_ctx, const kssdhip_select_opts_t *opts,
				     kssdhip_ani_row_t *d_rows, uint32_t cap, uint32_t *d_row_count, uint32_t *d_q_first,
				     uint32_t *d_q_count, void *stream);

/*
 * Diagnostic: the number of index records whose context occurs in a
 * device-resident block (relative offsets as kssdhip_count_block_dev),
 * i.e. the records a join of the block must read at least once.  No
 * reference counterpart; used to size work and to state the compulsory
 * traffic of a block.  Synchronous (allocates and frees a byte per slot of
 * the record table).
 */
int kssdhip_index_matched_records(kssdhip_index *index, const uint64_t *d_keys, const uint64_t *d_skidx,
				  int section_gnum, uint64_t total_keys, uint64_t *matched);

/*
 * Diagnostic: the layout of the handle's record table, the structure the
 * query-major join reads one 128-byte line of per query context: lines in
 * all (0: the handle has none, e.g. an index of 2^20 genomes), overflow
 * lines among them, home-line bits (2^home_bits home lines) and context
 * bits.  No reference counterpart.
 */
int kssdhip_index_table_info(const kssdhip_index *index, uint64_t *lines, uint64_t *overflow_lines, int *home_bits,
			     int *ctx_bits);

/*
 * Diagnostic: the device's rate of independent random 128-byte line reads
 * from a buffer of `bytes` (rounded down to a power of two; allocated and
 * freed by the call), `reads` lines (rounded up to a multiple of 64) per
 * timed launch, best of two after a warm-up.  Lines are read as the
 * query-major join reads its record table (eight lanes per line, 16-byte
 * nontemporal loads).  No reference counterpart: it measures the bound of
 * the join's lookups (one home line per query context), against which its
 * line requests are set.  Synchronous.
 */
int kssdhip_probe_random_reads(size_t bytes, uint64_t reads, double *reads_per_s);

/*
 * Serving loop (no reference counterpart: the reference answers one request
 * per process, stream_ref_sketches_one_qraw_lookup, command_ani.c:2490-2666).
 * A server keeps the index resident and answers query sketches submitted
 * from any host thread with continuous batching: whenever one of `slots`
 * streams is free, the queued requests at the head of the queue are admitted
 * into one launch while their keys and dense counter rows fit that slot's
 * share of `budget_bytes` of device memory (and pinned host memory of the
 * same size); requests arriving meanwhile go to the next free slot, so
 * batches overlap on the device.  Counter rows are exactly those of
 * kssdhip_count_block for the query alone.
 *
 * With `select` set the server returns results instead of counter rows:
 * each batch runs the index path's pass gate and FP64 value on the device
 * (kssdhip_select_block_dev with select_opts; the request's qry_ctx is the
 * number of distinct contexts among its keys, ref_ctx the totals given at
 * open), only the passing rows come back, and kssdhip_serve_wait_rows
 * hands each request its rows ranked by value, highest first (ties by
 * ascending reference id), cut to ntop when ntop > 0.  This replaces the
 * per-query selection of the stream path (command_ani.c:2110-2201 rows,
 * sorted and cut at :2634-2650) with the index path's gate
 * (command_ani.h:585-606).  Admission then budgets keys plus one 56-byte
 * row per reference per request.
 */
typedef struct kssdhip_server kssdhip_server;
typedef struct {
	size_t budget_bytes; /* device bytes for in-flight query state, split over the slots */
	int slots;           /* concurrent batches (HIP streams), 1..16 */
	int max_batch;       /* queries per launch */
	int max_queued;      /* host queue bound; submit returns KSSDHIP_EAGAIN beyond it */
	uint32_t flags;      /* KSSDHIP_UNIQUE, KSSDHIP_IGNORECONFLICT */
	int select;          /* nonzero: results (kssdhip_serve_wait_rows), not counter rows */
	kssdhip_select_opts_t select_opts; /* with select: gate and value (ctxcut >= 1) */
	const uint32_t *ref_ctx;           /* with select: ref_gnum context totals, copied at open */
	int ntop;                          /* with select: rows kept per request (0: all) */
} kssdhip_serve_opts_t;
typedef struct {
	double queue_ms;    /* submit -> launch */
	double device_ms;   /* launch -> rows on the host */
	double total_ms;    /* submit -> rows on the host */
	int batch_queries;  /* queries in the launch that carried the request */
} kssdhip_serve_stats_t;
typedef struct {
	int64_t batches, queries, rejected;
	int max_batch_seen;
	size_t slot_bytes;
} kssdhip_serve_info_t;

int kssdhip_serve_open(kssdhip_index *index, const kssdhip_serve_opts_t *opts, kssdhip_server **out);
/* Queue one query sketch (n sorted keys, copied); returns a ticket >= 0 or
 * a negative error (KSSDHIP_EAGAIN when max_queued requests are waiting). */
int64_t kssdhip_serve_submit(kssdhip_server *server, const uint64_t *keys, size_t n);
/* Wait for a ticket (timeout_ms < 0: no limit) and copy its ref_gnum-cell
 * counter rows (unique rows when the server counts them; NULL skips a row).
 * Each ticket is waited for once. */
int kssdhip_serve_wait(kssdhip_server *server, int64_t ticket, kssd_ctx_mut2_t *ctx,
		       kssd_obj_section_t *obj, kssd_ctx_mut2_t *unique_ctx,
		       kssd_obj_section_t *unique_obj, int timeout_ms, kssdhip_serve_stats_t *stats);
/* Results of a server opened with select: the request's passing rows
 * (query 0), ranked; *n_rows receives their number.  When it exceeds cap
 * nothing is copied, KSSDHIP_ENOSPC is returned and the ticket stays
 * waitable, so the caller can retry with a larger buffer. */
int kssdhip_serve_wait_rows(kssdhip_server *server, int64_t ticket, kssdhip_ani_row_t *rows, size_t cap,
			    size_t *n_rows, int timeout_ms, kssdhip_serve_stats_t *stats);
int kssdhip_serve_info(kssdhip_server *server, kssdhip_serve_info_t *info);
/* Stop, drain the launched batches, fail the queued ones, free everything.
 * No call on the server may follow or run concurrently. */
int kssdhip_serve_close(kssdhip_server *server);

#ifdef __cplusplus
}
#endif
#endif
