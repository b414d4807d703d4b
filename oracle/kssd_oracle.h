/*
 * CPU oracle for the KSSD3A sketch-comparison path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C + OpenMP restatement of the
 * reference algorithm, written from the reference sources cited beside each
 * function.  Only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it, and only as the checker or as the timed CPU
 * baseline; the product path never calls it.
 *
 * Every function here is pinned by at least one check in tests/ that does not
 * retype its formula: the reference binary's own outputs (golden TSVs built
 * by oracle/_ref from /root/reference), the brute-force closed form of
 * tests/test_unique_context.py:99-104 (min over object pairs of per-base
 * mismatches), exhaustive enumeration for dna_popcount, and the identity /
 * empty special cases.  See the test files named in each comment.
 */
#ifndef KSSD_ORACLE_H
#define KSSD_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 12-byte packed index record, sketch_rearrange.h:19 */
typedef struct __attribute__((packed)) {
	uint64_t ctxgid; /* ctx << 20 | gid */
	uint32_t obj;
} orc_rec_t;

/* command_ani.h:114-124 */
typedef struct { uint32_t num_ctx, num_mut2_ctx; } orc_ctx_t;
typedef struct { uint32_t diff_obj, diff_obj_section; } orc_obj_t;

/* model_ani.h:215-222 (field order kept) */
typedef struct {
	uint32_t XnY_ctx, X_ctx, N_diff_obj_section, N_mut2_ctx, N_diff_obj;
} orc_features_t;

#define ORC_LOWER 1u
#define ORC_UNIQUE 2u
#define ORC_IGNORECONFLICT 4u

/* dna_popcount.h:45-49 */
int orc_dna_popcount(uint32_t x);
void orc_dna_popcount_many(const uint32_t *x, size_t n, uint8_t *out);

/* command_ani.c:3656-3678 (F has 2^k + 1 entries) */
int orc_fenceposts_build(const orc_rec_t *b, size_t n, int k, size_t *F);
/* command_ani.c:3703-3738; idx[i] = leftmost record with equal ctx or SIZE_MAX */
void orc_find_first(const uint64_t *a, size_t na, const orc_rec_t *b, size_t nb,
		    const size_t *F, int k, unsigned obj_bits, size_t *idx);

/*
 * Index-join counting of one query block against a sorted index.
 * Restates count_ctx_obj_frm_comb_sketch_section_with_unique
 * (command_ani.h:463-575, counting part) and
 * count_ctx_obj_frm_comb_sketch_section_lower (command_ani.h:635-721).
 * section_skidx holds absolute offsets (nq+1 values); keys start at
 * section_skidx[0].  Outputs are dense [nq x ref_gnum] and are ACCUMULATED
 * into (the caller zeroes them), exactly as the reference does.
 * uctx/uobj are only touched with ORC_UNIQUE (never with ORC_LOWER).
 */
void orc_count_block(const orc_rec_t *index, size_t n_index, int ref_gnum,
		     unsigned obj_bits, const uint64_t *section_sk,
		     const uint64_t *section_skidx, int nq, int qry_gid_offset,
		     unsigned flags, orc_ctx_t *ctx, orc_obj_t *obj,
		     orc_ctx_t *uctx, orc_obj_t *uobj, int nthreads);

/* get_ani_features_ctx_min_over_conflicts_both_filtered, command_ani.h:283-346 */
void orc_pair_features(const uint64_t *a, size_t n, const uint64_t *b, size_t m,
		       unsigned obj_bits, int ignore_b_conflict, orc_features_t *out);

/*
 * Index build: ctxobj64_2ctxgidobj (sketch_rearrange.c:65-78) followed by a
 * sort on (ctxgid, obj) (kssdlib_sort.c:145-148).  out has skidx[ng] records.
 */
void orc_build_index(const uint64_t *keys, const uint64_t *skidx, int ngenomes,
		     unsigned obj_bits, orc_rec_t *out);

/* count_ctx_runs_sorted_ctxobj64_local, command_ani.c:337-349 */
uint32_t orc_ctx_runs(const uint64_t *a, size_t n, unsigned obj_bits, int ignoreconflict);
/* ref_ctx_counts_from_ctxgidobj, command_ani.c:351-370 */
void orc_ref_ctx_counts_from_index(const orc_rec_t *index, size_t n, int ref_gnum,
				   const uint64_t *ref_skidx, int ignoreconflict,
				   uint32_t *counts);

/* FP64 formulas */
double orc_naive_dist(const orc_features_t *f);                     /* command_ani.h:202-224 */
double orc_p_dist(const orc_features_t *f, unsigned obj_bits);       /* command_ani.h:226-232 */
double orc_lm3ways_dist(const orc_features_t *f, int drfold);        /* model_ani.h:285-305 */
double orc_af_anib(const orc_features_t *f, int drfold);             /* model_ani.h:356-367 */
double orc_mashD(int k, uint32_t ref_ctx, uint32_t qry_ctx, uint32_t xny); /* command_matrix.c:26-29 */
double orc_aafD(int k, uint32_t ref_ctx, uint32_t qry_ctx, uint32_t xny);  /* command_matrix.c:30-33 */

/* Dense matrix values for metrics 3..9 (command_ani.c:1321-1379, :1458-1496). */
void orc_matrix_values(const orc_ctx_t *ctx, const orc_obj_t *obj, int nq, int nr,
		       const uint32_t *qctx, const uint32_t *rctx, int metric, int values_ani,
		       double afcut, double e, int drfold, unsigned obj_bits, unsigned ctx_bits,
		       double *out, int nthreads);

/*
 * The pass gate and ranking of the index path's block function for fmt == 0
 * (command_ani.h:577-621), on dense counts [nq x R] (uctx/uobj with
 * unique_ctx).  Per query row i: pass_ids[i*R + k] / rank[i*R + k] are the
 * k-th passing reference in ascending id order and its ranking value
 * (k < n_pass[i]), and
 * ids[i*R + r] the printed order after the top-N cut (r < n_out[i]).
 * ctxcut_opt / afcut / anicut / ntop_opt as ani_opt_t holds them.
 */
void orc_index_rank(const orc_ctx_t *ctx, const orc_obj_t *obj, const orc_ctx_t *uctx, const orc_obj_t *uobj,
		    int nq, int R, const uint32_t *ref_ctx, const uint32_t *qry_ctx, int ctxcut_opt, float afcut,
		    float anicut, int ntop_opt, int unique_ctx, uint32_t *ids, uint32_t *pass_ids, double *rank,
		    uint32_t *n_pass, uint32_t *n_out);

/*
 * The reference's default path for one query (stream_ref_sketches_one_qraw_lookup,
 * command_ani.c:2490-2666; compute_streamed_ref_row_one_qry, :2110-2201): the
 * query's context runs in an open-addressed table, every reference streamed run
 * by run against it.  Per reference r: ctx[r] / obj[r] the pair's counters (as
 * a row of orc_count_block; a reference given up early by the AF-needed bound
 * keeps the counts up to that point), ref_ctx[r] the contexts walked, pass[r]
 * whether the row survives ctxcut, the AF-needed count and afcut (the
 * selected-ANI cut and the close filters are left out).  qry_ctx: the query's
 * context runs.
 */
int orc_stream_one_query(const uint64_t *qry, size_t qlen, const uint64_t *ref_keys, const uint64_t *ref_idx, int R,
			 unsigned obj_bits, int ignoreconflict, int ref_conflict, int ctxcut, float afcut,
			 orc_ctx_t *ctx, orc_obj_t *obj, uint32_t *ref_ctx, uint8_t *pass, uint32_t *qry_ctx,
			 int nthreads);

#ifdef __cplusplus
}
#endif
#endif
