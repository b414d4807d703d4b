/*
 * Host epilogue of the comparison path: from per-pair counts to the rows and
 * matrices the `ani` command prints.  Counts come from the HIP kernels (or
 * any source with the same layout); everything here is FP64 host code.
 *
 * Row semantics follow make_selected_output_row (command_ani.c:1349-1379)
 * and its callees; the per-path filters and orders follow
 *   index path   count_ctx_obj_frm_comb_sketch_section_with_unique ranking
 *                (command_ani.h:577-621) + ani_block_print_internal
 *                (command_ani.c:3437-3553)
 *   stream path  compute_streamed_ref_row_one_qry (command_ani.c:2110-2201)
 *                + cmp_ani_desc order (command_ani.c:1087-1095)
 *   hash path    comb_manysmall_..._filter_and_sort_survivors
 *                (command_ani.c:3888-4000)
 *   matrices     comb_sortedsketch64_self_matrix, _indexed_self_full,
 *                _indexed_self_triangle, comb_sortedsketch64Xcomb_sortedsketch64
 */
#ifndef KSSD_REPORT_H
#define KSSD_REPORT_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include "kssdhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KSSD_PATHLEN 256

/* lcofiles.infilemeta record, global_basic.h:182-193 */
typedef struct {
	uint64_t total_length_bp;
	uint32_t record_count;
	uint32_t median_length_bp;
	float asm_level;
	float length_cv;
	uint8_t meta_fmt_version;
	int8_t infile_fmt;
	int8_t create_type;
	uint8_t infile_flags;
} kssd_infile_meta_t;

/* The options of ani_opt_t (command_ani.h:18-68) that shape the output. */
typedef struct {
	int fmt;             /* 0 detail, 1 matrix, 2 triangle */
	int s;               /* selected metric, sign = matrix values as ANI */
	bool v;              /* naive model */
	bool unassembled;    /* --qraw */
	bool unified_metric;
	bool ignoreconflict;
	bool raw_output;
	bool max_pdist_set;
	double max_pdist;
	int max_diff_obj_section; /* -1 = off */
	int e;               /* exception value */
	int ntop;            /* <= 0: all */
	int ctxcut;
	float afcut;
	float anicut;
	bool unique_ctx;
	bool d;              /* diagonal in triangle output */
	int drfold;          /* ani_model_drfold */
	unsigned obj_bits;   /* Bitslen.obj */
	unsigned ctx_bits;   /* Bitslen.ctx */
	bool estimate_coverage; /* append the coverage columns */
} kssd_report_opts_t;

/* ani_row_t (command_ani.c:1060-1085), the fields that reach the output */
typedef struct {
	uint32_t rn;
	double af_qry, blastn_af_qry, af_ref, blastn_af_ref;
	double ani, calibrated_ani, best_ani;
	double moe_dist, naive_dist, p_dist, mash_dist, aaf_dist;
	double metric, selected_ani;
	unsigned char confidence;
	unsigned char best_guarded;
	bool af_pass;
	bool unique_ctx_selected;
	int XnY_ctx, N_diff_obj, N_diff_obj_section, N_mut2_ctx;
	/* --estimate-coverage: status 0 not computed, 1 markers unique among the
	 * reported rows, 2 query without counts, 4 markers from the reference's
	 * unique-marker bits */
	unsigned char coverage_status;
	uint32_t cov_markers, cov_observed;
	double cov_observed_fraction, cov_count_sum, est_depth, pos_mean_depth, pos_median_depth, count_cv,
		abundance_mass, abundance_fraction;
} kssd_row_t;

/* Load the calibrated model's tables (idempotent, thread-safe); 0 or -1. */
int kssd_report_init(void);
/* calibrated ANI model (refaf_hgb_predict_ani, model_refaf_hgb.h:7593-7628) */
double kssd_refaf_predict_ani(double raw_ani, double ref_af, unsigned xny, unsigned n_diff_obj,
			      unsigned n_diff_obj_section, unsigned n_mut2_ctx, int drfold);

/* make_selected_output_row */
kssd_row_t kssd_make_row(uint32_t rn, uint32_t xny, uint32_t dobj, uint32_t dsec, uint32_t mut2,
			 const kssd_report_opts_t *o, uint32_t qry_ctx, uint32_t ref_ctx,
			 double af_qry, double blastn_af_qry, double af_ref, double blastn_af_ref,
			 const kssd_infile_meta_t *qry_asm, const kssd_infile_meta_t *ref_asm);
/* selected metric after a counter swap (unique-context rows, command_ani.c:3503-3510) */
void kssd_finalize_row(kssd_row_t *r, const kssd_report_opts_t *o);
/* lm3ways_af_ANIb_from_features with X_ctx */
double kssd_blastn_af(uint32_t xny, uint32_t dobj, uint32_t dsec, uint32_t mut2, uint32_t x_ctx,
		      int drfold);
double kssd_naive_distance(uint32_t xny, uint32_t dobj, uint32_t dsec);

/* Append the detail row (append_unified_detail_row, command_ani.c:1381-1432)
 * to a growable buffer; returns 0 or -1 (allocation). */
typedef struct {
	char *s;
	size_t l, m;
} kssd_buf_t;
int kssd_buf_printf(kssd_buf_t *b, const char *fmt, ...);
void kssd_buf_free(kssd_buf_t *b);
void kssd_append_detail_row(kssd_buf_t *b, const kssd_report_opts_t *o, const char *qry_name,
			    const char *ref_name, const kssd_row_t *r, const char *annotation);
extern const char kssd_detail_header[];

extern const char kssd_coverage_header[];

/* What --estimate-coverage reads for one query (command_ani.c:2203-2455):
 * the query's keys and their read counts (NULL: the query has no
 * comblco.a), every reference sample's keys, and optionally the reference's
 * sorted index with its unique-marker bits. */
typedef struct {
	const uint64_t *qkeys;
	const uint32_t *qcounts;
	size_t qn;
	const uint64_t *ref_keys, *ref_index;
	const kssd_ctxgidobj_t *ref_records; /* NULL: markers from the reported rows */
	const uint8_t *ref_unique_bits;
	size_t ref_nrec;
	unsigned obj_bits;
} kssd_coverage_src_t;
/* populate_selected_coverage_for_query_arrays (command_ani.c:2318-2455):
 * a reported row's markers are its reference's keys that are unique — by
 * the unique-marker bits when given, else among the keys of all n reported
 * rows — and its depth figures come from the query's counts of the markers
 * it holds; abundance fractions are shares of the rows' summed counts. */
void kssd_populate_coverage(kssd_row_t *rows, size_t n, const kssd_coverage_src_t *src);

/* Rows with qry/ref context totals and counts for one query against all
 * refs -- or, when ids is not NULL, against the ncand references ids[0 ..
 * ncand) (ascending) only, ctx/obj/uctx/uobj then holding ncand cells (the
 * pairs a device-side gate let through; every other pair reads as zero). */
typedef struct {
	const kssd_ctx_mut2_t *ctx;  /* [nref] */
	const kssd_obj_section_t *obj;
	const kssd_ctx_mut2_t *uctx; /* optional (unique-context) */
	const kssd_obj_section_t *uobj;
	uint32_t qry_ctx;
	const uint32_t *ref_ctx;     /* [nref] */
	int nref;
	const kssd_infile_meta_t *qry_meta;  /* may be NULL */
	const kssd_infile_meta_t *ref_meta;  /* [nref] or NULL */
	const kssd_coverage_src_t *cov;      /* --estimate-coverage inputs (NULL: not computed) */
	const uint32_t *ids;                 /* candidate references (index path only) or NULL */
	int ncand;
	/* one-query stream path: [nref], 1 when the last context the reference's
	 * walk counts is one of the query's (NULL: treated as 0) */
	const uint8_t *last_shared;
} kssd_query_counts_t;

/* Index path (mem_eff) detail rows of one query, ranked and filtered as the
 * reference does; appends the printed rows to out.  With c->ids the pass
 * gate is re-evaluated on the candidates only; that gives the reference's
 * rows when no other pair could pass it (ctxcut >= 1: a pair with no shared
 * context is never a candidate). */
void kssd_emit_index_detail(kssd_buf_t *out, const kssd_report_opts_t *o, const kssd_query_counts_t *c,
			    const char *qry_name, char (*ref_names)[KSSD_PATHLEN],
			    char (*ref_anno)[KSSD_PATHLEN]);
/* Index path (mem_eff) matrix row: "name\t%lf..." with float-cast AFs. */
void kssd_emit_index_matrix_row(kssd_buf_t *out, const kssd_report_opts_t *o,
				const kssd_query_counts_t *c, const char *qry_name);
/* Stream paths: one query (single != 0, stream_ref_sketches_one_qraw_lookup,
 * with its AF and close-filter early exits) or a few queries
 * (stream_ref_sketches_multi_qraw_sortedindex) against many references. */
void kssd_emit_stream_detail(kssd_buf_t *out, const kssd_report_opts_t *o, const kssd_query_counts_t *c,
			     const char *qry_name, char (*ref_names)[KSSD_PATHLEN],
			     char (*ref_anno)[KSSD_PATHLEN], int single);
/* Hash path (many small references, no index). */
void kssd_emit_hash_detail(kssd_buf_t *out, const kssd_report_opts_t *o, const kssd_query_counts_t *c,
			   const char *qry_name, char (*ref_names)[KSSD_PATHLEN],
			   char (*ref_anno)[KSSD_PATHLEN]);

/* Matrix cell (ani_matrix_pair_value / ani_indexed_self_pair_metric). */
double kssd_matrix_value(const kssd_report_opts_t *o, uint32_t xny, uint32_t dobj, uint32_t dsec,
			 uint32_t mut2, uint32_t qry_ctx, uint32_t ref_ctx,
			 const kssd_infile_meta_t *qry_meta, const kssd_infile_meta_t *ref_meta);
double kssd_matrix_exception(const kssd_report_opts_t *o);
double kssd_matrix_diagonal(const kssd_report_opts_t *o);
/* ani_scale_fixed6 / ani_kput_scaled6 (command_ani.c:1593-1625) */
int32_t kssd_scale_fixed6(double v);
void kssd_put_fixed6(kssd_buf_t *b, int32_t scaled);

/* ---- the `matrix` command's values and cells (command_matrix.c) ---- */
enum { KSSD_MM_NAIVE = 0, KSSD_MM_MOE, KSSD_MM_PDIST, KSSD_MM_MASH, KSSD_MM_AAF };
/* One metric name as pairwise_metric_from_string (pairwise_graph.c:69-98)
 * accepts it; -1 when unknown. */
int kssd_matrix_metric(const char *name);
/* Indexed self matrices: matrix_context_feature_value for one metric
 * (command_matrix.c:400-419, :621-661): e when nothing is shared or a
 * total is 0 or the value is not finite, else clamped to [0, 1]. */
double kssd_matrix_context_value(int metric, double e, uint32_t xny, uint32_t dobj, uint32_t dsec,
				 uint32_t mut2, uint32_t qry_ctx, uint32_t ref_ctx, int drfold,
				 unsigned obj_bits, unsigned ctx_bits);
/* Dense reports: pairwise_eval_arrays' distance (pairwise_graph.c:301-360),
 * or e when the pair is not valid (an empty sample or nothing shared). */
double kssd_pairwise_value(int metric, double e, size_t qry_len, size_t ref_len, uint32_t xny,
			   uint32_t dobj, uint32_t dsec, uint32_t mut2, uint32_t qry_ctx, uint32_t ref_ctx,
			   int drfold, unsigned obj_bits, unsigned ctx_bits);
/* ---- sparse edge reports (matrix --format edges) ---- */
/* pairwise_eval_t (pairwise_graph.h): one pair's evaluation */
typedef struct {
	double distance, similarity, af_qry, af_ref, max_af;
	uint32_t xny, dobj, dsec, mut2;
	int valid;
} kssd_pair_eval_t;
#define KSSD_EXPR_MAX 4
enum { KSSD_EXPR_SINGLE = 0, KSSD_EXPR_AND, KSSD_EXPR_OR };
/* pairwise_metric_expr_t: up to 4 metrics joined by & (all valid, the
 * largest distance) or | (the smallest valid distance), with its label */
typedef struct {
	int metrics[KSSD_EXPR_MAX];
	int count, op;
	char label[96];
} kssd_metric_expr_t;
/* kssd_matrix_context_value over a metric expression (matrix_context_feature_value,
 * command_matrix.c:621-661): e when nothing is shared, a total is 0 or a
 * metric is not finite; else the largest (&) or smallest (|) clamped value. */
double kssd_matrix_context_expr_value(const kssd_metric_expr_t *x, double e, uint32_t xny, uint32_t dobj,
				      uint32_t dsec, uint32_t mut2, uint32_t qc, uint32_t rc, int drfold,
				      unsigned obj_bits, unsigned ctx_bits);
/* pairwise_metric_name (pairwise_graph.c:52-67) */
const char *kssd_metric_name(int metric);
/* pairwise_metric_expr_single / pairwise_metric_expr_from_string
 * (pairwise_graph.c:153-234); the latter returns 0 on success, -1 when
 * the text is not an expression. */
kssd_metric_expr_t kssd_metric_expr_single(int metric);
int kssd_metric_expr_parse(const char *text, kssd_metric_expr_t *out);
/* pairwise_eval_expr_samples for lco sketches (pairwise_graph.c:301-360,
 * :420-513) from the pair's counts: qry_len/ref_len the samples' entries,
 * qry_ctx/ref_ctx their context totals. */
kssd_pair_eval_t kssd_pair_eval(const kssd_metric_expr_t *e, size_t qry_len, size_t ref_len, uint32_t xny,
				uint32_t dobj, uint32_t dsec, uint32_t mut2, uint32_t qry_ctx, uint32_t ref_ctx,
				int drfold, unsigned obj_bits, unsigned ctx_bits);
/* pairwise_eval_passes_edge (pairwise_graph.c:518-530) */
int kssd_pair_passes_edge(const kssd_pair_eval_t *ev, double cut, uint32_t ctxcut, double max_afcut);
/* pairwise_quality_compare (pairwise_graph.c:532-603): > 0 when sample a
 * is the better cluster representative, < 0 when b is; meta may be NULL
 * (then only sketch entries and the index decide); index: comblco.index. */
int kssd_quality_compare(const kssd_infile_meta_t *meta, const uint64_t *index, int a, int b);
/* matrix_print_edge_header / matrix_print_edge (command_matrix.c:87-101) */
void kssd_put_edge_header(kssd_buf_t *b);
void kssd_put_edge(kssd_buf_t *b, const char *qry_name, const char *ref_name, const kssd_pair_eval_t *ev,
		   const char *label);

/* matrix_kput_fixed6 / matrix_scale_fixed6 / matrix_kput_scaled6
 * (command_matrix.c:103-149) */
void kssd_matrix_put_fixed6(kssd_buf_t *b, double v);
uint32_t kssd_matrix_scale_fixed6(double v);
void kssd_matrix_put_scaled6(kssd_buf_t *b, uint32_t scaled);

#ifdef __cplusplus
}
#endif
#endif
