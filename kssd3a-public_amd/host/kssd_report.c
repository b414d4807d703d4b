/*
 * Host epilogue: rows, ranking, matrices and TSV formatting.  See
 * kssd_report.h for the reference functions each part follows.
 */
#define _GNU_SOURCE
#include "kssd_report.h"

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "kssd_models.h"

/* ------------------------------------------------------------ buffers */
int kssd_buf_printf(kssd_buf_t *b, const char *fmt, ...)
{
	for (;;) {
		va_list ap;
		va_start(ap, fmt);
		const size_t room = b->m - b->l;
		const int n = vsnprintf(b->s ? b->s + b->l : NULL, room, fmt, ap);
		va_end(ap);
		if (n < 0)
			return -1;
		if ((size_t)n < room) {
			b->l += (size_t)n;
			return 0;
		}
		size_t m = b->m ? b->m : 256;
		while (m - b->l <= (size_t)n)
			m <<= 1;
		char *s = realloc(b->s, m);
		if (!s)
			return -1;
		b->s = s;
		b->m = m;
	}
}

void kssd_buf_free(kssd_buf_t *b)
{
	free(b->s);
	b->s = NULL;
	b->l = b->m = 0;
}

/* ------------------------------------------------------------ calibrated model */
extern const unsigned char kssd_refaf_blob[];
extern const unsigned char kssd_refaf_blob_end[];

typedef struct {
	unsigned n_features, n_trees, n_nodes;
	double baseline;
	const uint32_t *offsets;
	const double *threshold, *value;
	const uint16_t *left, *right;
	const int8_t *feature;
	const uint8_t *is_leaf, *missing_left;
} refaf_model_t;

static refaf_model_t g_model;
static int g_model_ok;
static pthread_once_t g_model_once = PTHREAD_ONCE_INIT;

static void refaf_parse(void)
{
	const unsigned char *p = kssd_refaf_blob;
	const size_t size = (size_t)(kssd_refaf_blob_end - kssd_refaf_blob);
	g_model_ok = 0;
	if (size < 32 || memcmp(p, "KSSDHGB1", 8) != 0)
		return;
	uint32_t hdr[4];
	memcpy(hdr, p + 8, sizeof hdr);
	refaf_model_t m;
	m.n_features = hdr[0];
	m.n_trees = hdr[1];
	m.n_nodes = hdr[2];
	memcpy(&m.baseline, p + 24, 8);
	size_t off = 32;
	const size_t need = off + 4 * ((size_t)m.n_trees + 1) + (size_t)m.n_nodes * (8 + 8 + 2 + 2 + 1 + 1 + 1);
	if (m.n_features != 11 || size < need)
		return;
	/* the blob is 8-byte aligned and laid out so every array is naturally aligned */
	m.offsets = (const uint32_t *)(p + off);
	off += 4 * ((size_t)m.n_trees + 1);
	off = (off + 7) & ~(size_t)7;
	m.threshold = (const double *)(p + off);
	off += 8 * (size_t)m.n_nodes;
	m.value = (const double *)(p + off);
	off += 8 * (size_t)m.n_nodes;
	m.left = (const uint16_t *)(p + off);
	off += 2 * (size_t)m.n_nodes;
	m.right = (const uint16_t *)(p + off);
	off += 2 * (size_t)m.n_nodes;
	m.feature = (const int8_t *)(p + off);
	off += m.n_nodes;
	m.is_leaf = (const uint8_t *)(p + off);
	off += m.n_nodes;
	m.missing_left = (const uint8_t *)(p + off);
	g_model = m;
	g_model_ok = 1;
}

static int refaf_load(void)
{
	pthread_once(&g_model_once, refaf_parse);
	return g_model_ok;
}

int kssd_report_init(void)
{
	return refaf_load() ? 0 : -1;
}

static inline double clamp01(double x)
{
	return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
}

double kssd_refaf_predict_ani(double raw_ani, double ref_af, unsigned xny, unsigned n_diff_obj,
			      unsigned n_diff_obj_section, unsigned n_mut2_ctx, int drfold)
{
	/* model_refaf_hgb.h:7593-7628 */
	if (xny == 0)
		return raw_ani;
	if (!refaf_load())
		return raw_ani;
	raw_ani = clamp01(raw_ani);
	if (ref_af < 1e-12)
		ref_af = 1e-12;
	if (ref_af > 1.0)
		ref_af = 1.0;
	const double xny_d = (double)xny;
	const double ref_aaf_ani = clamp01(1.0 + log(ref_af) / 11.0);
	double x[11];
	x[0] = raw_ani;
	x[1] = ref_af;
	x[2] = log1p(xny_d * kssd_fold_scale(drfold));
	x[3] = (double)n_diff_obj / xny_d;
	x[4] = (double)n_diff_obj_section / xny_d;
	x[5] = (double)n_mut2_ctx / xny_d;
	x[6] = ref_aaf_ani;
	x[7] = raw_ani - ref_aaf_ani;
	x[8] = ref_af < 0.2 ? 1.0 : 0.0;
	x[9] = (ref_af >= 0.2 && ref_af < 0.5) ? 1.0 : 0.0;
	x[10] = ref_af >= 0.5 ? 1.0 : 0.0;
	/* model_refaf_hgb.h:7575-7591 */
	const refaf_model_t *m = &g_model;
	double pred = m->baseline;
	for (unsigned t = 0; t < m->n_trees; ++t) {
		const unsigned begin = m->offsets[t];
		unsigned node = begin;
		while (!m->is_leaf[node]) {
			const double v = x[m->feature[node]];
			const int go_left = isnan(v) ? m->missing_left[node] : (v <= m->threshold[node]);
			node = begin + (go_left ? m->left[node] : m->right[node]);
		}
		pred += m->value[node];
	}
	return clamp01(pred);
}

/* ------------------------------------------------------------ rows */
enum { CONF_LOW = 0, CONF_MEDIUM = 1, CONF_HIGH = 2 };

static const char *const metric_names[9] = {"BestDist", "RecalDist", "CtxMoE", "Naive", "MashD",
					    "AafD", "MashD_if_far", "AafD_if_far", "p_dist"};

const char kssd_detail_header[] =
	"Qry\tRef\tANI\tDistance\tConfidence\tSelected_metric\tXnY_ctx\tQry_align_fraction\t"
	"blastn_Qry_align_fraction\tRef_align_fraction\tblastn_Ref_align_fraction\tN_diff_obj\t"
	"N_diff_obj_section\tN_mut2_ctx\tRef_annotation";
const char kssd_coverage_header[] =
	"Coverage_estimator\tCoverage_ref_markers\tCoverage_observed_markers\tCoverage_observed_fraction\t"
	"Coverage_count_sum\tEstimated_depth\tPositive_mean_depth\tPositive_median_depth\tCount_CV\t"
	"Estimated_abundance_mass\tEstimated_abundance_fraction\tAbundance_fraction_scope";

static inline bool af_pass(const kssd_report_opts_t *o, double afq, double afr)
{
	/* ani_report_af_pass, command_ani.h:70-78 */
	return (afq > afr ? afq : afr) >= (double)o->afcut;
}

static inline uint32_t af_needed_ctx(const kssd_report_opts_t *o, uint32_t qc, uint32_t rc)
{
	/* ani_report_af_needed_ctx, command_ani.h:80-90 */
	const uint32_t denom = qc < rc ? qc : rc;
	const double need = (double)o->afcut * (double)denom;
	if (need <= 0.0)
		return 0;
	if (need >= (double)UINT32_MAX)
		return UINT32_MAX;
	return (uint32_t)ceil(need);
}

static inline unsigned char confidence_from(double raw, double cal, double ref_af)
{
	/* ani_confidence_from_values, command_ani.c:100-108 */
	const double drop = raw - cal;
	if (ref_af < 0.5 || (raw >= 0.95 && drop > 0.02))
		return CONF_LOW;
	if (fabs(drop) > 0.01)
		return CONF_MEDIUM;
	return CONF_HIGH;
}

static inline double bounded_ani(double v)
{
	if (!isfinite(v))
		return NAN;
	return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

static inline double ctx_k(const kssd_report_opts_t *o)
{
	return o->ctx_bits > 0 ? (double)o->ctx_bits / 2.0 : 22.0;
}

static double mash_ani(const kssd_report_opts_t *o, double overlap, double q, double r)
{
	/* mash_ani_from_counts, command_ani.c:164-175 */
	if (overlap <= 0.0 || q <= 0.0 || r <= 0.0)
		return 0.0;
	const double denom = q + r - overlap;
	if (denom <= 0.0)
		return 0.0;
	const double j = overlap / denom;
	if (j <= 0.0)
		return 0.0;
	return bounded_ani(1.0 + log(2.0 * j / (1.0 + j)) / ctx_k(o));
}

static double aaf_ani(const kssd_report_opts_t *o, double overlap, double q, double r)
{
	/* aaf_ani_from_counts, command_ani.c:177-188 */
	if (overlap <= 0.0 || q <= 0.0 || r <= 0.0)
		return 0.0;
	const double mn = q < r ? q : r;
	if (mn <= 0.0)
		return 0.0;
	const double c = overlap / mn;
	if (c <= 0.0)
		return 0.0;
	return bounded_ani(1.0 + log(c) / ctx_k(o));
}

static bool meta_like_assembly(const kssd_infile_meta_t *s)
{
	/* infile_meta_complete_like_assembly, command_ani.c:118-128 */
	if (!s || s->meta_fmt_version != 1 || s->infile_fmt != 1 || s->total_length_bp == 0 ||
	    s->median_length_bp == 0)
		return false;
	return s->total_length_bp >= 3500000ULL && s->record_count < 5 && s->median_length_bp > 100000U &&
	       (double)s->total_length_bp / (double)s->median_length_bp < 5.0;
}

static bool meta_has_assembly(const kssd_infile_meta_t *s)
{
	return s && s->meta_fmt_version == 1 && s->infile_fmt == 1 && s->total_length_bp > 0 &&
	       s->record_count > 0;
}

static inline bool best_guard_enabled(const kssd_report_opts_t *o)
{
	return o->fmt == 0 && !o->unassembled && !o->v;
}

static void fill_calibration(kssd_row_t *r, const kssd_report_opts_t *o, const kssd_infile_meta_t *qa)
{
	/* fill_row_calibration, command_ani.c:1129-1184 */
	if (o->unassembled) {
		r->calibrated_ani = r->ani;
		r->best_ani = r->ani;
		r->confidence = CONF_HIGH;
		r->best_guarded = 0;
		return;
	}
	if (r->XnY_ctx > 0 && r->N_diff_obj == 0 && r->N_diff_obj_section == 0 && r->N_mut2_ctx == 0 &&
	    r->af_qry == 1.0 && r->af_ref == 1.0) {
		r->calibrated_ani = 1.0;
		r->best_ani = 1.0;
		r->confidence = CONF_HIGH;
		r->best_guarded = 0;
		return;
	}
	r->calibrated_ani = kssd_refaf_predict_ani(r->ani, r->af_ref, (unsigned)r->XnY_ctx,
						   (unsigned)r->N_diff_obj, (unsigned)r->N_diff_obj_section,
						   (unsigned)r->N_mut2_ctx, o->drfold);
	r->confidence = confidence_from(r->ani, r->calibrated_ani, r->af_ref);
	r->best_ani = r->calibrated_ani;
	r->best_guarded = 0;
	const double min_af = r->af_qry < r->af_ref ? r->af_qry : r->af_ref;
	if (!best_guard_enabled(o) || !meta_has_assembly(qa) || !meta_like_assembly(qa) || r->ani < 0.958 ||
	    r->calibrated_ani < 0.962 || min_af < 0.25 || min_af >= 0.50 || r->XnY_ctx <= 0 ||
	    r->af_qry <= 0.0 || r->af_ref <= 0.0)
		return;
	const double qc = (double)r->XnY_ctx / r->af_qry;
	const double rc = (double)r->XnY_ctx / r->af_ref;
	double exact = (double)r->XnY_ctx - (double)r->N_diff_obj;
	if (exact < 0.0)
		exact = 0.0;
	const double mean = (mash_ani(o, (double)r->XnY_ctx, qc, rc) + aaf_ani(o, (double)r->XnY_ctx, qc, rc) +
			     mash_ani(o, exact, qc, rc) + aaf_ani(o, exact, qc, rc)) / 4.0;
	if (r->ani - mean < 0.015)
		return;
	r->best_ani = aaf_ani(o, exact, qc, rc);
	r->best_guarded = 1;
}

static inline int request_id(const kssd_report_opts_t *o)
{
	int s = abs(o->s);
	return (s < 1 || s > 9) ? 1 : s;
}

static inline bool value_available(double v) { return v > 0.0 && v <= 1.0; }

static int effective_id(const kssd_row_t *r, const kssd_report_opts_t *o)
{
	/* selected_metric_effective_id, command_ani.c:1260-1270 */
	int s = request_id(o);
	if (o->unassembled && !o->unified_metric && s <= 4)
		return 4;
	if (s == 1 && !(!o->raw_output && !o->unassembled && value_available(r->best_ani)))
		s = 2;
	if (s == 2 && !(!o->raw_output && !o->unassembled && value_available(r->calibrated_ani)))
		s = 3;
	return s;
}

static double selected_distance(const kssd_row_t *r, const kssd_report_opts_t *o)
{
	/* selected_metric_distance_from_row, command_ani.c:1272-1296 */
	switch (effective_id(r, o)) {
	case 1: return 1.0 - r->best_ani;
	case 2: return 1.0 - r->calibrated_ani;
	case 3: return r->moe_dist;
	case 4: return r->naive_dist;
	case 5: return r->mash_dist;
	case 6: return r->aaf_dist;
	case 7: return r->af_pass ? r->moe_dist : r->mash_dist;
	case 8: return r->af_pass ? r->moe_dist : r->aaf_dist;
	case 9: return r->p_dist;
	default: return r->metric;
	}
}

void kssd_finalize_row(kssd_row_t *r, const kssd_report_opts_t *o)
{
	r->metric = selected_distance(r, o);
	r->selected_ani = 1.0 - r->metric;
}

double kssd_blastn_af(uint32_t xny, uint32_t dobj, uint32_t dsec, uint32_t mut2, uint32_t x_ctx, int drfold)
{
	return kssd_af_anib(xny, dobj, dsec, mut2, x_ctx, drfold);
}

double kssd_naive_distance(uint32_t xny, uint32_t dobj, uint32_t dsec)
{
	return kssd_naive_dist(xny, dobj, dsec);
}

kssd_row_t kssd_make_row(uint32_t rn, uint32_t xny, uint32_t dobj, uint32_t dsec, uint32_t mut2,
			 const kssd_report_opts_t *o, uint32_t qry_ctx, uint32_t ref_ctx,
			 double af_qry, double blastn_af_qry, double af_ref, double blastn_af_ref,
			 const kssd_infile_meta_t *qry_asm, const kssd_infile_meta_t *ref_asm)
{
	/* make_selected_output_row, command_ani.c:1349-1379, with
	 * fill_row_base_distances :1321-1341 */
	(void)ref_asm;
	kssd_row_t r;
	memset(&r, 0, sizeof r);
	r.rn = rn;
	r.af_qry = af_qry;
	r.blastn_af_qry = blastn_af_qry;
	r.af_ref = af_ref;
	r.blastn_af_ref = blastn_af_ref;
	r.XnY_ctx = (int)xny;
	r.N_diff_obj = (int)dobj;
	r.N_diff_obj_section = (int)dsec;
	r.N_mut2_ctx = (int)mut2;
	r.af_pass = af_pass(o, af_qry, af_ref);
	if (r.af_pass) {
		r.moe_dist = kssd_moe_dist(xny, dobj, dsec, mut2, o->drfold);
		r.naive_dist = kssd_naive_dist(xny, dobj, dsec);
		r.p_dist = kssd_p_dist(xny, dsec, o->obj_bits);
	} else {
		r.moe_dist = o->e;
		r.naive_dist = o->e;
		r.p_dist = o->e;
	}
	r.mash_dist = kssd_mashD(o->ctx_bits / 2, ref_ctx, qry_ctx, xny);
	r.aaf_dist = kssd_aafD(o->ctx_bits / 2, ref_ctx, qry_ctx, xny);
	r.ani = 1.0 - (o->v ? r.naive_dist : r.moe_dist);
	if (!o->raw_output && r.af_pass)
		fill_calibration(&r, o, qry_asm);
	kssd_finalize_row(&r, o);
	return r;
}

static const char *confidence_label(unsigned char c)
{
	return c == CONF_HIGH ? "high" : (c == CONF_MEDIUM ? "medium" : "low");
}

static const char *row_confidence(const kssd_row_t *r, const kssd_report_opts_t *o)
{
	/* selected_metric_confidence_label, command_ani.c:1304-1319 */
	if (o->raw_output)
		return "raw";
	if (o->unassembled && !o->unified_metric && request_id(o) <= 4)
		return "unassembled";
	switch (effective_id(r, o)) {
	case 1: return r->best_guarded ? "guarded_low_confidence" : confidence_label(r->confidence);
	case 2: return confidence_label(r->confidence);
	default: return "raw";
	}
}

static bool passes_close_filter(const kssd_row_t *r, const kssd_report_opts_t *o)
{
	/* ani_row_passes_close_filter, command_ani.c:1232-1246 */
	if (!o->max_pdist_set && o->max_diff_obj_section < 0)
		return true;
	if (o->max_diff_obj_section >= 0 && r->N_diff_obj_section > o->max_diff_obj_section)
		return false;
	if (o->max_pdist_set && (!isfinite(r->p_dist) || r->p_dist > o->max_pdist))
		return false;
	return true;
}

static bool close_filter_impossible(const kssd_report_opts_t *o, uint32_t dsec, uint32_t max_shared)
{
	/* ani_close_filter_impossible, command_ani.c:1209-1230, evaluated on the
	 * final counts */
	if (!o->max_pdist_set && o->max_diff_obj_section < 0)
		return false;
	double limit = INFINITY;
	if (o->max_diff_obj_section >= 0)
		limit = (double)o->max_diff_obj_section;
	if (o->max_pdist_set) {
		const double pos = o->obj_bits > 0 ? (double)o->obj_bits / 2.0 : 0.0;
		if (pos <= 0.0)
			return dsec > 0;
		const double pl = o->max_pdist * (double)max_shared * pos;
		if (pl < limit)
			limit = pl;
	}
	return (double)dsec > floor(limit + 1e-12);
}

static inline double bounded_af(double v)
{
	if (!isfinite(v))
		return v;
	return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

/* append_coverage_double, command_ani.c:1118-1126 */
static void put_cov(kssd_buf_t *b, double v)
{
	if (!isfinite(v))
		kssd_buf_printf(b, "\tNA");
	else if (fabs(v) >= 1000.0)
		kssd_buf_printf(b, "\t%.3f", v);
	else
		kssd_buf_printf(b, "\t%.6g", v);
}

static const char *coverage_label(unsigned char status) /* coverage_estimator_label, :1106-1116 */
{
	switch (status) {
	case 0: return "NA:not_computed";
	case 2: return "NA:query_missing_comblco.a";
	case 3: return "NA:ref_read_error";
	case 4: return "refdb_unique_ctxobj_bitset";
	default: return "reported_unique_ctxobj_sum";
	}
}

void kssd_append_detail_row(kssd_buf_t *b, const kssd_report_opts_t *o, const char *qry_name,
			    const char *ref_name, const kssd_row_t *r, const char *annotation)
{
	/* append_unified_detail_row, command_ani.c:1381-1432 */
	if (!passes_close_filter(r, o))
		return;
	const char *metric = o->unique_ctx ? (r->unique_ctx_selected ? "NaiveUnique" : "NaiveFull")
					   : metric_names[effective_id(r, o) - 1];
	kssd_buf_printf(b, "%s\t%s\t%lf\t%lf\t%s\t%s\t%d\t%f\t%f\t%f\t%f\t%d\t%d\t%d\t%s", qry_name, ref_name,
			r->selected_ani, r->metric, row_confidence(r, o), metric, r->XnY_ctx, r->af_qry,
			bounded_af(r->blastn_af_qry), r->af_ref, bounded_af(r->blastn_af_ref), r->N_diff_obj,
			r->N_diff_obj_section, r->N_mut2_ctx,
			(annotation && annotation[0]) ? annotation : "NA");
	if (o->estimate_coverage) {
		const unsigned char st = r->coverage_status;
		kssd_buf_printf(b, "\t%s", coverage_label(st));
		if (st == 1 || st == 4) {
			kssd_buf_printf(b, "\t%u\t%u", r->cov_markers, r->cov_observed);
			put_cov(b, r->cov_observed_fraction);
			put_cov(b, r->cov_count_sum);
			put_cov(b, r->est_depth);
			put_cov(b, r->pos_mean_depth);
			put_cov(b, r->pos_median_depth);
			put_cov(b, r->count_cv);
			put_cov(b, r->abundance_mass);
			put_cov(b, r->abundance_fraction);
			kssd_buf_printf(b, "\t%s", st == 4 ? "refdb" : "reported_rows");
		} else {
			for (int k = 0; k < 11; ++k)
				kssd_buf_printf(b, "\tNA");
		}
	}
	kssd_buf_printf(b, "\n");
}

/* ------------------------------------------------------------ coverage */
static size_t lower_u64(const uint64_t *a, size_t n, uint64_t v)
{
	size_t lo = 0, hi = n;
	while (lo < hi) {
		const size_t mid = lo + (hi - lo) / 2;
		if (a[mid] < v)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

static int cmp_u64_asc(const void *a, const void *b)
{
	const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return (x > y) - (x < y);
}

static int cmp_u32_asc(const void *a, const void *b)
{
	const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
	return (x > y) - (x < y);
}

/* the record (ctxgid, obj) is in the sorted index with its unique bit set
 * (coverage_unique_bitset_get_key, command_ani.c:2261-2280) */
static bool refdb_marker(const kssd_coverage_src_t *s, uint64_t ctxgid, uint32_t obj)
{
	size_t lo = 0, hi = s->ref_nrec;
	while (lo < hi) {
		const size_t mid = lo + (hi - lo) / 2;
		const kssd_ctxgidobj_t *v = &s->ref_records[mid];
		if (v->ctxgid < ctxgid || (v->ctxgid == ctxgid && v->obj < obj))
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo < s->ref_nrec && s->ref_records[lo].ctxgid == ctxgid && s->ref_records[lo].obj == obj &&
	       (s->ref_unique_bits[lo >> 3] >> (lo & 7)) & 1u;
}

void kssd_populate_coverage(kssd_row_t *rows, size_t n, const kssd_coverage_src_t *s)
{
	if (!n)
		return;
	if (!s->qcounts) {
		for (size_t i = 0; i < n; ++i)
			rows[i].coverage_status = 2;
		return;
	}
	const bool refdb = s->ref_records && s->ref_unique_bits;
	uint64_t *pool = NULL;
	size_t npool = 0;
	if (!refdb) { /* keys of every reported row's reference, sorted */
		for (size_t i = 0; i < n; ++i)
			npool += (size_t)(s->ref_index[rows[i].rn + 1] - s->ref_index[rows[i].rn]);
		pool = malloc((npool ? npool : 1) * sizeof *pool);
		if (!pool)
			abort();
		size_t w = 0;
		for (size_t i = 0; i < n; ++i) {
			const uint64_t b = s->ref_index[rows[i].rn], e = s->ref_index[rows[i].rn + 1];
			memcpy(pool + w, s->ref_keys + b, (size_t)(e - b) * sizeof *pool);
			w += (size_t)(e - b);
		}
		qsort(pool, npool, sizeof *pool, cmp_u64_asc);
	}
	const uint64_t objmask = (s->obj_bits >= 64) ? UINT64_MAX : ((1ULL << s->obj_bits) - 1ULL);
	double total = 0.0;
	for (size_t i = 0; i < n; ++i) {
		kssd_row_t *r = &rows[i];
		const uint64_t b = s->ref_index[r->rn], e = s->ref_index[r->rn + 1];
		uint32_t *pos = malloc((size_t)(e - b ? e - b : 1) * sizeof *pos);
		if (!pos)
			abort();
		uint32_t markers = 0, seen = 0;
		double sum = 0.0, sumsq = 0.0;
		for (uint64_t k = b; k < e; ++k) {
			const uint64_t key = s->ref_keys[k];
			bool marker;
			if (refdb) {
				const uint64_t ctx = s->obj_bits >= 64 ? 0 : key >> s->obj_bits;
				marker = refdb_marker(s, (ctx << 20) | r->rn, (uint32_t)(key & objmask));
			} else {
				const size_t lo = lower_u64(pool, npool, key);
				marker = lo < npool && pool[lo] == key && (lo + 1 == npool || pool[lo + 1] != key);
			}
			if (!marker)
				continue;
			markers++;
			const size_t qi = lower_u64(s->qkeys, s->qn, key);
			if (qi < s->qn && s->qkeys[qi] == key) {
				const uint32_t c = s->qcounts[qi];
				pos[seen++] = c;
				sum += (double)c;
				sumsq += (double)c * (double)c;
			}
		}
		r->coverage_status = refdb ? 4 : 1;
		r->cov_markers = markers;
		r->cov_observed = seen;
		r->cov_observed_fraction = markers ? (double)seen / (double)markers : 0.0;
		r->cov_count_sum = sum;
		r->est_depth = markers ? sum / (double)markers : 0.0;
		r->pos_mean_depth = seen ? sum / (double)seen : 0.0;
		if (seen) {
			qsort(pos, seen, sizeof *pos, cmp_u32_asc);
			r->pos_median_depth = (seen & 1u) ? pos[seen / 2u]
							  : ((double)pos[seen / 2u - 1u] + (double)pos[seen / 2u]) / 2.0;
			const double mean = r->pos_mean_depth;
			double var = sumsq / (double)seen - mean * mean;
			if (var < 0.0 && var > -1e-9)
				var = 0.0;
			r->count_cv = mean > 0.0 && var > 0.0 ? sqrt(var) / mean : 0.0;
		}
		r->abundance_mass = sum;
		total += sum;
		free(pos);
	}
	if (total > 0.0)
		for (size_t i = 0; i < n; ++i)
			rows[i].abundance_fraction = rows[i].abundance_mass / total;
	free(pool);
}

/* ------------------------------------------------------------ sorting */
/* Stable top-down merge sort taking from the left run while cmp <= 0: the
 * element order glibc's qsort produces (msort, glibc 2.35), which the
 * reference's outputs depend on for ties. */
typedef int (*cmp_fn)(const void *, const void *);

static void msort_rec(char *b, size_t n, size_t sz, cmp_fn cmp, char *tmp)
{
	if (n <= 1)
		return;
	const size_t n1 = n / 2, n2 = n - n1;
	char *b1 = b, *b2 = b + n1 * sz;
	msort_rec(b1, n1, sz, cmp, tmp);
	msort_rec(b2, n2, sz, cmp, tmp);
	char *t = tmp;
	size_t i = n1, j = n2;
	while (i > 0 && j > 0) {
		if (cmp(b1, b2) <= 0) {
			memcpy(t, b1, sz);
			b1 += sz;
			--i;
		} else {
			memcpy(t, b2, sz);
			b2 += sz;
			--j;
		}
		t += sz;
	}
	if (i > 0)
		memcpy(t, b1, i * sz);
	memcpy(b, tmp, (n - j) * sz);
}

static void stable_sort(void *base, size_t n, size_t sz, cmp_fn cmp)
{
	if (n <= 1)
		return;
	char *tmp = malloc(n * sz);
	if (!tmp)
		abort();
	msort_rec(base, n, sz, cmp, tmp);
	free(tmp);
}

typedef struct {
	uint32_t id;
	float ani;
} idani_t;

static int cmp_idani_desc(const void *a, const void *b)
{
	/* compare_idani_desc, command_ani.h:361-366 */
	const idani_t *x = a, *y = b;
	return (x->ani < y->ani) - (x->ani > y->ani);
}

static void heap_sift_down(idani_t *h, int n, int i)
{
	for (;;) {
		const int l = i * 2 + 1, r = l + 1;
		int s = i;
		if (l < n && h[l].ani < h[s].ani)
			s = l;
		if (r < n && h[r].ani < h[s].ani)
			s = r;
		if (s == i)
			break;
		const idani_t t = h[i];
		h[i] = h[s];
		h[s] = t;
		i = s;
	}
}

static int select_topN_sorted(idani_t *arr, int m, int topN)
{
	/* idani_select_topN_sorted and partial_topN_inplace_idani,
	 * command_ani.h:369-433 */
	if (m <= 0 || topN <= 0)
		return 0;
	if (topN > m)
		topN = m;
	if (m <= 128 || topN >= m / 2) {
		stable_sort(arr, (size_t)m, sizeof *arr, cmp_idani_desc);
		return topN;
	}
	if (m > 1) {
		for (int i = (topN >> 1) - 1; i >= 0; --i)
			heap_sift_down(arr, topN, i);
		for (int i = topN; i < m; ++i) {
			if (arr[i].ani <= arr[0].ani)
				continue;
			const idani_t t = arr[i];
			arr[i] = arr[0];
			arr[0] = t;
			heap_sift_down(arr, topN, 0);
		}
		stable_sort(arr, (size_t)topN, sizeof *arr, cmp_idani_desc);
	}
	return topN;
}

typedef struct {
	double ani;
	uint32_t id;
} unique_rank_t;

static int cmp_unique_rank(const void *a, const void *b)
{
	/* compare_unique_ctx_rank, command_ani.h:456-461 */
	const unique_rank_t *x = a, *y = b;
	if (x->ani != y->ani)
		return x->ani < y->ani ? 1 : -1;
	return (x->id > y->id) - (x->id < y->id);
}

static int cmp_row_desc(const void *pa, const void *pb)
{
	/* cmp_ani_desc, command_ani.c:1087-1095 */
	const kssd_row_t *a = pa, *b = pb;
	if (a->selected_ani < b->selected_ani)
		return 1;
	if (a->selected_ani > b->selected_ani)
		return -1;
	if (a->af_qry < b->af_qry)
		return 1;
	if (a->af_qry > b->af_qry)
		return -1;
	return (a->rn > b->rn) - (a->rn < b->rn);
}

/* ------------------------------------------------------------ emitters */
static inline const char *anno_at(char (*anno)[KSSD_PATHLEN], uint32_t i)
{
	return anno ? anno[i] : NULL;
}

static inline double naive_of(kssd_ctx_mut2_t c, kssd_obj_section_t ob)
{
	return kssd_naive_dist(c.num_ctx, ob.diff_obj, ob.diff_obj_section);
}

static inline bool use_unique(uint32_t matched, uint32_t full)
{
	/* ani_use_unique_context, command_ani.h:437-440 */
	return full > 0 && (uint64_t)matched * 20 >= full;
}

void kssd_emit_index_detail(kssd_buf_t *out, const kssd_report_opts_t *o, const kssd_query_counts_t *c,
			    const char *qry_name, char (*ref_names)[KSSD_PATHLEN],
			    char (*ref_anno)[KSSD_PATHLEN])
{
	const int R = c->nref;
	const uint32_t ctxcut = (uint32_t)o->ctxcut;
	const int ntop = o->ntop > 0 ? o->ntop : R;
	if (R <= 0)
		return;
	/* ranking, command_ani.h:577-621; the ids stored while ranking are
	 * candidate positions k, ascending with the reference ids j = cand(k),
	 * so every comparison on them orders as on the ids */
	const int ncand = c->ids ? c->ncand : R;
	idani_t *ids = malloc((size_t)(ncand > 0 ? ncand : 1) * sizeof *ids);
	unique_rank_t *ranks = o->unique_ctx ? malloc((size_t)(ncand > 0 ? ncand : 1) * sizeof *ranks) : NULL;
	if (!ids || (o->unique_ctx && !ranks))
		abort();
	int npass = 0;
	for (int k = 0; k < ncand; ++k) {
		const uint32_t j = c->ids ? c->ids[k] : (uint32_t)k;
		const uint32_t shared = c->ctx[k].num_ctx;
		if (shared < ctxcut)
			continue;
		if (c->ref_ctx[j] == 0 || c->qry_ctx == 0 ||
		    !af_pass(o, (double)shared / (double)c->qry_ctx, (double)shared / (double)c->ref_ctx[j]))
			continue;
		const float dist = (float)c->obj[k].diff_obj / c->ctx[k].num_ctx / 7;
		double ani = (float)(1 - dist);
		if (ranks) {
			const bool u = use_unique(c->uctx[k].num_ctx, c->ref_ctx[j]);
			ani = 1.0 - naive_of(u ? c->uctx[k] : c->ctx[k], u ? c->uobj[k] : c->obj[k]);
		}
		if (ani < o->anicut)
			continue;
		if (ranks)
			ranks[npass] = (unique_rank_t){ani, (uint32_t)k};
		ids[npass].id = (uint32_t)k;
		ids[npass].ani = (float)ani;
		npass++;
	}
	int out_n;
	if (ranks) {
		stable_sort(ranks, (size_t)npass, sizeof *ranks, cmp_unique_rank);
		out_n = npass < ntop ? npass : ntop;
		for (int r = 0; r < out_n; ++r)
			ids[r].id = ranks[r].id;
	} else {
		out_n = select_topN_sorted(ids, npass, ntop);
	}
	/* printing, ani_block_print_internal command_ani.c:3470-3539; with
	 * coverage the printed rows are held back until their coverage is known */
	const bool cov = o->estimate_coverage && c->cov;
	kssd_row_t *held = cov ? malloc((size_t)(out_n > 0 ? out_n : 1) * sizeof *held) : NULL;
	size_t nheld = 0;
	if (cov && !held)
		abort();
	const int qsz = (int)c->qry_ctx;
	for (int n = 0; n < out_n; ++n) {
		const uint32_t k = ids[n].id;
		const uint32_t j = c->ids ? c->ids[k] : k;
		const int rsz = (int)c->ref_ctx[j];
		if (qsz == 0 || rsz == 0)
			continue;
		const uint32_t xny = c->ctx[k].num_ctx, mut2 = c->ctx[k].num_mut2_ctx;
		const uint32_t dobj = c->obj[k].diff_obj, dsec = c->obj[k].diff_obj_section;
		const float afq = (float)xny / qsz, afr = (float)xny / rsz;
		const float bq = (float)kssd_af_anib(xny, dobj, dsec, mut2, (uint32_t)qsz, o->drfold);
		const float br = (float)kssd_af_anib(xny, dobj, dsec, mut2, (uint32_t)rsz, o->drfold);
		kssd_row_t row = kssd_make_row(j, xny, dobj, dsec, mut2, o, (uint32_t)qsz, (uint32_t)rsz, afq, bq,
					       afr, br, c->qry_meta, c->ref_meta ? &c->ref_meta[j] : NULL);
		if (o->unique_ctx) {
			row.unique_ctx_selected = use_unique(c->uctx[k].num_ctx, c->ref_ctx[j]);
			row.naive_dist = naive_of(row.unique_ctx_selected ? c->uctx[k] : c->ctx[k],
						  row.unique_ctx_selected ? c->uobj[k] : c->obj[k]);
			kssd_finalize_row(&row, o);
		}
		if (row.selected_ani < o->anicut)
			continue;
		if (cov)
			held[nheld++] = row;
		else
			kssd_append_detail_row(out, o, qry_name, ref_names[j], &row, anno_at(ref_anno, j));
	}
	if (cov) {
		kssd_populate_coverage(held, nheld, c->cov);
		for (size_t r = 0; r < nheld; ++r)
			kssd_append_detail_row(out, o, qry_name, ref_names[held[r].rn], &held[r],
					       anno_at(ref_anno, held[r].rn));
		free(held);
	}
	free(ranks);
	free(ids);
}

void kssd_emit_index_matrix_row(kssd_buf_t *out, const kssd_report_opts_t *o,
				const kssd_query_counts_t *c, const char *qry_name)
{
	/* matrix branch of ani_block_print_internal, command_ani.c:3465-3551 */
	kssd_buf_printf(out, "%s", qry_name);
	const int qsz = (int)c->qry_ctx;
	for (int j = 0; j < c->nref; ++j) {
		const int rsz = (int)c->ref_ctx[j];
		if (qsz == 0 || rsz == 0) {
			kssd_buf_printf(out, "\t%lf", kssd_matrix_exception(o));
			continue;
		}
		const uint32_t xny = c->ctx[j].num_ctx, mut2 = c->ctx[j].num_mut2_ctx;
		const uint32_t dobj = c->obj[j].diff_obj, dsec = c->obj[j].diff_obj_section;
		const float afq = (float)xny / qsz, afr = (float)xny / rsz;
		const float bq = (float)kssd_af_anib(xny, dobj, dsec, mut2, (uint32_t)qsz, o->drfold);
		const float br = (float)kssd_af_anib(xny, dobj, dsec, mut2, (uint32_t)rsz, o->drfold);
		kssd_row_t row = kssd_make_row((uint32_t)j, xny, dobj, dsec, mut2, o, (uint32_t)qsz, (uint32_t)rsz,
					       afq, bq, afr, br, c->qry_meta,
					       c->ref_meta ? &c->ref_meta[j] : NULL);
		if (o->unique_ctx) {
			row.unique_ctx_selected = use_unique(c->uctx[j].num_ctx, c->ref_ctx[j]);
			row.naive_dist = naive_of(row.unique_ctx_selected ? c->uctx[j] : c->ctx[j],
						  row.unique_ctx_selected ? c->uobj[j] : c->obj[j]);
			kssd_finalize_row(&row, o);
		}
		kssd_buf_printf(out, "\t%lf", o->s < 0 ? row.selected_ani : row.metric);
	}
	kssd_buf_printf(out, "\n");
}

static void emit_sorted(kssd_buf_t *out, const kssd_report_opts_t *o, kssd_row_t *rows, size_t n,
			size_t limit, const char *qry_name, char (*ref_names)[KSSD_PATHLEN],
			char (*ref_anno)[KSSD_PATHLEN], const kssd_coverage_src_t *cov)
{
	stable_sort(rows, n, sizeof *rows, cmp_row_desc);
	if (limit < n)
		n = limit;
	if (o->estimate_coverage && cov) /* :2634-2650, :2983-2998 */
		kssd_populate_coverage(rows, n, cov);
	for (size_t i = 0; i < n; ++i)
		kssd_append_detail_row(out, o, qry_name, ref_names[rows[i].rn], &rows[i],
				       anno_at(ref_anno, rows[i].rn));
}

void kssd_emit_stream_detail(kssd_buf_t *out, const kssd_report_opts_t *o, const kssd_query_counts_t *c,
			     const char *qry_name, char (*ref_names)[KSSD_PATHLEN],
			     char (*ref_anno)[KSSD_PATHLEN], int single)
{
	/* one query: compute_streamed_ref_row_one_qry, command_ani.c:2110-2201;
	 * several: the per-reference loop of command_ani.c:2892-2968 (no AF or
	 * close-filter early exit).  Order and top N as :2634-2650, :2983-2998 */
	kssd_row_t *rows = malloc((size_t)(c->nref > 0 ? c->nref : 1) * sizeof *rows);
	if (!rows)
		abort();
	size_t n = 0;
	const uint32_t qc = c->qry_ctx;
	for (int j = 0; j < c->nref; ++j) {
		const uint32_t rc = c->ref_ctx[j];
		if (qc == 0 || rc == 0)
			continue;
		const uint32_t need = af_needed_ctx(o, qc, rc);
		const uint32_t xny = c->ctx[j].num_ctx, mut2 = c->ctx[j].num_mut2_ctx;
		const uint32_t dobj = c->obj[j].diff_obj, dsec = c->obj[j].diff_obj_section;
		if (single && (xny < need || close_filter_impossible(o, dsec, xny)))
			continue;
		/* the walk gives a reference up once the shared contexts so far plus
		 * the contexts after the current one fall short of `need`, before it
		 * looks the current one up (command_ani.c:2154-2156): a pair that
		 * reaches `need` exactly on its last context is given up too */
		if (single && need > 0 && xny == need && c->last_shared && c->last_shared[j])
			continue;
		if (xny < (uint32_t)o->ctxcut)
			continue;
		const double afq = (double)xny / (double)qc, afr = (double)xny / (double)rc;
		if (!af_pass(o, afq, afr))
			continue;
		const double bq = kssd_af_anib(xny, dobj, dsec, mut2, qc, o->drfold);
		const double br = kssd_af_anib(xny, dobj, dsec, mut2, rc, o->drfold);
		kssd_row_t row = kssd_make_row((uint32_t)j, xny, dobj, dsec, mut2, o, qc, rc, afq, bq, afr, br,
					       c->qry_meta, c->ref_meta ? &c->ref_meta[j] : NULL);
		if ((single && !passes_close_filter(&row, o)) || row.selected_ani < o->anicut)
			continue;
		rows[n++] = row;
	}
	emit_sorted(out, o, rows, n, o->ntop > 0 ? (size_t)o->ntop : n, qry_name, ref_names, ref_anno, c->cov);
	free(rows);
}

void kssd_emit_hash_detail(kssd_buf_t *out, const kssd_report_opts_t *o, const kssd_query_counts_t *c,
			   const char *qry_name, char (*ref_names)[KSSD_PATHLEN],
			   char (*ref_anno)[KSSD_PATHLEN])
{
	/* comb_manysmall_..._filter_and_sort_survivors, command_ani.c:3888-4000:
	 * no ctxcut, ANI must exceed anicut, every survivor printed */
	kssd_row_t *rows = malloc((size_t)(c->nref > 0 ? c->nref : 1) * sizeof *rows);
	if (!rows)
		abort();
	size_t n = 0;
	const uint32_t qc = c->qry_ctx;
	for (int j = 0; j < c->nref; ++j) {
		const uint32_t rc = c->ref_ctx[j];
		if (rc == 0 || qc == 0)
			continue;
		const uint32_t need = af_needed_ctx(o, qc, rc);
		const uint32_t xny = c->ctx[j].num_ctx, mut2 = c->ctx[j].num_mut2_ctx;
		const uint32_t dobj = c->obj[j].diff_obj, dsec = c->obj[j].diff_obj_section;
		if (xny < need || close_filter_impossible(o, dsec, xny))
			continue;
		const double afq = (double)xny / (double)qc, afr = (double)xny / (double)rc;
		if (!af_pass(o, afq, afr))
			continue;
		const double bq = kssd_af_anib(xny, dobj, dsec, mut2, qc, o->drfold);
		const double br = kssd_af_anib(xny, dobj, dsec, mut2, rc, o->drfold);
		kssd_row_t row = kssd_make_row((uint32_t)j, xny, dobj, dsec, mut2, o, qc, rc, afq, bq, afr, br,
					       c->qry_meta, c->ref_meta ? &c->ref_meta[j] : NULL);
		if (!passes_close_filter(&row, o) || row.selected_ani <= o->anicut)
			continue;
		rows[n++] = row;
	}
	emit_sorted(out, o, rows, n, n, qry_name, ref_names, ref_anno, NULL); /* no coverage here */
	free(rows);
}

/* ------------------------------------------------------------ matrices */
double kssd_matrix_exception(const kssd_report_opts_t *o)
{
	/* ani_matrix_exception_value, command_ani.c:1448-1451 */
	return o->s < 0 ? 1.0 - o->e : o->e;
}

double kssd_matrix_diagonal(const kssd_report_opts_t *o)
{
	return o->s < 0 ? 1.0 : 0.0;
}

double kssd_matrix_value(const kssd_report_opts_t *o, uint32_t xny, uint32_t dobj, uint32_t dsec,
			 uint32_t mut2, uint32_t qry_ctx, uint32_t ref_ctx,
			 const kssd_infile_meta_t *qry_meta, const kssd_infile_meta_t *ref_meta)
{
	/* ani_matrix_pair_value :1458-1496 / ani_indexed_self_pair_metric :1627-1655 */
	if (qry_ctx == 0 || ref_ctx == 0)
		return kssd_matrix_exception(o);
	const double afq = (double)xny / (double)qry_ctx, afr = (double)xny / (double)ref_ctx;
	const double bq = kssd_af_anib(xny, dobj, dsec, mut2, qry_ctx, o->drfold);
	const double br = kssd_af_anib(xny, dobj, dsec, mut2, ref_ctx, o->drfold);
	kssd_row_t row = kssd_make_row(0, xny, dobj, dsec, mut2, o, qry_ctx, ref_ctx, afq, bq, afr, br,
				       qry_meta, ref_meta);
	return o->s < 0 ? row.selected_ani : row.metric;
}

int32_t kssd_scale_fixed6(double v)
{
	/* ani_scale_fixed6, command_ani.c:1593-1604 */
	if (!isfinite(v))
		v = 1.0;
	if (v > 2147.0)
		v = 2147.0;
	if (v < -2147.0)
		v = -2147.0;
	return (int32_t)(v >= 0.0 ? v * 1000000.0 + 0.5 : v * 1000000.0 - 0.5);
}

void kssd_put_fixed6(kssd_buf_t *b, int32_t scaled)
{
	/* ani_kput_scaled6, command_ani.c:1606-1625 */
	uint32_t mag;
	if (scaled < 0) {
		kssd_buf_printf(b, "-");
		mag = (uint32_t)(-(int64_t)scaled);
	} else {
		mag = (uint32_t)scaled;
	}
	kssd_buf_printf(b, "%u.%06u", mag / 1000000U, mag % 1000000U);
}

/* ------------------------------------------------------------ matrix command */
int kssd_matrix_metric(const char *a)
{
	/* pairwise_metric_from_string, pairwise_graph.c:69-98 */
	if (!a)
		return -1;
	if (!strcmp(a, "ctx-moe") || !strcmp(a, "moe"))
		return KSSD_MM_MOE;
	if (!strcmp(a, "ctx-naive") || !strcmp(a, "naive"))
		return KSSD_MM_NAIVE;
	if (!strcmp(a, "p_dist") || !strcmp(a, "pdist") || !strcmp(a, "p-distance") || !strcmp(a, "micro-obj") ||
	    !strcmp(a, "micro_obj_rate"))
		return KSSD_MM_PDIST;
	if (!strcmp(a, "mash") || !strcmp(a, "mashd") || !strcmp(a, "0"))
		return KSSD_MM_MASH;
	if (!strcmp(a, "aaf") || !strcmp(a, "aafd") || !strcmp(a, "1"))
		return KSSD_MM_AAF;
	return -1;
}

double kssd_matrix_context_value(int metric, double e, uint32_t xny, uint32_t dobj, uint32_t dsec,
				 uint32_t mut2, uint32_t qc, uint32_t rc, int drfold, unsigned obj_bits,
				 unsigned ctx_bits)
{
	/* matrix_context_feature_value + matrix_context_metric_distance,
	 * command_matrix.c:400-419, :621-661 (one metric) */
	if (xny == 0 || qc == 0 || rc == 0)
		return e;
	const uint32_t k = ctx_bits / 2;
	double d;
	switch (metric) {
	case KSSD_MM_NAIVE:
		d = kssd_naive_dist(xny, dobj, dsec);
		break;
	case KSSD_MM_MOE:
		d = kssd_moe_dist(xny, dobj, dsec, mut2, drfold);
		break;
	case KSSD_MM_PDIST:
		d = kssd_p_dist(xny, dsec, obj_bits);
		break;
	case KSSD_MM_MASH:
		d = kssd_mashD(k, rc, qc, xny);
		break;
	case KSSD_MM_AAF:
		d = kssd_aafD(k, rc, qc, xny);
		break;
	default:
		d = 1.0;
	}
	if (!isfinite(d))
		return e;
	return d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
}

double kssd_matrix_context_expr_value(const kssd_metric_expr_t *x, double e, uint32_t xny, uint32_t dobj,
				      uint32_t dsec, uint32_t mut2, uint32_t qc, uint32_t rc, int drfold,
				      unsigned obj_bits, unsigned ctx_bits)
{
	/* matrix_context_feature_value over an expression, command_matrix.c:621-661:
	 * each metric clamped to [0, 1], & keeps the largest, | the smallest */
	if (xny == 0 || qc == 0 || rc == 0 || x->count == 0)
		return e;
	double sel = 0.0;
	for (int i = 0; i < x->count; ++i) {
		const double d = kssd_matrix_context_value(x->metrics[i], NAN, xny, dobj, dsec, mut2, qc, rc, drfold,
							   obj_bits, ctx_bits);
		if (isnan(d))
			return e;
		if (i == 0 || (x->op == KSSD_EXPR_AND ? d > sel : d < sel))
			sel = d;
	}
	return sel;
}

static double clean_distance(double d)
{
	/* pairwise_clean_distance, pairwise_graph.c:14-21 */
	if (!isfinite(d))
		return 1.0;
	return d < 0.0 ? 0.0 : d;
}

double kssd_pairwise_value(int metric, double e, size_t qry_len, size_t ref_len, uint32_t xny, uint32_t dobj,
			   uint32_t dsec, uint32_t mut2, uint32_t qc, uint32_t rc, int drfold, unsigned obj_bits,
			   unsigned ctx_bits)
{
	/* pairwise_eval_arrays, pairwise_graph.c:301-360, read by
	 * matrix_eval_value (command_matrix.c:82-85) */
	if (qry_len == 0 || ref_len == 0 || xny == 0)
		return e;
	const double k = ctx_bits > 0 ? (double)(ctx_bits / 2) : 32.0;
	switch (metric) {
	case KSSD_MM_MASH: {
		/* pairwise_mash_distance, pairwise_graph.c:23-36 */
		const double denom = (double)qc + (double)rc - (double)xny;
		if (denom <= 0.0)
			return 1.0;
		const double j = (double)xny / denom;
		const double t = (2.0 * j) / (1.0 + j);
		if (t <= 0.0)
			return 1.0;
		return clean_distance(-log(t) / k);
	}
	case KSSD_MM_AAF: {
		/* pairwise_aaf_distance, pairwise_graph.c:38-50 */
		const uint32_t mn = qc < rc ? qc : rc;
		if (mn == 0)
			return 1.0;
		const double c = (double)xny / (double)mn;
		if (c <= 0.0)
			return 1.0;
		return clean_distance(-log(c) / k);
	}
	case KSSD_MM_NAIVE:
		return clean_distance(kssd_naive_dist(xny, dobj, dsec));
	case KSSD_MM_PDIST:
		return clean_distance(kssd_p_dist(xny, dsec, obj_bits));
	default:
		return clean_distance(kssd_moe_dist(xny, dobj, dsec, mut2, drfold));
	}
}

const char *kssd_metric_name(int metric)
{
	switch (metric) {
	case KSSD_MM_MOE:
		return "ctx-moe";
	case KSSD_MM_NAIVE:
		return "ctx-naive";
	case KSSD_MM_PDIST:
		return "p_dist";
	case KSSD_MM_MASH:
		return "mash";
	case KSSD_MM_AAF:
		return "aaf";
	}
	return "ctx-moe";
}

static int expr_label(kssd_metric_expr_t *e)
{
	const char *sep = e->op == KSSD_EXPR_AND ? "&" : e->op == KSSD_EXPR_OR ? "|" : "";
	size_t used = 0;
	e->label[0] = '\0';
	for (int i = 0; i < e->count; ++i) {
		const int w = snprintf(e->label + used, sizeof e->label - used, "%s%s", i ? sep : "",
				       kssd_metric_name(e->metrics[i]));
		if (w < 0 || (size_t)w >= sizeof e->label - used)
			return -1;
		used += (size_t)w;
	}
	return 0;
}

kssd_metric_expr_t kssd_metric_expr_single(int metric)
{
	kssd_metric_expr_t e;
	memset(&e, 0, sizeof e);
	e.metrics[0] = metric;
	e.count = 1;
	e.op = KSSD_EXPR_SINGLE;
	(void)expr_label(&e);
	return e;
}

int kssd_metric_expr_parse(const char *arg, kssd_metric_expr_t *out)
{
	/* whitespace dropped; one kind of operator; no leading, trailing or
	 * doubled operator; every term a metric name */
	if (!arg || !out || !arg[0])
		return -1;
	char compact[96];
	size_t n = 0;
	for (const char *p = arg; *p; ++p) {
		if (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')
			continue;
		if (n + 1 >= sizeof compact)
			return -1;
		compact[n++] = *p;
	}
	compact[n] = '\0';
	if (!n)
		return -1;
	const int has_and = strchr(compact, '&') != NULL, has_or = strchr(compact, '|') != NULL;
	if (has_and && has_or)
		return -1;
	kssd_metric_expr_t e;
	memset(&e, 0, sizeof e);
	e.op = has_and ? KSSD_EXPR_AND : has_or ? KSSD_EXPR_OR : KSSD_EXPR_SINGLE;
	if (e.op != KSSD_EXPR_SINGLE) {
		const char op = has_and ? '&' : '|';
		if (compact[0] == op || compact[n - 1] == op)
			return -1;
		for (size_t i = 1; i < n; ++i)
			if (compact[i] == op && compact[i - 1] == op)
				return -1;
	}
	const char delim[2] = {has_and ? '&' : has_or ? '|' : '\0', '\0'};
	char *save = NULL;
	char *tok = e.op == KSSD_EXPR_SINGLE ? compact : strtok_r(compact, delim, &save);
	while (tok) {
		if (e.count >= KSSD_EXPR_MAX)
			return -1;
		const int m = kssd_matrix_metric(tok);
		if (m < 0)
			return -1;
		e.metrics[e.count++] = m;
		tok = e.op == KSSD_EXPR_SINGLE ? NULL : strtok_r(NULL, delim, &save);
	}
	if (!e.count)
		return -1;
	if (e.count == 1)
		e.op = KSSD_EXPR_SINGLE;
	if (expr_label(&e))
		return -1;
	*out = e;
	return 0;
}

static kssd_pair_eval_t pair_eval_one(int metric, size_t qry_len, size_t ref_len, uint32_t xny, uint32_t dobj,
				      uint32_t dsec, uint32_t mut2, uint32_t qc, uint32_t rc, int drfold,
				      unsigned obj_bits, unsigned ctx_bits)
{
	/* pairwise_eval_arrays, pairwise_graph.c:301-360 */
	kssd_pair_eval_t ev;
	memset(&ev, 0, sizeof ev);
	ev.distance = 1.0;
	if (qry_len == 0 || ref_len == 0)
		return ev;
	ev.xny = xny;
	ev.dobj = dobj;
	ev.dsec = dsec;
	ev.mut2 = mut2;
	ev.af_qry = qc > 0 ? (double)xny / (double)qc : 0.0;
	ev.af_ref = rc > 0 ? (double)xny / (double)rc : 0.0;
	ev.max_af = ev.af_qry > ev.af_ref ? ev.af_qry : ev.af_ref;
	if (xny == 0)
		return ev;
	ev.distance = kssd_pairwise_value(metric, 1.0, qry_len, ref_len, xny, dobj, dsec, mut2, qc, rc, drfold,
					  obj_bits, ctx_bits);
	ev.similarity = 1.0 - ev.distance;
	ev.valid = 1;
	return ev;
}

kssd_pair_eval_t kssd_pair_eval(const kssd_metric_expr_t *e, size_t qry_len, size_t ref_len, uint32_t xny,
				uint32_t dobj, uint32_t dsec, uint32_t mut2, uint32_t qc, uint32_t rc, int drfold,
				unsigned obj_bits, unsigned ctx_bits)
{
	if (!e || e->count == 0)
		return pair_eval_one(KSSD_MM_MOE, qry_len, ref_len, xny, dobj, dsec, mut2, qc, rc, drfold, obj_bits,
				     ctx_bits);
	if (e->count == 1 || e->op == KSSD_EXPR_SINGLE)
		return pair_eval_one(e->metrics[0], qry_len, ref_len, xny, dobj, dsec, mut2, qc, rc, drfold, obj_bits,
				     ctx_bits);
	/* pairwise_eval_expr_merge, pairwise_graph.c:420-448 */
	kssd_pair_eval_t merged;
	memset(&merged, 0, sizeof merged);
	merged.distance = 1.0;
	for (int i = 0; i < e->count; ++i) {
		const kssd_pair_eval_t ev = pair_eval_one(e->metrics[i], qry_len, ref_len, xny, dobj, dsec, mut2, qc,
							  rc, drfold, obj_bits, ctx_bits);
		if (e->op == KSSD_EXPR_AND) {
			if (!ev.valid)
				return merged;
			if (!merged.valid || ev.distance > merged.distance)
				merged = ev;
		} else if (ev.valid && (!merged.valid || ev.distance < merged.distance)) {
			merged = ev;
		}
	}
	if (merged.valid)
		merged.similarity = 1.0 - merged.distance;
	return merged;
}

int kssd_pair_passes_edge(const kssd_pair_eval_t *ev, double cut, uint32_t ctxcut, double max_afcut)
{
	return ev && ev->valid && ev->distance < cut && ev->xny >= ctxcut && ev->max_af >= max_afcut;
}

static int meta_rankable(const kssd_infile_meta_t *meta, int i)
{
	/* pairwise_meta_rankable: a current-version record with a length */
	return meta && meta[i].meta_fmt_version == 1u && meta[i].total_length_bp > 0;
}

static double finite_or(double v, double fallback)
{
	return isfinite(v) ? v : fallback;
}

static int positive_lower_cmp(uint32_t a, uint32_t b)
{
	/* fewer records is better, but 0 (unknown) loses */
	if (a == b)
		return 0;
	if (a == 0)
		return -1;
	if (b == 0)
		return 1;
	return a < b ? 1 : -1;
}

int kssd_quality_compare(const kssd_infile_meta_t *meta, const uint64_t *index, int a, int b)
{
	if (a == b)
		return 0;
	const int ra = meta_rankable(meta, a), rb = meta_rankable(meta, b);
	if (ra != rb)
		return ra ? 1 : -1;
	if (ra && rb) {
		const kssd_infile_meta_t *ma = &meta[a], *mb = &meta[b];
		const int fa = ma->infile_fmt == 1, fb = mb->infile_fmt == 1; /* FASTA first */
		if (fa != fb)
			return fa ? 1 : -1;
		const double aa = finite_or((double)ma->asm_level, -1.0), ab = finite_or((double)mb->asm_level, -1.0);
		if (aa != ab)
			return aa > ab ? 1 : -1;
		if (ma->total_length_bp != mb->total_length_bp)
			return ma->total_length_bp > mb->total_length_bp ? 1 : -1;
		const int rc = positive_lower_cmp(ma->record_count, mb->record_count);
		if (rc)
			return rc;
		if (ma->median_length_bp != mb->median_length_bp)
			return ma->median_length_bp > mb->median_length_bp ? 1 : -1;
		const double ca = finite_or((double)ma->length_cv, INFINITY), cb = finite_or((double)mb->length_cv, INFINITY);
		if (ca != cb)
			return ca < cb ? 1 : -1;
	}
	const uint64_t ea = index[a + 1] - index[a], eb = index[b + 1] - index[b];
	if (ea != eb)
		return ea > eb ? 1 : -1;
	return a < b ? 1 : -1;
}

void kssd_put_edge_header(kssd_buf_t *b)
{
	kssd_buf_printf(b, "Qry\tRef\tDistance\tSimilarity\tMetric\tXnY_ctx\tQry_align_fraction\tRef_align_fraction\t"
			   "max_AF\tN_diff_obj\tN_diff_obj_section\tN_mut2_ctx\n");
}

void kssd_put_edge(kssd_buf_t *b, const char *qry_name, const char *ref_name, const kssd_pair_eval_t *ev,
		   const char *label)
{
	kssd_buf_printf(b, "%s\t%s\t%.12g\t%.12g\t%s\t%u\t%.12g\t%.12g\t%.12g\t%u\t%u\t%u\n", qry_name, ref_name,
			ev->distance, ev->similarity, label, ev->xny, ev->af_qry, ev->af_ref, ev->max_af, ev->dobj,
			ev->dsec, ev->mut2);
}

static void put_frac6(kssd_buf_t *b, uint32_t frac)
{
	char f[7];
	for (int i = 5; i >= 0; --i) {
		f[i] = (char)('0' + frac % 10);
		frac /= 10;
	}
	f[6] = '\0';
	kssd_buf_printf(b, "%s", f);
}

void kssd_matrix_put_fixed6(kssd_buf_t *b, double v)
{
	/* matrix_kput_fixed6, command_matrix.c:103-124 */
	if (!isfinite(v) || v > 900000000000.0 || v < -900000000000.0) {
		kssd_buf_printf(b, "%lf", v);
		return;
	}
	if (v < 0.0) {
		kssd_buf_printf(b, "-");
		v = -v;
	}
	const uint64_t scaled = (uint64_t)(v * 1000000.0 + 0.5);
	kssd_buf_printf(b, "%ld.", (long)(scaled / 1000000ULL));
	put_frac6(b, (uint32_t)(scaled % 1000000ULL));
}

uint32_t kssd_matrix_scale_fixed6(double v)
{
	/* matrix_scale_fixed6, command_matrix.c:126-135 */
	if (!isfinite(v))
		v = 1.0;
	if (v < 0.0)
		v = 0.0;
	if (v > 4294.0)
		v = 4294.0;
	return (uint32_t)(v * 1000000.0 + 0.5);
}

void kssd_matrix_put_scaled6(kssd_buf_t *b, uint32_t scaled)
{
	/* matrix_kput_scaled6, command_matrix.c:137-149 */
	kssd_buf_printf(b, "%u.", scaled / 1000000U);
	put_frac6(b, scaled % 1000000U);
}
