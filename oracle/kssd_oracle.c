/*
 * CPU oracle for the KSSD3A sketch-comparison path -- TEST INFRASTRUCTURE.
 * See kssd_oracle.h for the contract and for which tests pin each function.
 * Citations are to the reference tree (yhg926/KSSD3A-public v3.2.0-rc.1).
 */
#include "kssd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_GID_NBITS 20                    /* sketch_rearrange.h:7 */
#define ORC_GIDMASK ((1ULL << ORC_GID_NBITS) - 1ULL)
#define ORC_MIN_INIT 12                     /* NUM_CODENS + 1, command_ani.h:535 */
#define ORC_EPS 1e-8                        /* model_ani.h:224 */

int orc_dna_popcount(uint32_t x)
{
	/* dna_popcount.h:45-49: one bit per 2-bit base lane that differs */
	x |= x >> 1;
	x &= 0x55555555u;
	return __builtin_popcount(x);
}

void orc_dna_popcount_many(const uint32_t *x, size_t n, uint8_t *out)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = (uint8_t)orc_dna_popcount(x[i]);
}

int orc_fenceposts_build(const orc_rec_t *b, size_t n, int k, size_t *F)
{
	/* command_ani.c:3656-3678.  The bucket of a record is the top k bits of
	 * the 64-bit value ctx = ctxgid >> 20 (command_ani.c:3634-3641). */
	if (!b || !F || k < 0 || k > 32)
		return -1;
	const size_t buckets = (size_t)1 << k;
	for (size_t t = 0; t <= buckets; ++t)
		F[t] = n;
	size_t next = 0;
	for (size_t i = 0; i < n; ++i) {
		const uint64_t key = b[i].ctxgid >> ORC_GID_NBITS;
		const uint32_t top = k == 0 ? 0u : (uint32_t)(key >> (64 - k));
		while (next <= top)
			F[next++] = i;
		if (next > buckets)
			break;
	}
	while (next <= buckets)
		F[next++] = n;
	return 0;
}

static size_t orc_lower_bound_ctx(const orc_rec_t *b, size_t lo, size_t hi, uint64_t target)
{
	while (lo < hi) {
		const size_t mid = lo + ((hi - lo) >> 1);
		if ((b[mid].ctxgid >> ORC_GID_NBITS) < target)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

void orc_find_first(const uint64_t *a, size_t na, const orc_rec_t *b, size_t nb,
		    const size_t *F, int k, unsigned obj_bits, size_t *idx)
{
	/* command_ani.c:3680-3738 */
	uint64_t prev_key = UINT64_MAX;
	size_t prev_idx = SIZE_MAX;
	const size_t buckets = (size_t)1 << k;
	for (size_t i = 0; i < na; ++i) {
		const uint64_t target = a[i] >> obj_bits;
		if (i && target == prev_key) {
			idx[i] = prev_idx;
			continue;
		}
		uint32_t t = k == 0 ? 0u : (uint32_t)(target >> (64 - k));
		if (t > buckets)
			t = (uint32_t)buckets;
		const size_t pos = orc_lower_bound_ctx(b, F[t], F[t + 1], target);
		idx[i] = (pos < nb && (b[pos].ctxgid >> ORC_GID_NBITS) == target) ? pos : SIZE_MAX;
		prev_key = target;
		prev_idx = idx[i];
	}
}

/* min over (query objs of one ctx group) x (ref objs of one gid run) of the
 * per-base mismatch count; 0 stops the search (command_ani.h:535-550). */
static inline int orc_min_cross(const uint64_t *a, size_t s, uint64_t objmask,
				const orc_rec_t *index, size_t rb, size_t re)
{
	int m = ORC_MIN_INIT;
	for (size_t n = 0; n < s; ++n) {
		const uint32_t oq = (uint32_t)(a[n] & objmask);
		for (size_t r = rb; r < re; ++r) {
			const uint32_t x = oq ^ index[r].obj;
			if (x == 0) {
				m = 0;
				break;
			}
			const int d = orc_dna_popcount(x);
			if (d < m)
				m = d;
		}
		if (m == 0)
			break;
	}
	return m;
}

void orc_count_block(const orc_rec_t *index, size_t n_index, int ref_gnum,
		     unsigned obj_bits, const uint64_t *section_sk,
		     const uint64_t *section_skidx, int nq, int qry_gid_offset,
		     unsigned flags, orc_ctx_t *ctx, orc_obj_t *obj,
		     orc_ctx_t *uctx, orc_obj_t *uobj, int nthreads)
{
	const uint64_t objmask = (1ULL << obj_bits) - 1ULL;
	const int lower = (flags & ORC_LOWER) != 0;
	const int unique = !lower && (flags & ORC_UNIQUE) != 0 && uctx && uobj;
	const int ignoreconflict = (flags & ORC_IGNORECONFLICT) != 0;
	/* Fenceposts are rebuilt per call with k = 17, as the reference does
	 * (command_ani.h:472-479, :645-653). */
	const int k = 17;
	size_t *F = malloc((((size_t)1 << k) + 1) * sizeof *F);
	if (!F)
		abort();
	orc_fenceposts_build(index, n_index, k, F);
	if (nthreads <= 0)
		nthreads = 1;

#pragma omp parallel for num_threads(nthreads) schedule(guided)
	for (int i = 0; i < nq; ++i) {
		const size_t a_begin = (size_t)(section_skidx[i] - section_skidx[0]);
		const size_t a_size = (size_t)(section_skidx[i + 1] - section_skidx[i]);
		if (a_size == 0)
			continue;
		const uint64_t *a = section_sk + a_begin;
		size_t *idx = malloc(a_size * sizeof *idx);
		if (!idx)
			abort();
		orc_find_first(a, a_size, index, n_index, F, k, obj_bits, idx);
		const int qry_gid = qry_gid_offset + i;
		size_t s = 1;
		for (size_t j = 0; j < a_size; j += s) {
			/* group query entries that resolved to the same index run
			 * (command_ani.h:506-511) */
			for (s = 1; j + s < a_size && idx[j] == idx[j + s]; ++s)
				;
			if (idx[j] == SIZE_MAX)
				continue;
			const uint64_t qctx = a[j] >> obj_bits;
			size_t d = idx[j];
			size_t owners = 0;
			uint32_t sole_gid = 0;
			int sole_diff = 0, sole_allowed = 0;
			while (d < n_index && (index[d].ctxgid >> ORC_GID_NBITS) == qctx) {
				const uint64_t cg = index[d].ctxgid;
				const uint32_t gid = (uint32_t)(cg & ORC_GIDMASK);
				const size_t rb = d;
				do {
					++d;
				} while (d < n_index && index[d].ctxgid == cg);
				const size_t re = d;
				if (lower) {
					/* command_ani.h:685-689 */
					if ((int)gid >= qry_gid) {
						if ((int)gid > qry_gid)
							break;
						continue;
					}
				}
				if (unique) {
					++owners;
					sole_gid = gid;
				}
				if (ignoreconflict && re - rb > 1)
					continue;
				const size_t cell = (size_t)i * (size_t)ref_gnum + gid;
				ctx[cell].num_ctx++;
				const int m = orc_min_cross(a + j, s, objmask, index, rb, re);
				if (unique) {
					sole_diff = m;
					sole_allowed = 1;
				}
				if (m) {
					obj[cell].diff_obj++;
					obj[cell].diff_obj_section += (uint32_t)m;
					if (m > 1)
						ctx[cell].num_mut2_ctx++;
				}
			}
			/* command_ani.h:564-574 */
			if (unique && owners == 1 && sole_allowed) {
				const size_t cell = (size_t)i * (size_t)ref_gnum + sole_gid;
				uctx[cell].num_ctx++;
				if (sole_diff) {
					uobj[cell].diff_obj++;
					uobj[cell].diff_obj_section += (uint32_t)sole_diff;
					if (sole_diff > 1)
						uctx[cell].num_mut2_ctx++;
				}
			}
		}
		free(idx);
	}
	free(F);
}

void orc_pair_features(const uint64_t *a, size_t n, const uint64_t *b, size_t m,
		       unsigned obj_bits, int ignore_b_conflict, orc_features_t *out)
{
	/* command_ani.h:303-346 */
	const uint64_t objmask = (1ULL << obj_bits) - 1ULL;
	memset(out, 0, sizeof *out);
	size_t i = 0, j = 0;
	while (i < n && j < m) {
		const uint64_t ca = a[i] >> obj_bits, cb = b[j] >> obj_bits;
		if (ca < cb) {
			do {
				++i;
			} while (i < n && (a[i] >> obj_bits) == ca);
			continue;
		}
		if (cb < ca) {
			do {
				++j;
			} while (j < m && (b[j] >> obj_bits) == cb);
			continue;
		}
		const size_t ab = i;
		do {
			++i;
		} while (i < n && (a[i] >> obj_bits) == ca);
		const size_t ae = i;
		const size_t bb = j;
		do {
			++j;
		} while (j < m && (b[j] >> obj_bits) == cb);
		const size_t be = j;
		if (ignore_b_conflict && be - bb > 1)
			continue;
		out->XnY_ctx++;
		int mn = ORC_MIN_INIT;
		for (size_t x = ab; x < ae && mn; ++x) {
			const uint32_t oa = (uint32_t)(a[x] & objmask);
			for (size_t y = bb; y < be; ++y) {
				const uint32_t d = oa ^ (uint32_t)(b[y] & objmask);
				if (d == 0) {
					mn = 0;
					break;
				}
				const int c = orc_dna_popcount(d);
				if (c < mn)
					mn = c;
			}
		}
		if (mn > 0) {
			out->N_diff_obj++;
			out->N_diff_obj_section += (uint32_t)mn;
			if (mn > 1)
				out->N_mut2_ctx++;
		}
	}
}

static int orc_rec_cmp(const void *pa, const void *pb)
{
	/* kssdlib_sort.c:145-148: order by ctxgid, then obj */
	const orc_rec_t *a = pa, *b = pb;
	if (a->ctxgid != b->ctxgid)
		return a->ctxgid < b->ctxgid ? -1 : 1;
	return (a->obj > b->obj) - (a->obj < b->obj);
}

void orc_build_index(const uint64_t *keys, const uint64_t *skidx, int ngenomes,
		     unsigned obj_bits, orc_rec_t *out)
{
	/* sketch_rearrange.h:95-100 / sketch_rearrange.c:65-78 */
	const uint64_t objmask = (1ULL << obj_bits) - 1ULL;
	for (int g = 0; g < ngenomes; ++g)
		for (uint64_t i = skidx[g]; i < skidx[g + 1]; ++i) {
			out[i].ctxgid = (keys[i] >> obj_bits << ORC_GID_NBITS) | ((uint64_t)g & ORC_GIDMASK);
			out[i].obj = (uint32_t)(keys[i] & objmask);
		}
	qsort(out, (size_t)skidx[ngenomes], sizeof *out, orc_rec_cmp);
}

uint32_t orc_ctx_runs(const uint64_t *a, size_t n, unsigned obj_bits, int ignoreconflict)
{
	uint32_t runs = 0;
	for (size_t i = 0; i < n;) {
		const uint64_t c = a[i] >> obj_bits;
		const size_t b = i;
		do {
			++i;
		} while (i < n && (a[i] >> obj_bits) == c);
		if (!ignoreconflict || i - b == 1)
			runs++;
	}
	return runs;
}

void orc_ref_ctx_counts_from_index(const orc_rec_t *index, size_t n, int ref_gnum,
				   const uint64_t *ref_skidx, int ignoreconflict,
				   uint32_t *counts)
{
	for (int g = 0; g < ref_gnum; ++g)
		counts[g] = (uint32_t)(ref_skidx[g + 1] - ref_skidx[g]);
	for (size_t i = 0; i < n;) {
		const uint64_t cg = index[i].ctxgid;
		const uint32_t gid = (uint32_t)(cg & ORC_GIDMASK);
		size_t run = 0;
		do {
			++i;
			++run;
		} while (i < n && index[i].ctxgid == cg);
		if (run > 1)
			counts[gid] -= (uint32_t)(ignoreconflict ? run : run - 1);
	}
}

/* ------------------------------------------------------------------ FP64 */

double orc_naive_dist(const orc_features_t *f)
{
	/* command_ani.h:200-224, NUM_CODENS == 11 constants */
	if (f->XnY_ctx == 0)
		return 1;
	const double ratio = ((double)f->N_diff_obj_section + ORC_EPS) / ((double)f->N_diff_obj + ORC_EPS);
	const double dist0 = (double)f->N_diff_obj / (double)(uint32_t)(f->XnY_ctx + f->N_diff_obj);
	const double fd = 1 - pow(1 - dist0, ratio);
	if (fd <= 0)
		return 0;
	double w = fd / 0.01;
	if (w > 1.0)
		w = 1.0;
	return fd * 0.1544286 + 7.133e-4 * w;
}

double orc_p_dist(const orc_features_t *f, unsigned obj_bits)
{
	/* command_ani.h:226-232 */
	if (f->XnY_ctx == 0 || obj_bits == 0)
		return 1;
	return (double)f->N_diff_obj_section / ((double)f->XnY_ctx * ((double)obj_bits / 2.0));
}

static double orc_fold_scale(int drfold)
{
	/* model_ani.h:233-241, reference fold 8 */
	const int delta = drfold - 8;
	if (delta < -60)
		return ldexp(1.0, -60);
	if (delta > 60)
		return ldexp(1.0, 60);
	return ldexp(1.0, delta);
}

/* model_ani.h:16-20, :29-33, :56-74, :94-116 (NUM_CODENS == 11 sets) */
static const double orc_p_main[12] = {
	1.0178887, 1.0135146, 1.0302181, 1.1205122,
	1.0136223, 0.9803396, 1.0021964, 0.9839395,
	0.9988935, 0.9690589, 0.9888171, 0.9817371};
static const double orc_p_n95[12] = {
	1.0145656, 1.2708196, 1.2169829, 0.3904416,
	0.8147527, 1.3389150, 2.0581364, -1.4624736,
	1.6877678, 2.4390067, 0.7878482, 0.6697437};
static const double orc_c_main[17] = {
	6.86905236142636, -5.85271800517096e-08, 7.43351941016881e-06,
	-1.95859228059721e-05, -8.44336369105416e-06, -11726.5478433625,
	-111109.682632437, 121000.895632267, -741.538955661909,
	-1894.53741972785, 2585.81964981999, -7744.5216892858,
	-18190.2514369096, 23306.3090989115, 8352.82760552325,
	19782.4294837312, -25518.2493259614};
static const double orc_c_n95[17] = {
	-1.69173585242774, -3.64716315423252e-08, 2.85880264052975e-05,
	-5.5597924264054e-05, -2.96208613804271e-05, 1106.96550282022,
	-83.197177767225, -1671.88720051458, 152.75590068948,
	212.734805232395, 208.388739816682, -0.27860125939004,
	2.27065806912198, -10.2977013026837, -250.681324281127,
	-403.751360812353, -137.96628693833};
/* model_ani.h:183-213 */
static const double orc_p_af[15] = {
	2.36775520439829, 2.17036643548818, -0.0977124271437818, -0.251063182363729, 2.19491057829433,
	1.41850238263675, 2.15224961402778, -2.18587619877498, -2.21943802445092, 0.418474287083788,
	0.990424249133548, 0.502163745411501, 5.7787290524088, 0.626839640715682, 0.638706160147675};
static const double orc_c_af[21] = {
	0.127498260534708, -1.51730137145957e-06, 9.48114616179122e-05,
	-0.000257916039708839, -7.85253629216324e-05, 1.96482915215685e-06,
	-1475.58996322769, 26.4698377745167, 529.387641078915,
	42.0602258298213, 92.3637001980455, -306.00755719743,
	-86.3482066935661, 2.16630645286942, -0.674746089642152,
	48.1147600903939, 3.39101087242059, -15.5481688297158,
	-26.7068442368566, 18.5142668221724, 20.4430437567379};

static double orc_lm_core(double x, double s, double m2, double d, const double *p, const double *c)
{
	/* model_ani.h:243-273 (evaluation order kept: sums left to right) */
	const double d1 = 1 / (p[0] * x + p[1] * s + p[2] * m2 + p[3] * d + ORC_EPS);
	const double d2 = 1 / (p[4] * x + p[5] * s + p[6] * m2 + p[7] * d + ORC_EPS);
	const double d3 = 1 / (p[8] * x + p[9] * s + p[10] * m2 + p[11] * d + ORC_EPS);
	return c[0] + c[1] * x + c[2] * s + c[3] * m2 + c[4] * d + c[5] * d1 + c[6] * d2 +
	       c[7] * d3 + c[8] * x * d1 + c[9] * s * d1 + c[10] * m2 * d1 + c[11] * x * d2 +
	       c[12] * s * d2 + c[13] * m2 * d2 + c[14] * x * d3 + c[15] * s * d3 + c[16] * m2 * d3;
}

double orc_lm3ways_dist(const orc_features_t *f, int drfold)
{
	/* model_ani.h:285-305 */
	const double sc = orc_fold_scale(drfold);
	const double x = (double)f->XnY_ctx * sc, s = (double)f->N_diff_obj_section * sc;
	const double m2 = (double)f->N_mut2_ctx * sc, d = (double)f->N_diff_obj * sc;
	if (x <= 0.0)
		return 1;
	if (d <= 0.0)
		return 0;
	double dist = orc_lm_core(x, s, m2, d, orc_p_main, orc_c_main);
	if (dist < 0.05)
		dist = orc_lm_core(x, s, m2, d, orc_p_n95, orc_c_n95);
	if (dist < 0)
		dist = 0;
	return dist;
}

double orc_af_anib(const orc_features_t *f, int drfold)
{
	/* model_ani.h:307-367 */
	const double sc = orc_fold_scale(drfold);
	const double x = (double)f->XnY_ctx * sc, s = (double)f->N_diff_obj_section * sc;
	const double m2 = (double)f->N_mut2_ctx * sc, d = (double)f->N_diff_obj * sc;
	const double X = (double)f->X_ctx * sc;
	const double *p = orc_p_af, *c = orc_c_af;
	const double d1 = 1 / (p[0] * x + p[1] * s + p[2] * m2 + p[3] * d + p[4] * X + ORC_EPS);
	const double d2 = 1 / (p[5] * x + p[6] * s + p[7] * m2 + p[8] * d + p[9] * X + ORC_EPS);
	const double d3 = 1 / (p[10] * x + p[11] * s + p[12] * m2 + p[13] * d + p[14] * X + ORC_EPS);
	return c[0] + c[1] * x + c[2] * s + c[3] * m2 + c[4] * d + c[5] * X + c[6] * d1 +
	       c[7] * d2 + c[8] * d3 + c[9] * x * d1 + c[10] * s * d1 + c[11] * m2 * d1 +
	       c[12] * d * d1 + c[13] * x * d2 + c[14] * s * d2 + c[15] * m2 * d2 +
	       c[16] * d * d2 + c[17] * x * d3 + c[18] * s * d3 + c[19] * m2 * d3 + c[20] * d * d3;
}

double orc_mashD(int k, uint32_t X, uint32_t Y, uint32_t XnY)
{
	/* command_matrix.c:26-29 with command_matrix.h:10-12 (unsigned union) */
	const double j = (double)XnY / (double)(uint32_t)(X + Y - XnY);
	return -log(2 * j / (1 + j)) / (double)(uint32_t)k;
}

double orc_aafD(int k, uint32_t X, uint32_t Y, uint32_t XnY)
{
	/* command_matrix.c:30-33 with command_matrix.h:13 */
	const uint32_t mn = X < Y ? X : Y;
	return -log((double)XnY / (double)mn) / (double)(uint32_t)k;
}

void orc_matrix_values(const orc_ctx_t *ctx, const orc_obj_t *obj, int nq, int nr,
		       const uint32_t *qctx, const uint32_t *rctx, int metric, int values_ani,
		       double afcut, double e, int drfold, unsigned obj_bits, unsigned ctx_bits,
		       double *out, int nthreads)
{
	/* make_selected_output_row for metrics 3..9 in the matrix outputs
	 * (command_ani.c:1321-1379, :1458-1496): AF gate, exception cells. */
	const uint32_t k = ctx_bits / 2;
	if (nthreads <= 0)
		nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (long c = 0; c < (long)nq * nr; ++c) {
		const int q = (int)(c / nr), r = (int)(c % nr);
		if (qctx[q] == 0 || rctx[r] == 0) {
			out[c] = values_ani ? 1.0 - e : e;
			continue;
		}
		const orc_features_t f = {ctx[c].num_ctx, 0, obj[c].diff_obj_section, ctx[c].num_mut2_ctx,
					  obj[c].diff_obj};
		const double afq = (double)f.XnY_ctx / qctx[q], afr = (double)f.XnY_ctx / rctx[r];
		const int pass = (afq > afr ? afq : afr) >= afcut;
		double d;
		switch (metric) {
		case 4: d = pass ? orc_naive_dist(&f) : e; break;
		case 5: d = orc_mashD((int)k, rctx[r], qctx[q], f.XnY_ctx); break;
		case 6: d = orc_aafD((int)k, rctx[r], qctx[q], f.XnY_ctx); break;
		case 7: d = pass ? orc_lm3ways_dist(&f, drfold) : orc_mashD((int)k, rctx[r], qctx[q], f.XnY_ctx); break;
		case 8: d = pass ? orc_lm3ways_dist(&f, drfold) : orc_aafD((int)k, rctx[r], qctx[q], f.XnY_ctx); break;
		case 9: d = pass ? orc_p_dist(&f, obj_bits) : e; break;
		default: d = pass ? orc_lm3ways_dist(&f, drfold) : e; break;
		}
		out[c] = values_ani ? 1.0 - d : d;
	}
}

/* ------------------------------------------------------------ index-path ranking */
typedef struct {
	uint32_t id;
	float ani;
} orc_idani_t; /* idani_t, command_ani.h:353-356 */

static int orc_cmp_idani_desc(const void *a, const void *b)
{
	/* compare_idani_desc, command_ani.h:361-366 */
	const orc_idani_t *x = a, *y = b;
	return (x->ani < y->ani) - (x->ani > y->ani);
}

static void orc_heap_down(orc_idani_t *h, int n, int i)
{
	/* heap_sift_down_min_idani, command_ani.h:369-383 */
	for (;;) {
		const int l = 2 * i + 1, r = l + 1;
		int s = i;
		if (l < n && h[l].ani < h[s].ani)
			s = l;
		if (r < n && h[r].ani < h[s].ani)
			s = r;
		if (s == i)
			break;
		const orc_idani_t t = h[i];
		h[i] = h[s];
		h[s] = t;
		i = s;
	}
}

static int orc_topN(orc_idani_t *arr, int m, int topN)
{
	/* idani_select_topN_sorted + partial_topN_inplace_idani,
	 * command_ani.h:391-433; qsort is the C library's, as the reference's */
	if (m <= 0 || topN <= 0)
		return 0;
	if (topN > m)
		topN = m;
	if (m <= 128 || topN >= m / 2) {
		qsort(arr, (size_t)m, sizeof *arr, orc_cmp_idani_desc);
		return topN;
	}
	if (m > 1) {
		for (int i = (topN >> 1) - 1; i >= 0; --i)
			orc_heap_down(arr, topN, i);
		for (int i = topN; i < m; ++i) {
			if (arr[i].ani <= arr[0].ani)
				continue;
			const orc_idani_t t = arr[i];
			arr[i] = arr[0];
			arr[0] = t;
			orc_heap_down(arr, topN, 0);
		}
		qsort(arr, (size_t)topN, sizeof *arr, orc_cmp_idani_desc);
	}
	return topN;
}

typedef struct {
	double ani;
	uint32_t id;
} orc_urank_t;

static int orc_cmp_urank(const void *a, const void *b)
{
	/* compare_unique_ctx_rank, command_ani.h:456-461 */
	const orc_urank_t *x = a, *y = b;
	if (x->ani != y->ani)
		return x->ani < y->ani ? 1 : -1;
	return (x->id > y->id) - (x->id < y->id);
}

static double orc_naive_of(orc_ctx_t c, orc_obj_t o)
{
	/* ani_context_naive_dist, command_ani.h:442-450 */
	orc_features_t f = {0};
	f.XnY_ctx = c.num_ctx;
	f.N_mut2_ctx = c.num_mut2_ctx;
	f.N_diff_obj = o.diff_obj;
	f.N_diff_obj_section = o.diff_obj_section;
	return orc_naive_dist(&f);
}

void orc_index_rank(const orc_ctx_t *ctx, const orc_obj_t *obj, const orc_ctx_t *uctx, const orc_obj_t *uobj,
		    int nq, int R, const uint32_t *ref_ctx, const uint32_t *qry_ctx, int ctxcut_opt, float afcut,
		    float anicut, int ntop_opt, int unique_ctx, uint32_t *ids, uint32_t *pass_ids, double *rank,
		    uint32_t *n_pass, uint32_t *n_out)
{
	/* the fmt == 0 branch of count_ctx_obj_frm_comb_sketch_section_with_unique,
	 * command_ani.h:577-621 */
	const uint32_t ctxcut = (uint32_t)ctxcut_opt;
	const int ntop = ntop_opt > 0 ? ntop_opt : R;
	orc_idani_t *ida = malloc((size_t)(R > 0 ? R : 1) * sizeof *ida);
	orc_urank_t *ranks = unique_ctx ? malloc((size_t)(R > 0 ? R : 1) * sizeof *ranks) : NULL;
	for (int i = 0; i < nq; ++i) {
		const size_t row = (size_t)i * (size_t)R;
		int np = 0;
		for (int j = 0; j < R; ++j) {
			const uint32_t shared = ctx[row + j].num_ctx;
			if (shared < ctxcut)
				continue;
			if (ref_ctx[j] == 0 || qry_ctx[i] == 0)
				continue;
			const double afq = (double)shared / (double)qry_ctx[i], afr = (double)shared / (double)ref_ctx[j];
			if (!((afq > afr ? afq : afr) >= (double)afcut))
				continue;
			const float dist = (float)obj[row + j].diff_obj / ctx[row + j].num_ctx / 7;
			double ani = (float)(1 - dist);
			if (ranks) {
				/* ani_use_unique_context, command_ani.h:437-440 */
				const int use = (uint64_t)uctx[row + j].num_ctx * 20 >= ref_ctx[j];
				ani = 1.0 - orc_naive_of(use ? uctx[row + j] : ctx[row + j], use ? uobj[row + j] : obj[row + j]);
			}
			if (ani < anicut)
				continue;
			if (ranks)
				ranks[np] = (orc_urank_t){ani, (uint32_t)j};
			ida[np].id = (uint32_t)j;
			ida[np].ani = (float)ani;
			rank[row + (size_t)np] = ani;
			pass_ids[row + (size_t)np] = (uint32_t)j;
			np++;
		}
		n_pass[i] = (uint32_t)np;
		int out;
		if (ranks) {
			qsort(ranks, (size_t)np, sizeof *ranks, orc_cmp_urank);
			out = np < ntop ? np : ntop;
			for (int r = 0; r < out; ++r)
				ida[r].id = ranks[r].id;
		} else {
			out = orc_topN(ida, np, ntop);
		}
		for (int r = 0; r < out; ++r)
			ids[row + (size_t)r] = ida[r].id;
		n_out[i] = (uint32_t)out;
	}
	free(ida);
	free(ranks);
}

/* ------------------------------------------------------------------ one-query stream path */

static inline uint64_t orc_fmix64(uint64_t k)
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdULL;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ULL;
	k ^= k >> 33;
	return k;
}

int orc_stream_one_query(const uint64_t *qry, size_t qlen, const uint64_t *ref_keys, const uint64_t *ref_idx, int R,
			 unsigned obj_bits, int ignoreconflict, int ref_conflict, int ctxcut, float afcut,
			 orc_ctx_t *ctx, orc_obj_t *obj, uint32_t *ref_ctx_out, uint8_t *pass, uint32_t *qry_ctx_out,
			 int nthreads)
{
	const uint64_t objmask = obj_bits >= 64 ? ~0ULL : ((1ULL << obj_bits) - 1ULL);
	memset(ctx, 0, (size_t)R * sizeof *ctx);
	memset(obj, 0, (size_t)R * sizeof *obj);
	memset(ref_ctx_out, 0, (size_t)R * sizeof *ref_ctx_out);
	memset(pass, 0, (size_t)R);
	/* build_qry_ctx_hash_lookup (command_ani.c:2024-2051): key ctx + 1,
	 * value the run [beg, end), capacity the power of two >= 2 qlen */
	uint32_t cap = 1;
	while (cap < 2 * qlen)
		cap <<= 1;
	uint64_t *hk = calloc(cap, sizeof *hk);
	uint64_t *hv = malloc((size_t)cap * sizeof *hv);
	if (!hk || !hv)
		abort();
	const uint32_t mask = cap - 1;
	uint32_t qry_ctx = 0;
	for (size_t i = 0; i < qlen;) {
		const uint64_t c = qry[i] >> obj_bits;
		const size_t beg = i;
		do {
			++i;
		} while (i < qlen && (qry[i] >> obj_bits) == c);
		uint32_t pos = (uint32_t)orc_fmix64(c + 1) & mask;
		while (hk[pos] && hk[pos] != c + 1)
			pos = (pos + 1) & mask;
		hk[pos] = c + 1;
		hv[pos] = ((uint64_t)beg << 32) | (uint64_t)i;
		++qry_ctx;
	}
	*qry_ctx_out = qry_ctx;
	if (nthreads <= 0)
		nthreads = 1;
	/* compute_streamed_ref_row_one_qry (command_ani.c:2110-2201) per
	 * reference, dynamic schedule as the block loop (:2607-2622) */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
	for (int rn = 0; rn < R; ++rn) {
		const uint64_t *ref = ref_keys + ref_idx[rn] - ref_idx[0];
		const size_t ref_len = (size_t)(ref_idx[rn + 1] - ref_idx[rn]);
		if (qry_ctx == 0 || ref_len == 0)
			continue;
		uint32_t total = 0; /* count_ctx_runs_sorted_ctxobj64_local (:337-349) */
		if (ref_conflict) {
			for (size_t j = 0; j < ref_len;) {
				const uint64_t c = ref[j] >> obj_bits;
				const size_t b = j;
				do {
					++j;
				} while (j < ref_len && (ref[j] >> obj_bits) == c);
				if (!ignoreconflict || j - b == 1)
					++total;
			}
		} else {
			total = (uint32_t)ref_len;
		}
		if (total == 0)
			continue;
		/* ani_report_af_needed_ctx (command_ani.h:80-90) */
		const uint32_t denom = qry_ctx < total ? qry_ctx : total;
		const double need_d = (double)afcut * (double)denom;
		const uint32_t need = need_d <= 0.0 ? 0 : (uint32_t)ceil(need_d);
		uint32_t xny = 0, mut2 = 0, dobj = 0, dsec = 0, walked = 0;
		int gave_up = 0;
		for (size_t j = 0; j < ref_len;) {
			const uint64_t c = ref[j] >> obj_bits;
			const size_t rb = j;
			do {
				++j;
			} while (j < ref_len && (ref[j] >> obj_bits) == c);
			if (ignoreconflict && j - rb > 1)
				continue;
			++walked;
			if (xny + (total - walked) < need) { /* :2155-2156 */
				gave_up = 1;
				break;
			}
			uint32_t pos = (uint32_t)orc_fmix64(c + 1) & mask;
			while (hk[pos] && hk[pos] != c + 1)
				pos = (pos + 1) & mask;
			if (!hk[pos])
				continue;
			const size_t qb = (size_t)(hv[pos] >> 32), qe = (size_t)(uint32_t)hv[pos];
			++xny;
			/* min_diff_sections_ctxobj64_runs (command_ani.h:283-299) */
			int m = ORC_MIN_INIT;
			for (size_t a = qb; a < qe && m; ++a)
				for (size_t b = rb; b < j; ++b) {
					const uint32_t x = (uint32_t)(qry[a] & objmask) ^ (uint32_t)(ref[b] & objmask);
					if (!x) {
						m = 0;
						break;
					}
					const int d = orc_dna_popcount(x);
					if (d < m)
						m = d;
				}
			if (m > 0) {
				++dobj;
				dsec += (uint32_t)m;
				if (m > 1)
					++mut2;
			}
		}
		ctx[rn].num_ctx = xny;
		ctx[rn].num_mut2_ctx = mut2;
		obj[rn].diff_obj = dobj;
		obj[rn].diff_obj_section = dsec;
		ref_ctx_out[rn] = walked;
		if (gave_up || walked == 0 || xny < (uint32_t)ctxcut || xny < need)
			continue;
		const double afq = (double)xny / (double)qry_ctx, afr = (double)xny / (double)walked;
		pass[rn] = (afq > afr ? afq : afr) >= (double)afcut; /* ani_report_af_pass (command_ani.h:75-78) */
	}
	free(hk);
	free(hv);
	return 0;
}
