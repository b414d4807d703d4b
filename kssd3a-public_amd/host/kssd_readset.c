/*
 * k-mer count filters of the read-set sketch modes: the -n lower cutoff,
 * the count-weighted percentile cutoff, the inferred read-QC count range
 * and the cap (apply_kmer_count_filters, command_sketch.c:717-767).
 *
 * The read-QC range comes from the histogram of per-key counts (how many
 * keys were seen c times): a coverage peak is chosen, a negative-binomial
 * (or Poisson) shape is fitted around it, and the range is the run of
 * contiguous counts whose frequencies stay within 3x of that shape
 * (infer_reads_qc_count_range, command_sketch.c:585-669, and its helpers
 * :250-583).  The arithmetic is long double, in the order the reference
 * evaluates it, so that the chosen bins are the same.
 */
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kssd_host.h"

/* ---------------------------------------------------------- histogram */
typedef struct {
	uint32_t *c; /* distinct counts, ascending */
	uint64_t *f; /* keys with that count */
	size_t len;
} hist_t;

static int cmp_u32(const void *a, const void *b)
{
	const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
	return (x > y) - (x < y);
}

/* counts >= min_count, sorted ascending (a copy; NULL when none) */
static uint32_t *selected_counts(const uint32_t *counts, size_t n, uint32_t min_count, size_t *m)
{
	uint32_t *v = malloc((n ? n : 1) * sizeof *v);
	if (!v)
		abort();
	size_t k = 0;
	for (size_t i = 0; i < n; ++i)
		if (counts[i] >= min_count)
			v[k++] = counts[i];
	if (!k) {
		free(v);
		*m = 0;
		return NULL;
	}
	qsort(v, k, sizeof *v, cmp_u32);
	*m = k;
	return v;
}

/* build_count_histogram, command_sketch.c:250-296 */
static hist_t hist_build(const uint32_t *counts, size_t n, uint32_t min_count)
{
	hist_t h = {0};
	size_t m;
	uint32_t *v = selected_counts(counts, n, min_count, &m);
	if (!v)
		return h;
	uint64_t *f = malloc(m * sizeof *f);
	if (!f)
		abort();
	size_t b = 0;
	f[0] = 1;
	for (size_t i = 1; i < m; ++i) {
		if (v[i] == v[b]) {
			f[b]++;
		} else {
			v[++b] = v[i];
			f[b] = 1;
		}
	}
	h.c = v;
	h.f = f;
	h.len = b + 1;
	return h;
}

static inline bool adjacent(const hist_t *h, size_t i) /* bins i and i+1 are counts c, c+1 */
{
	return h->c[i] + 1 == h->c[i + 1];
}

/* frequency of bin i averaged with its contiguous neighbours, the bin
 * itself weighted 2 (histogram_smooth_freq, :351-369) */
static long double smooth(const hist_t *h, size_t i)
{
	if (h->len == 1)
		return (long double)h->f[0];
	long double sum = 2.0L * (long double)h->f[i], w = 2.0L;
	if (i > 0 && adjacent(h, i - 1)) {
		sum += (long double)h->f[i - 1];
		w += 1.0L;
	}
	if (i + 1 < h->len && adjacent(h, i)) {
		sum += (long double)h->f[i + 1];
		w += 1.0L;
	}
	return sum / w;
}

static bool is_valley(const hist_t *h, size_t i) /* :371-383 */
{
	if (i == 0 || i + 1 >= h->len || !adjacent(h, i - 1) || !adjacent(h, i))
		return false;
	const long double l = smooth(h, i - 1), m = smooth(h, i), r = smooth(h, i + 1);
	return m <= l && m <= r;
}

static bool is_peak(const hist_t *h, size_t i) /* :385-396 */
{
	const long double m = smooth(h, i);
	const long double l = i > 0 ? smooth(h, i - 1) : -1.0L;
	const long double r = i + 1 < h->len ? smooth(h, i + 1) : -1.0L;
	if (i > 0 && !adjacent(h, i - 1))
		return false;
	if (i + 1 < h->len && !adjacent(h, i))
		return false;
	return m >= l && m >= r;
}

/* first bin of the largest smoothed frequency in [a, b] (:398-410) */
static size_t highest_bin(const hist_t *h, size_t a, size_t b)
{
	size_t best = a;
	long double bf = -1.0L;
	for (size_t i = a; i <= b && i < h->len; ++i) {
		const long double v = smooth(h, i);
		if (v > bf) {
			bf = v;
			best = i;
		}
	}
	return best;
}

/* first bin of the smallest smoothed frequency in [a, b] (:453-465) */
static size_t lowest_bin(const hist_t *h, size_t a, size_t b)
{
	size_t best = a;
	long double bf = smooth(h, a);
	for (size_t i = a + 1; i <= b && i < h->len; ++i) {
		const long double v = smooth(h, i);
		if (v < bf) {
			bf = v;
			best = i;
		}
	}
	return best;
}

/* count at which the count-weighted mass reaches q of the total (:412-433) */
static uint32_t weighted_quantile(const hist_t *h, long double q)
{
	long double total = 0.0L;
	for (size_t i = 0; i < h->len; ++i)
		total += (long double)h->c[i] * (long double)h->f[i];
	if (total <= 0.0L)
		return h->c[0];
	const long double target = total * q;
	long double cum = 0.0L;
	for (size_t i = 0; i < h->len; ++i) {
		cum += (long double)h->c[i] * (long double)h->f[i];
		if (cum >= target)
			return h->c[i];
	}
	return h->c[h->len - 1];
}

static size_t bin_at_least(const hist_t *h, uint32_t c) /* :435-442 */
{
	for (size_t i = 0; i < h->len; ++i)
		if (h->c[i] >= c)
			return i;
	return h->len - 1;
}

static size_t bin_at_most(const hist_t *h, uint32_t c) /* :444-451 */
{
	for (size_t i = h->len; i > 0; --i)
		if (h->c[i - 1] <= c)
			return i - 1;
	return 0;
}

typedef struct {
	size_t mode, lower_limit;
	uint32_t median;
	bool anchored;
} peak_t;

/* The coverage peak.  First choice: the highest local maximum between the
 * count-weighted quartiles, the error valley being the lowest bin below it
 * (select_weighted_anchor_peak, :467-514).  When the quartiles give no
 * span: the right-hand peak after a valley with the best frequency x
 * log1p(count) among those standing clear of noise (select_reads_qc_peak,
 * :516-562). */
static peak_t find_peak(const hist_t *h)
{
	peak_t p = {0};
	p.mode = highest_bin(h, 0, h->len - 1);
	const uint32_t q25 = weighted_quantile(h, 0.25L);
	const uint32_t q50 = weighted_quantile(h, 0.50L);
	const uint32_t q75 = weighted_quantile(h, 0.75L);
	p.median = q50;
	if (q50 && q25 <= q75) {
		const size_t a = bin_at_least(h, q25), b = bin_at_most(h, q75);
		if (a <= b) {
			size_t best = a;
			long double bf = -1.0L;
			bool found = false;
			for (size_t i = a; i <= b && i < h->len; ++i) {
				if (!is_peak(h, i))
					continue;
				const long double v = smooth(h, i);
				if (v > bf) {
					bf = v;
					best = i;
					found = true;
				}
			}
			if (!found)
				best = highest_bin(h, a, b);
			p.mode = best;
			p.lower_limit = 0;
			p.anchored = true;
			if (best > 1)
				p.lower_limit = lowest_bin(h, 0, best - 1) + 1;
			return p;
		}
	}
	const long double global = smooth(h, p.mode);
	long double best_score = -1.0L;
	for (size_t v = 1; v + 1 < h->len; ++v) {
		if (!is_valley(h, v))
			continue;
		const size_t r = highest_bin(h, v + 1, h->len - 1);
		if (r <= v)
			continue;
		const long double vf = smooth(h, v), rf = smooth(h, r);
		if (rf < 25.0L || rf < global * 0.001L || rf < vf * 1.5L)
			continue;
		if (!is_peak(h, r) && r + 1 < h->len)
			continue;
		const long double score = rf * log1pl((long double)h->c[r]);
		if (score > best_score) {
			best_score = score;
			p.mode = r;
			p.lower_limit = v + 1;
		}
	}
	return p;
}

/* log P(count) under Poisson(mean) (size <= 0) or the negative binomial of
 * that mean and size (log_count_pmf, :337-349) */
static long double log_pmf(uint32_t count, long double mean, long double size)
{
	const long double x = (long double)count;
	if (mean <= 0.0L)
		return -INFINITY;
	if (size <= 0.0L)
		return -mean + x * logl(mean) - lgammal(x + 1.0L);
	const long double p = size / (size + mean);
	return lgammal(x + size) - lgammal(size) - lgammal(x + 1.0L) + size * logl(p) + x * logl(1.0L - p);
}

/* bin i's frequency relative to the mode's within 3x of the fitted
 * shape's ratio (histogram_bin_is_nb_like, :564-583) */
static bool fits_shape(const hist_t *h, size_t i, size_t mode, long double mean, long double size)
{
	const long double seen = (long double)h->f[i] / (long double)h->f[mode];
	const long double log_ratio = log_pmf(h->c[i], mean, size) - log_pmf(h->c[mode], mean, size);
	if (!isfinite((double)log_ratio))
		return false;
	const long double expect = expl(log_ratio);
	if (expect < 0.005L && seen < 0.005L)
		return false;
	return seen >= expect / 3.0L && seen <= expect * 3.0L;
}

typedef struct {
	bool valid;
	uint32_t lower, upper, mode;
} qc_range_t;

static qc_range_t infer_range(const uint32_t *counts, size_t n, uint32_t min_count)
{
	qc_range_t r = {0};
	hist_t h = hist_build(counts, n, min_count);
	if (!h.len)
		return r;
	const peak_t p = find_peak(&h);
	const size_t mode = p.mode, floor_bin = p.lower_limit;

	/* fit window: contiguous bins holding >= 1/20 of the peak (at least 2) */
	const uint64_t min_f = h.f[mode] / 20 > 2 ? h.f[mode] / 20 : 2;
	size_t fl = mode, fr = mode;
	while (fl > floor_bin && adjacent(&h, fl - 1) && h.f[fl - 1] >= min_f)
		--fl;
	while (fr + 1 < h.len && adjacent(&h, fr) && h.f[fr + 1] >= min_f)
		++fr;
	while (fr - fl + 1 < 3 && (fl > floor_bin || fr + 1 < h.len)) {
		if (fl > floor_bin)
			--fl;
		if (fr + 1 < h.len)
			++fr;
	}
	/* weighted mean and variance of the window (histogram_weighted_stats, :298-328) */
	uint64_t mass = 0;
	long double sum = 0.0L, sumsq = 0.0L;
	for (size_t i = fl; i <= fr; ++i) {
		const long double f = (long double)h.f[i], c = (long double)h.c[i];
		mass += h.f[i];
		sum += c * f;
		sumsq += c * c * f;
	}
	if (mass) {
		const long double mean = sum / (long double)mass;
		long double var = sumsq / (long double)mass - mean * mean;
		if (var < 0.0L)
			var = 0.0L;
		/* negative_binomial_size, :330-335 */
		const long double size = (mean <= 0.0L || var <= mean) ? 0.0L : (mean * mean) / (var - mean);

		size_t lo = mode, hi = mode;
		while (lo > floor_bin && adjacent(&h, lo - 1) && fits_shape(&h, lo - 1, mode, mean, size))
			--lo;
		while (hi + 1 < h.len && adjacent(&h, hi) && fits_shape(&h, hi + 1, mode, mean, size))
			++hi;
		if (lo == hi) {
			lo = fl;
			hi = fr;
		}
		if (p.anchored) {
			const size_t w = bin_at_least(&h, p.median);
			if (w < lo)
				lo = w < floor_bin ? floor_bin : w;
			if (w > hi)
				hi = w;
		}
		r.valid = true;
		r.lower = h.c[lo];
		r.upper = h.c[hi];
		r.mode = h.c[mode];
	}
	free(h.c);
	free(h.f);
	return r;
}

/* ---------------------------------------------------------- filters */
/* smallest count whose cumulative count-weighted mass (counts >= base,
 * ascending) reaches ceil(p * total) (kmer_count_percentile_threshold,
 * command_sketch.c:88-134) */
static uint32_t percentile_cutoff(const uint32_t *counts, size_t n, double p, uint32_t base)
{
	size_t m;
	uint32_t *v = selected_counts(counts, n, base, &m);
	if (!v)
		return base;
	uint64_t total = 0;
	for (size_t i = 0; i < m; ++i)
		total = UINT64_MAX - total < v[i] ? UINT64_MAX : total + v[i];
	const long double want = ceill((long double)p * (long double)total);
	uint64_t rank = want >= (long double)UINT64_MAX ? UINT64_MAX : (uint64_t)want;
	if (rank < 1)
		rank = 1;
	uint32_t t = v[m - 1];
	uint64_t cum = 0;
	for (size_t i = 0; i < m; ++i) {
		cum = UINT64_MAX - cum < v[i] ? UINT64_MAX : cum + v[i];
		if (cum >= rank) {
			t = v[i];
			break;
		}
	}
	free(v);
	return t;
}

int kssd_count_opts_need_counts(const kssd_count_opts_t *o)
{
	return o->record_qc || o->min_count > 1 || o->percentile > 0.0 || o->reads_qc;
}

kssd_count_filter_t kssd_apply_count_filters(uint64_t *keys, uint32_t *counts, size_t *n,
					     const kssd_count_opts_t *o)
{
	kssd_count_filter_t r = {0};
	if (!*n || !counts)
		return r;
	uint32_t cut = o->min_count > 1 ? (uint32_t)o->min_count : 1u;
	if (o->percentile > 0.0) {
		const uint32_t pc = percentile_cutoff(counts, *n, o->percentile, cut);
		if (pc > cut)
			cut = pc;
	}
	uint32_t upper = 0;
	if (o->record_qc || o->reads_qc) {
		const uint32_t range_min = cut < 2 ? 2u : cut;
		if (o->reads_qc && cut < 2)
			cut = 2;
		const qc_range_t q = infer_range(counts, *n, range_min);
		if (q.valid) {
			r.mode = q.mode;
			r.qc.flags = KSSD_QC_RANGE_VALID | (o->reads_qc ? KSSD_QC_RANGE_APPLIED : 0u);
			r.qc.lower = q.lower;
			r.qc.upper = q.upper;
			r.qc.mode = q.mode;
			if (o->reads_qc) {
				if (q.lower > cut)
					cut = q.lower;
				upper = q.upper;
				r.used_reads_qc = 1;
			}
		}
	}
	r.raw_cutoff = cut;
	if (o->cap > 0 && cut > (uint32_t)o->cap)
		cut = (uint32_t)o->cap;
	r.cutoff = cut;
	r.upper = upper;
	if (cut > 1 || upper) {
		size_t w = 0;
		for (size_t i = 0; i < *n; ++i) {
			if (counts[i] >= cut && (!upper || counts[i] <= upper)) {
				keys[w] = keys[i];
				counts[w] = counts[i];
				++w;
			}
		}
		*n = w;
	}
	return r;
}

void kssd_print_count_filter(FILE *fp, const kssd_count_opts_t *o, const char *name,
			     const kssd_count_filter_t *r, int leading_newline)
{
	if (leading_newline)
		fputc('\n', fp);
	if (r->used_reads_qc) {
		fprintf(fp, "--readsQC count range for %s: mode=%u lower=%u upper=%u", name, r->mode, r->cutoff,
			r->upper);
		if (o->cap > 0)
			fprintf(fp, " ncap=%u", (uint32_t)o->cap);
		fputc('\n', fp);
	} else if (o->reads_qc && o->percentile <= 0.0) {
		fprintf(fp, "--readsQC count range for %s: insufficient count histogram; lower=%u\n", name,
			r->cutoff);
	} else if (o->cap > 0) {
		fprintf(fp, "--npercentile %.6g count cutoff for %s: raw=%u capped=%u effective=%u\n",
			o->percentile, name, r->raw_cutoff, (uint32_t)o->cap, r->cutoff);
	} else {
		fprintf(fp, "--npercentile %.6g count cutoff for %s: %u\n", o->percentile, name, r->cutoff);
	}
}

int kssd_sketch_pooled(char *const *paths, int n, int drfold, int threads, int keep_conflict,
		       const kssd_count_opts_t *co, kssd_u64vec_t *keys, uint32_t **counts, kssd_count_filter_t *filter,
		       int *bad)
{
	/* sketch_few_files_with_intrafile_parallel (command_sketch.c:5435-5497)
	 * with finalize_filter_conflict_and_write (:5194-5214) */
	memset(keys, 0, sizeof *keys);
	if (counts)
		*counts = NULL;
	if (filter)
		memset(filter, 0, sizeof *filter);
	kssd_u64vec_t *part = calloc((size_t)(n > 0 ? n : 1), sizeof *part);
	if (!part)
		return -1;
	if (kssd_collect_files(paths, n, drfold, threads, part, bad, NULL, NULL) != 0) {
		for (int i = 0; i < n; ++i)
			kssd_u64vec_free(&part[i]);
		free(part);
		return -1;
	}
	size_t total = 0;
	for (int i = 0; i < n; ++i)
		total += part[i].n;
	keys->a = malloc((total ? total : 1) * sizeof *keys->a);
	if (!keys->a) {
		for (int i = 0; i < n; ++i)
			kssd_u64vec_free(&part[i]);
		free(part);
		return -1;
	}
	keys->cap = total ? total : 1;
	for (int i = 0; i < n; ++i) {
		if (part[i].n)
			memcpy(keys->a + keys->n, part[i].a, part[i].n * sizeof *part[i].a);
		keys->n += part[i].n;
		kssd_u64vec_free(&part[i]);
	}
	free(part);
	if (!kssd_count_opts_need_counts(co)) {
		kssd_finalize_keys(keys, keep_conflict);
		return 0;
	}
	uint32_t *c = NULL;
	if (kssd_count_keys(keys, &c) != 0)
		return -1;
	const kssd_count_filter_t f = kssd_apply_count_filters(keys->a, c, &keys->n, co);
	if (filter)
		*filter = f;
	if (!keep_conflict)
		kssd_drop_conflicts(keys->a, c, &keys->n);
	if (counts)
		*counts = c;
	else
		free(c);
	return 0;
}
