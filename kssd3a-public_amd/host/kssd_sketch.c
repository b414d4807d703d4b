/*
 * Sketcher (sequence -> sorted context-object keys) and seeded genome
 * simulator.  See kssd_host.h for the contract.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include "kssd_host.h"

void kssd_u64vec_free(kssd_u64vec_t *v)
{
	free(v->a);
	v->a = NULL;
	v->n = v->cap = 0;
}

static int vec_reserve(kssd_u64vec_t *v, size_t want)
{
	if (want <= v->cap)
		return 0;
	size_t cap = v->cap ? v->cap : 4096;
	while (cap < want)
		cap <<= 1;
	uint64_t *a = realloc(v->a, cap * sizeof *a);
	if (!a)
		return -1;
	v->a = a;
	v->cap = cap;
	return 0;
}

/* SplitMix64 finalizer, the sketch selection hash (kssdlib_hash.h:18-23). */
static inline uint64_t mix64(uint64_t x)
{
	x += 0x9e3779b97f4a7c15ULL;
	x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
	x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
	return x ^ (x >> 31);
}

/* Extract the 44 context bits and 20 object bits of a 32-mer under the
 * coden mask: from the least significant end the mask alternates 4 context
 * bits and 2 object bits ten times, then 4 context bits (0xf3cf...cf). */
static inline uint64_t coden_key(uint64_t kmer)
{
	uint64_t ctx = 0, obj = 0;
	for (int i = 0; i < 10; ++i) {
		ctx |= (kmer & 0xFULL) << (4 * i);
		kmer >>= 4;
		obj |= (kmer & 0x3ULL) << (2 * i);
		kmer >>= 2;
	}
	ctx |= (kmer & 0xFULL) << 40;
	return (ctx << KSSD_OBJ_BITS) | obj;
}

void kssd_sketch_codes(const uint8_t *codes, size_t len, int drfold, kssd_u64vec_t *out)
{
	const uint32_t filter = drfold >= 32 ? 0u : (uint32_t)(UINT32_MAX >> drfold);
	const unsigned len_mv = 2 * KSSD_KLEN - 2;
	uint64_t fwd = 0, rev = 0;
	int base = 0;
	if (len < KSSD_KLEN)
		return;
	for (size_t pos = 0; pos < len; ++pos) {
		const uint8_t b = codes[pos];
		if (b > 3) {
			base = 0;
			fwd = rev = 0;
			continue;
		}
		fwd = (fwd << 2) | b;
		rev = (rev >> 2) | ((uint64_t)(b ^ 3u) << len_mv);
		if (++base < KSSD_KLEN)
			continue;
		const uint64_t tf = fwd & KSSD_CODEN_CTXMASK, tr = rev & KSSD_CODEN_CTXMASK;
		const int use_fwd = tf < tr;
		if ((uint32_t)mix64(use_fwd ? tf : tr) > filter)
			continue;
		if (out->n == out->cap && vec_reserve(out, out->n + 1))
			abort();
		out->a[out->n++] = coden_key(use_fwd ? fwd : rev);
	}
}

/* LSD radix sort of u64 (8 passes of 8 bits, skipping constant bytes). */
static void radix_sort_u64(uint64_t *a, size_t n)
{
	if (n < 64) {
		for (size_t i = 1; i < n; ++i) {
			uint64_t v = a[i];
			size_t j = i;
			while (j && a[j - 1] > v) {
				a[j] = a[j - 1];
				--j;
			}
			a[j] = v;
		}
		return;
	}
	uint64_t *tmp = malloc(n * sizeof *tmp);
	if (!tmp)
		abort();
	uint64_t *src = a, *dst = tmp;
	for (int shift = 0; shift < 64; shift += 8) {
		size_t cnt[256] = {0};
		for (size_t i = 0; i < n; ++i)
			cnt[(src[i] >> shift) & 0xFF]++;
		if (cnt[(src[0] >> shift) & 0xFF] == n)
			continue;
		size_t sum = 0;
		for (int b = 0; b < 256; ++b) {
			size_t c = cnt[b];
			cnt[b] = sum;
			sum += c;
		}
		for (size_t i = 0; i < n; ++i)
			dst[cnt[(src[i] >> shift) & 0xFF]++] = src[i];
		uint64_t *t = src;
		src = dst;
		dst = t;
	}
	if (src != a)
		memcpy(a, src, n * sizeof *a);
	free(tmp);
}

void kssd_finalize_keys(kssd_u64vec_t *v, int keep_conflict)
{
	if (!v->n)
		return;
	radix_sort_u64(v->a, v->n);
	size_t w = 1;
	for (size_t i = 1; i < v->n; ++i)
		if (v->a[i] != v->a[w - 1])
			v->a[w++] = v->a[i];
	v->n = w;
	if (keep_conflict)
		return;
	/* remove_ctx_with_conflict_obj_noabund, kssdlib_sort.c:388-407 */
	size_t i = 0;
	w = 0;
	while (i < v->n) {
		size_t c = 1;
		while (i + c < v->n && ((v->a[i] ^ v->a[i + c]) >> KSSD_OBJ_BITS) == 0)
			++c;
		if (c == 1)
			v->a[w++] = v->a[i];
		i += c;
	}
	v->n = w;
}

void kssd_ascii_to_codes(const char *s, size_t len, uint8_t *codes)
{
	for (size_t i = 0; i < len; ++i) {
		switch (s[i]) {
		case 'A': case 'a': codes[i] = 0; break;
		case 'C': case 'c': codes[i] = 1; break;
		case 'G': case 'g': codes[i] = 2; break;
		case 'T': case 't': codes[i] = 3; break;
		default: codes[i] = 4; break;
		}
	}
}

int kssd_sketch_fasta_file(const char *path, int drfold, int keep_conflict, kssd_u64vec_t *out)
{
	if (kssd_collect_fastx_file(path, drfold, out) != 0)
		return -1;
	kssd_finalize_keys(out, keep_conflict);
	return 0;
}

int kssd_count_keys(kssd_u64vec_t *v, uint32_t **counts)
{
	*counts = NULL;
	if (!v->n)
		return 0;
	radix_sort_u64(v->a, v->n);
	uint32_t *c = malloc(v->n * sizeof *c);
	if (!c)
		return -1;
	size_t w = 0;
	c[0] = 1;
	for (size_t i = 1; i < v->n; ++i) {
		if (v->a[i] == v->a[w]) {
			c[w]++;
		} else {
			v->a[++w] = v->a[i];
			c[w] = 1;
		}
	}
	v->n = w + 1;
	*counts = c;
	return 0;
}

void kssd_drop_conflicts(uint64_t *keys, uint32_t *counts, size_t *n)
{
	/* remove_ctx_with_conflict_obj, kssdlib_sort.c:358-386: a context kept
	 * only when one object carries it; counts stay aligned */
	size_t i = 0, w = 0;
	while (i < *n) {
		size_t c = 1;
		while (i + c < *n && ((keys[i] ^ keys[i + c]) >> KSSD_OBJ_BITS) == 0)
			++c;
		if (c == 1) {
			keys[w] = keys[i];
			if (counts)
				counts[w] = counts[i];
			++w;
		}
		i += c;
	}
	*n = w;
}

int kssd_collect_fastx_file(const char *path, int drfold, kssd_u64vec_t *out)
{
	gzFile fp = gzopen(path, "rb");
	if (!fp)
		return -1;
	size_t cap = 1 << 20, n = 0;
	uint8_t *codes = malloc(cap);
	char *line = malloc(1 << 16);
	if (!codes || !line) {
		gzclose(fp);
		free(codes);
		free(line);
		errno = ENOMEM;
		return -1;
	}
	/* lines longer than the buffer arrive in pieces: `start` marks the
	 * first piece of a line, `skip` the rest of a header or quality line */
	int fastq = -1, line_no = 0;
	bool start = true, skip = false;
	while (gzgets(fp, line, 1 << 16)) {
		size_t l = strlen(line);
		const bool ends = l && line[l - 1] == '\n';
		while (l && (line[l - 1] == '\n' || line[l - 1] == '\r'))
			line[--l] = '\0';
		const bool first_piece = start;
		start = ends;
		if (!first_piece) {
			if (skip)
				continue;
		} else {
			skip = false;
			if (fastq < 0 && l)
				fastq = line[0] == '@';
			if (fastq == 1) {
				/* 4-line FASTQ records: header, sequence, '+', quality */
				const int k = line_no++ % 4;
				if (k != 1) {
					if (k == 0 && n) {
						kssd_sketch_codes(codes, n, drfold, out);
						n = 0;
					}
					skip = true;
					continue;
				}
			} else if (l && (line[0] == '>' || line[0] == ';')) {
				if (n)
					kssd_sketch_codes(codes, n, drfold, out);
				n = 0;
				skip = true;
				continue;
			}
		}
		if (n + l > cap) {
			while (n + l > cap)
				cap <<= 1;
			uint8_t *c2 = realloc(codes, cap);
			if (!c2)
				abort();
			codes = c2;
		}
		kssd_ascii_to_codes(line, l, codes + n);
		n += l;
	}
	if (n)
		kssd_sketch_codes(codes, n, drfold, out);
	gzclose(fp);
	free(codes);
	free(line);
	return 0;
}

uint32_t kssd_sketching_id(int hclen, int holen, int iolen, int drfold)
{
	const uint32_t filter = drfold >= 32 ? 0u : (uint32_t)(UINT32_MAX >> drfold);
	return (uint32_t)mix64((uint64_t)(uint32_t)hclen) ^ (uint32_t)mix64((uint64_t)(uint32_t)holen) ^
	       (uint32_t)mix64((uint64_t)(uint32_t)iolen) ^ (uint32_t)mix64((uint64_t)(uint32_t)drfold) ^
	       (uint32_t)mix64((uint64_t)filter) ^ (uint32_t)mix64(31415926ULL);
}

/* ---------------------------------------------------------------- simulator */
typedef struct {
	uint64_t s[4];
} rng_t;

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void rng_seed(rng_t *r, uint64_t seed)
{
	for (int i = 0; i < 4; ++i) {
		seed += 0x9e3779b97f4a7c15ULL;
		uint64_t z = seed;
		z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
		z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
		r->s[i] = z ^ (z >> 31);
	}
}

/* xoshiro256** */
static inline uint64_t rng_next(rng_t *r)
{
	const uint64_t res = rotl(r->s[1] * 5, 7) * 9;
	const uint64_t t = r->s[1] << 17;
	r->s[2] ^= r->s[0];
	r->s[3] ^= r->s[1];
	r->s[1] ^= r->s[2];
	r->s[0] ^= r->s[3];
	r->s[2] ^= t;
	r->s[3] = rotl(r->s[3], 45);
	return res;
}

static inline double rng_unif(rng_t *r) { return (rng_next(r) >> 11) * 0x1.0p-53; }

static void random_bases(rng_t *r, uint8_t *out, size_t n)
{
	size_t i = 0;
	while (i < n) {
		uint64_t x = rng_next(r);
		for (int k = 0; k < 32 && i < n; ++k, x >>= 2)
			out[i++] = (uint8_t)(x & 3);
	}
}

/* i.i.d. substitutions: geometric skipping keeps this O(mutations). */
static void substitute(rng_t *r, uint8_t *g, size_t n, double rate)
{
	if (rate <= 0)
		return;
	if (rate >= 1)
		rate = 0.999999;
	const double lq = log1p(-rate);
	size_t pos = 0;
	for (;;) {
		const double u = rng_unif(r);
		const double skip = floor(log1p(-u) / lq);
		if (skip >= (double)(n - pos))
			break;
		pos += (size_t)skip;
		g[pos] = (uint8_t)((g[pos] + 1 + (rng_next(r) % 3)) & 3);
		if (++pos >= n)
			break;
	}
}

long kssd_sim_fastq_pair(const uint8_t *codes, size_t len, double coverage, uint64_t seed, int read_len,
			 int insert, int sd, double err, const char *path1, const char *path2)
{
	if (!codes || read_len < 1 || len < (size_t)read_len + 1 || coverage <= 0 || !path1 || !path2)
		return -1;
	const long n = (long)(coverage * (double)len / (2.0 * read_len));
	FILE *f[2] = {fopen(path1, "w"), fopen(path2, "w")};
	const size_t rec = 14 + (size_t)read_len + 3 + (size_t)read_len + 1;
	const size_t chunk = 4096;
	char *buf[2] = {malloc(rec * chunk), malloc(rec * chunk)};
	uint8_t *mate = malloc((size_t)read_len * 2);
	if (!f[0] || !f[1] || !buf[0] || !buf[1] || !mate) {
		for (int m = 0; m < 2; ++m) {
			if (f[m])
				fclose(f[m]);
			free(buf[m]);
		}
		free(mate);
		return -1;
	}
	static const char acgt[4] = {'A', 'C', 'G', 'T'};
	rng_t r;
	rng_seed(&r, seed);
	const double lq = err > 0 && err < 1 ? log1p(-err) : 0.0;
	double next_err = -1; /* bases to the next error (geometric) */
	size_t fill = 0;
	int ok = 1;
	for (long i = 0; i < n && ok; ++i) {
		/* fragment length N(insert, sd) by Box-Muller, clipped */
		const double u1 = rng_unif(&r), u2 = rng_unif(&r);
		long fl = (long)(insert + sd * sqrt(-2.0 * log1p(-u1)) * cos(6.283185307179586 * u2));
		if (fl < read_len)
			fl = read_len;
		if ((size_t)fl > len - 1)
			fl = (long)len - 1;
		const size_t start = (size_t)(rng_unif(&r) * (double)(len - (size_t)fl));
		uint8_t *a = mate, *b = mate + read_len;
		for (int k = 0; k < read_len; ++k) {
			a[k] = codes[start + (size_t)k];
			b[k] = (uint8_t)(3 - codes[start + (size_t)fl - 1 - (size_t)k]);
		}
		if (rng_next(&r) & 1) { /* a fragment of the reverse strand: mates swapped */
			uint8_t *t = a;
			a = b;
			b = t;
		}
		uint8_t *m2[2] = {a, b};
		for (int m = 0; m < 2; ++m) {
			uint8_t *q = m2[m];
			if (lq < 0) { /* i.i.d. per-base errors by geometric skipping */
				for (int k = 0;;) {
					if (next_err < 0)
						next_err = floor(log1p(-rng_unif(&r)) / lq);
					if (next_err >= (double)(read_len - k)) {
						next_err -= (double)(read_len - k);
						break;
					}
					k += (int)next_err;
					q[k] = (uint8_t)((q[k] + 1 + rng_next(&r) % 3) & 3);
					next_err = -1;
					++k;
				}
			}
			char *o = buf[m] + fill * rec;
			snprintf(o, 15, "@r%09ld/%c", i % 1000000000L, m ? '2' : '1');
			o[13] = '\n';
			for (int k = 0; k < read_len; ++k)
				o[14 + k] = acgt[q[k]];
			memcpy(o + 14 + read_len, "\n+\n", 3);
			memset(o + 17 + read_len, 'I', (size_t)read_len);
			o[rec - 1] = '\n';
		}
		if (++fill == chunk || i + 1 == n) {
			for (int m = 0; m < 2; ++m)
				if (fwrite(buf[m], rec, fill, f[m]) != fill)
					ok = 0;
			fill = 0;
		}
	}
	for (int m = 0; m < 2; ++m) {
		if (fclose(f[m]) != 0)
			ok = 0;
		free(buf[m]);
	}
	free(mate);
	return ok ? n : -1;
}

int kssd_sim_genome(const kssd_simspec_t *spec, uint8_t **codes, size_t *len)
{
	rng_t r;
	size_t n = (size_t)spec->anc_len;
	uint8_t *g = malloc(n ? n : 1);
	if (!g)
		return -1;
	rng_seed(&r, spec->anc_seed);
	random_bases(&r, g, n);
	if (spec->anc2_seed) {
		rng_seed(&r, spec->anc2_seed);
		substitute(&r, g, n, spec->anc2_rate);
	}
	rng_seed(&r, spec->mut_seed);
	substitute(&r, g, n, spec->sub_rate);
	if (spec->indel_frac > 0 && n > 100) {
		/* indel events of U{1..50} bp until ~indel_frac*n bases are touched */
		const size_t budget = (size_t)(spec->indel_frac * (double)n);
		const size_t cap = n + budget + 64;
		uint8_t *o = malloc(cap);
		if (!o) {
			free(g);
			return -1;
		}
		/* pick event positions as a sorted uniform sample */
		size_t nev = budget / 25 + 1, touched = 0, ip = 0, op = 0;
		double gap_mean = (double)n / (double)(nev + 1);
		while (ip < n) {
			size_t gap = (size_t)(-log1p(-rng_unif(&r)) * gap_mean);
			if (gap > n - ip)
				gap = n - ip;
			memcpy(o + op, g + ip, gap);
			op += gap;
			ip += gap;
			if (ip >= n || touched >= budget)
				break;
			const size_t el = 1 + rng_next(&r) % 50;
			if (rng_next(&r) & 1) { /* insertion */
				if (op + el > cap)
					break;
				random_bases(&r, o + op, el);
				op += el;
			} else {                 /* deletion */
				ip += el < n - ip ? el : n - ip;
			}
			touched += el;
		}
		if (ip < n && op + (n - ip) <= cap) {
			memcpy(o + op, g + ip, n - ip);
			op += n - ip;
		}
		free(g);
		g = o;
		n = op;
	}
	*codes = g;
	*len = n;
	return 0;
}

int kssd_simsketch(const kssd_simspec_t *specs, int n, int drfold, int keep_conflict,
		   int nthreads, uint64_t **keys_out, uint64_t *index)
{
	kssd_u64vec_t *v = calloc((size_t)(n > 0 ? n : 1), sizeof *v);
	if (!v)
		return -1;
	int fail = 0;
	if (nthreads <= 0)
		nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
	for (int i = 0; i < n; ++i) {
		uint8_t *g = NULL;
		size_t len = 0;
		if (kssd_sim_genome(&specs[i], &g, &len)) {
			fail = 1;
			continue;
		}
		kssd_sketch_codes(g, len, drfold, &v[i]);
		free(g);
		kssd_finalize_keys(&v[i], keep_conflict);
	}
	index[0] = 0;
	for (int i = 0; i < n; ++i)
		index[i + 1] = index[i] + v[i].n;
	uint64_t *keys = malloc((index[n] ? index[n] : 1) * sizeof *keys);
	if (!keys || fail) {
		for (int i = 0; i < n; ++i)
			kssd_u64vec_free(&v[i]);
		free(v);
		free(keys);
		return -1;
	}
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int i = 0; i < n; ++i) {
		if (v[i].n)
			memcpy(keys + index[i], v[i].a, v[i].n * sizeof *keys);
		kssd_u64vec_free(&v[i]);
	}
	free(v);
	*keys_out = keys;
	return 0;
}

/* ------------------------------------------------------------ plain files, in parallel */
/* Base code of an input byte (the Basemap of global_basic.c:131-142): 0..3
 * for ACGT/acgt, 4 for anything else (a window reset), 5 for the line ends
 * a record's lines are joined over. */
static uint8_t g_code[256];
static void init_codes(void)
{
	for (int c = 0; c < 256; ++c)
		g_code[c] = 4;
	g_code['A'] = g_code['a'] = 0;
	g_code['C'] = g_code['c'] = 1;
	g_code['G'] = g_code['g'] = 2;
	g_code['T'] = g_code['t'] = 3;
	g_code['\n'] = 5;
}

/* A run of sequence bytes: a FASTA record's lines (line ends joined; a CR
 * before a line end, or at the end, dropped) or one FASTQ read line. */
typedef struct {
	const char *s;
	size_t from, emit, to; /* sketch [from, to), keep k-mers ending at or past emit;
				* for a FASTQ chunk, emit is its first line's number */
	int file;
	int line; /* SPAN_LINE: one line, no line ends inside (a FASTQ read) */
} span_t;

enum { SPAN_BYTES = 0, SPAN_LINE = 1, SPAN_FASTQ_CHUNK = 2 };

typedef struct {
	span_t *v;
	size_t n, cap;
} spans_t;

static void span_push(spans_t *x, span_t sp)
{
	if (x->n == x->cap) {
		x->cap = x->cap ? 2 * x->cap : 1024;
		x->v = realloc(x->v, x->cap * sizeof *x->v);
		if (!x->v)
			abort();
	}
	x->v[x->n++] = sp;
}

/* a CR that only precedes the line end (or the end) is dropped with it */
static inline bool cr_at_line_end(const char *s, size_t i, size_t to)
{
	size_t j = i + 1;
	while (j < to && s[j] == '\r')
		++j;
	return j == to || s[j] == '\n';
}

/* kssd_sketch_codes over one line of raw bytes (a FASTQ read): every
 * non-ACGT byte resets the window */
static void sketch_line(const char *s, size_t from, size_t to, int drfold, kssd_u64vec_t *out)
{
	const uint32_t filter = drfold >= 32 ? 0u : (uint32_t)(UINT32_MAX >> drfold);
	const unsigned len_mv = 2 * KSSD_KLEN - 2;
	uint64_t fwd = 0, rev = 0;
	int base = 0;
	for (size_t i = from; i < to; ++i) {
		const uint64_t b = g_code[(unsigned char)s[i]];
		if (__builtin_expect(b > 3, 0)) {
			base = 0;
			fwd = rev = 0;
			continue;
		}
		fwd = (fwd << 2) | b;
		rev = (rev >> 2) | ((b ^ 3u) << len_mv);
		if (__builtin_expect(++base < KSSD_KLEN, 0))
			continue;
		const uint64_t tf = fwd & KSSD_CODEN_CTXMASK, tr = rev & KSSD_CODEN_CTXMASK;
		const int use_fwd = tf < tr;
		if (__builtin_expect((uint32_t)mix64(use_fwd ? tf : tr) > filter, 1))
			continue;
		if (out->n == out->cap && vec_reserve(out, out->n + 1))
			abort();
		out->a[out->n++] = coden_key(use_fwd ? fwd : rev);
	}
}

/* kssd_sketch_codes over raw bytes [from, to), line ends skipped */
static void sketch_bytes(const char *s, size_t from, size_t emit, size_t to, int drfold, kssd_u64vec_t *out)
{
	const uint32_t filter = drfold >= 32 ? 0u : (uint32_t)(UINT32_MAX >> drfold);
	const unsigned len_mv = 2 * KSSD_KLEN - 2;
	uint64_t fwd = 0, rev = 0;
	int base = 0;
	for (size_t i = from; i < to; ++i) {
		uint8_t b = g_code[(unsigned char)s[i]];
		if (b == 5)
			continue;
		if (b == 4) {
			if (s[i] == '\r' && cr_at_line_end(s, i, to))
				continue;
			base = 0;
			fwd = rev = 0;
			continue;
		}
		fwd = (fwd << 2) | b;
		rev = (rev >> 2) | ((uint64_t)(b ^ 3u) << len_mv);
		if (++base < KSSD_KLEN || i < emit)
			continue;
		const uint64_t tf = fwd & KSSD_CODEN_CTXMASK, tr = rev & KSSD_CODEN_CTXMASK;
		const int use_fwd = tf < tr;
		if ((uint32_t)mix64(use_fwd ? tf : tr) > filter)
			continue;
		if (out->n == out->cap && vec_reserve(out, out->n + 1))
			abort();
		out->a[out->n++] = coden_key(use_fwd ? fwd : rev);
	}
}

/* start of a piece that begins sketching 31 positions before `at` */
static size_t back_31(const char *s, size_t lo, size_t at)
{
	size_t i = at;
	int pos = 0;
	while (i > lo && pos < KSSD_KLEN - 1) {
		--i;
		const char c = s[i];
		if (c == '\n' || (c == '\r' && cr_at_line_end(s, i, at)))
			continue;
		++pos;
	}
	return i;
}

#define PIECE_BYTES (1u << 20)

/* Spans of one mapped plain file: the same reading as the line parser of
 * kssd_collect_fastx_file (format from the first non-empty line, FASTQ in
 * 4-line records, FASTA records split at '>' or ';' header lines, sequence
 * before the first header kept), FASTA records cut into pieces of about
 * PIECE_BYTES, FASTQ reads grouped about as much per piece. */
static void lens_push(kssd_u64vec_t *v, uint64_t x)
{
	if (v->n == v->cap && vec_reserve(v, v->n + 1))
		abort();
	v->a[v->n++] = x;
}

static void file_spans(const char *s, size_t n, int file, spans_t *out, kssd_u64vec_t *lens)
{
	int fastq = -1;
	long line_no = 0;
	size_t rec = 0; /* FASTA: start of the current record's lines */
	bool in_rec = false;
	/* record lengths as kssd_fasta_record_lengths reads them: records open
	 * at '>' lines, every other line adds its length less one trailing CR */
	bool len_rec = false;
	uint64_t cur = 0;
	size_t i = 0;
	while (i < n) {
		const char *nl = memchr(s + i, '\n', n - i);
		const size_t le = nl ? (size_t)(nl - s) : n; /* line [i, le) */
		size_t l = le;
		while (l > i && s[l - 1] == '\r')
			--l;
		const size_t next = nl ? le + 1 : n;
		if (fastq < 0 && l > i)
			fastq = s[i] == '@';
		if (lens) {
			if (le > i && s[i] == '>') {
				if (len_rec && cur)
					lens_push(lens, cur);
				len_rec = true;
				cur = 0;
			} else if (len_rec) { /* one CR before the line's LF is not counted */
				cur += le - i - (nl && le > i && s[le - 1] == '\r');
			}
		}
		if (fastq == 1) {
			if (line_no++ % 4 == 1 && l > i)
				span_push(out, (span_t){s, i, i, l, file, 1});
		} else if (l > i && (s[i] == '>' || s[i] == ';')) {
			if (in_rec && rec < i)
				span_push(out, (span_t){s, rec, rec, i, file, 0});
			rec = next;
			in_rec = true;
		} else if (!in_rec) {
			rec = i; /* sequence before any header */
			in_rec = true;
		}
		i = next;
	}
	if (fastq != 1 && in_rec && rec < n)
		span_push(out, (span_t){s, rec, rec, n, file, 0});
	if (lens && len_rec && cur)
		lens_push(lens, cur);
}

/* First byte of the first non-empty line (a line of CRs counts as empty),
 * n when there is none. */
static size_t first_nonempty(const char *s, size_t n)
{
	size_t i = 0;
	while (i < n) {
		const char *nl = memchr(s + i, '\n', n - i);
		const size_t le = nl ? (size_t)(nl - s) : n;
		size_t l = le;
		while (l > i && s[l - 1] == '\r')
			--l;
		if (l > i)
			return i;
		i = nl ? le + 1 : n;
	}
	return n;
}

/* A large plain FASTQ file cut into line-aligned chunks of about
 * PIECE_BYTES for the sketching threads: the line ends of each raw cut
 * counted in parallel, each cut moved forward past the end of the line it
 * falls in, and each chunk (span kind SPAN_FASTQ_CHUNK) told the line number
 * of its first line counted from the first non-empty line, so its worker
 * knows which lines are sequence lines (4-line records, as file_spans reads
 * them). */
static void fastq_chunks(const char *s, size_t n, size_t f0, int file, int threads, spans_t *out)
{
	const size_t nc = (n - f0 + PIECE_BYTES - 1) / PIECE_BYTES;
	size_t *cut = malloc((nc + 1) * sizeof *cut), *nl = calloc(nc + 1, sizeof *nl);
	if (!cut || !nl)
		abort();
	for (size_t k = 0; k < nc; ++k)
		cut[k] = f0 + k * PIECE_BYTES;
	cut[nc] = n;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
	for (size_t k = 0; k < nc; ++k) {
		size_t c = 0;
		for (const char *p = s + cut[k], *e = s + cut[k + 1]; (p = memchr(p, '\n', (size_t)(e - p))) != NULL; ++p)
			++c;
		nl[k + 1] = c;
	}
	for (size_t k = 0; k < nc; ++k)
		nl[k + 1] += nl[k]; /* line ends in [f0, cut[k]) */
	size_t prev = f0, prev_line = 0;
	for (size_t k = 1; k <= nc; ++k) {
		size_t at = cut[k], line = nl[k];
		if (k < nc && s[at - 1] != '\n') {
			const char *q = memchr(s + at, '\n', n - at);
			at = q ? (size_t)(q - s) + 1 : n;
			line += q != NULL;
		}
		if (at > prev) {
			span_push(out, (span_t){s, prev, prev_line, at, file, SPAN_FASTQ_CHUNK});
			prev = at;
			prev_line = line;
		}
	}
	free(cut);
	free(nl);
}

/* The sequence lines (line number 1 mod 4) of the whole lines in [from, to),
 * `line` being the number of the first, trailing CRs dropped. */
static void sketch_fastq_chunk(const char *s, size_t from, size_t to, size_t line, int drfold, kssd_u64vec_t *out)
{
	for (size_t i = from; i < to; ++line) {
		const char *q = memchr(s + i, '\n', to - i);
		const size_t le = q ? (size_t)(q - s) : to;
		if (line % 4 == 1) {
			size_t l = le;
			while (l > i && s[l - 1] == '\r')
				--l;
			if (l > i)
				sketch_line(s, i, l, drfold, out);
		}
		i = q ? le + 1 : to;
	}
}

static void split_spans(const spans_t *in, spans_t *out)
{
	for (size_t k = 0; k < in->n; ++k) {
		const span_t sp = in->v[k];
		if (sp.to - sp.from <= PIECE_BYTES || sp.line == SPAN_FASTQ_CHUNK) {
			span_push(out, sp);
			continue;
		}
		for (size_t at = sp.from; at < sp.to; at += PIECE_BYTES) {
			const size_t to = at + PIECE_BYTES < sp.to ? at + PIECE_BYTES : sp.to;
			span_push(out, (span_t){sp.s, back_31(sp.s, sp.from, at), at, to, sp.file, sp.line});
		}
	}
}

static bool gzip_magic(const char *path)
{
	FILE *fp = fopen(path, "rb");
	if (!fp)
		return false;
	unsigned char m[2] = {0, 0};
	const bool gz = fread(m, 1, 2, fp) == 2 && m[0] == 0x1f && m[1] == 0x8b;
	fclose(fp);
	return gz;
}

int kssd_collect_files(char *const *paths, int n, int drfold, int threads, kssd_u64vec_t *out, int *bad,
		       kssd_u64vec_t *fasta_lens, int *lens_done)
{
	static int inited;
	if (!__atomic_load_n(&inited, __ATOMIC_ACQUIRE)) {
#pragma omp critical(kssd_codes)
		if (!inited) {
			init_codes();
			__atomic_store_n(&inited, 1, __ATOMIC_RELEASE);
		}
	}
	*bad = -1;
	char **map = calloc((size_t)(n ? n : 1), sizeof *map);
	size_t *len = calloc((size_t)(n ? n : 1), sizeof *len);
	bool *gz = calloc((size_t)(n ? n : 1), sizeof *gz);
	if (!map || !len || !gz)
		abort();
	spans_t spans = {0}, pieces = {0};
	spans_t *fspans = calloc((size_t)(n ? n : 1), sizeof *fspans);
	if (!fspans)
		abort();
	int rc = 0;
	/* map the files and find their spans, several files at once (FASTQ
	 * files over 8 MB are cut into chunks by all threads afterwards) */
	bool *fq_file = calloc((size_t)(n ? n : 1), sizeof *fq_file);
	size_t *fq0 = calloc((size_t)(n ? n : 1), sizeof *fq0);
	if (!fq_file || !fq0)
		abort();
	int first_bad = n;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
	for (int f = 0; f < n; ++f) {
		gz[f] = gzip_magic(paths[f]);
		if (gz[f])
			continue;
		const int fd = open(paths[f], O_RDONLY);
		struct stat st;
		void *m = MAP_FAILED;
		if (fd >= 0 && fstat(fd, &st) == 0) {
			len[f] = (size_t)st.st_size;
			m = len[f] ? mmap(NULL, len[f], PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0) : NULL;
		}
		if (fd >= 0)
			close(fd);
		if (fd < 0 || m == MAP_FAILED) {
#pragma omp critical(kssd_collect_fail)
			if (f < first_bad)
				first_bad = f;
			continue;
		}
		if (!len[f])
			continue;
		madvise(m, len[f], MADV_SEQUENTIAL | MADV_WILLNEED);
		map[f] = m;
		fq0[f] = first_nonempty(map[f], len[f]);
		if (fq0[f] < len[f] && map[f][fq0[f]] == '@') {
			if (len[f] - fq0[f] <= (8u << 20))
				fastq_chunks(map[f], len[f], fq0[f], f, 1, &fspans[f]);
			else
				fq_file[f] = true;
		} else {
			file_spans(map[f], len[f], f, &fspans[f], fasta_lens ? &fasta_lens[f] : NULL);
			if (lens_done)
				lens_done[f] = fasta_lens != NULL;
		}
	}
	if (first_bad < n) {
		*bad = first_bad;
		rc = -1;
	}
	for (int f = 0; f < n && !rc; ++f) {
		if (fq_file[f])
			fastq_chunks(map[f], len[f], fq0[f], f, threads, &fspans[f]);
		for (size_t k = 0; k < fspans[f].n; ++k)
			span_push(&spans, fspans[f].v[k]);
	}
	for (int f = 0; f < n; ++f)
		free(fspans[f].v);
	free(fspans);
	free(fq_file);
	free(fq0);
	if (!rc) {
		split_spans(&spans, &pieces);
		/* FASTQ reads grouped into pieces of about PIECE_BYTES */
		size_t ng = 0;
		size_t *gstart = malloc((pieces.n + 1) * sizeof *gstart);
		if (!gstart)
			abort();
		for (size_t k = 0; k < pieces.n;) {
			size_t e = k, bytes = 0;
			while (e < pieces.n && pieces.v[e].file == pieces.v[k].file &&
			       (e == k || bytes + (pieces.v[e].to - pieces.v[e].from) <= PIECE_BYTES)) {
				bytes += pieces.v[e].to - pieces.v[e].from;
				++e;
			}
			gstart[ng++] = k;
			k = e;
		}
		gstart[ng] = pieces.n;
		kssd_u64vec_t *part = calloc(ng ? ng : 1, sizeof *part);
		int fail = -1;
		if (!part)
			abort();
#pragma omp parallel num_threads(threads)
		{
#pragma omp for schedule(dynamic, 1) nowait
			for (size_t g = 0; g < ng; ++g)
				for (size_t k = gstart[g]; k < gstart[g + 1]; ++k) {
					const span_t *sp = &pieces.v[k];
					if (sp->line == SPAN_FASTQ_CHUNK)
						sketch_fastq_chunk(sp->s, sp->from, sp->to, sp->emit, drfold, &part[g]);
					else if (sp->line && sp->emit == sp->from)
						sketch_line(sp->s, sp->from, sp->to, drfold, &part[g]);
					else
						sketch_bytes(sp->s, sp->from, sp->emit, sp->to, drfold, &part[g]);
				}
			/* gzip files: one per thread through the line reader */
#pragma omp for schedule(dynamic, 1)
			for (int f = 0; f < n; ++f)
				if (gz[f] && kssd_collect_fastx_file(paths[f], drfold, &out[f]) != 0) {
#pragma omp critical(kssd_collect_fail)
					fail = f;
				}
		}
		if (fail >= 0) {
			*bad = fail;
			rc = -1;
		}
		for (size_t g = 0; g < ng && !rc; ++g) {
			kssd_u64vec_t *o = &out[pieces.v[gstart[g]].file];
			if (part[g].n) {
				if (vec_reserve(o, o->n + part[g].n))
					abort();
				memcpy(o->a + o->n, part[g].a, part[g].n * sizeof *o->a);
				o->n += part[g].n;
			}
		}
		for (size_t g = 0; g < ng; ++g)
			kssd_u64vec_free(&part[g]);
		free(part);
		free(gstart);
	}
	for (int f = 0; f < n; ++f)
		if (map[f])
			munmap(map[f], len[f]);
	free(spans.v);
	free(pieces.v);
	free(map);
	free(len);
	free(gz);
	if (rc)
		errno = errno ? errno : EIO;
	return rc;
}
