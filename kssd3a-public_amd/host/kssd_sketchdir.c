/*
 * Sketch-directory I/O: lcofiles.stat, comblco, comblco.index,
 * sortedcomb_ctxgid64obj32 (+ .unique_bits), lcofiles.anno,
 * lcofiles.infilemeta (global_basic.h:138-193, command_sketch.c:29-39,
 * :850-938, :989-1005).
 */
#define _GNU_SOURCE
#include "kssd_sketchdir.h"

#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <fcntl.h>
#include <omp.h>
#include <stdbool.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

static char *path_join(const char *dir, const char *name)
{
	char *p = NULL;
	if (asprintf(&p, "%s/%s", dir, name) < 0)
		return NULL;
	return p;
}

bool kssd_file_in_dir(const char *dir, const char *name)
{
	char *p = path_join(dir, name);
	struct stat st;
	const bool ok = p && stat(p, &st) == 0 && S_ISREG(st.st_mode);
	free(p);
	return ok;
}

void *kssd_read_file(const char *path, size_t *size)
{
	const int fd = open(path, O_RDONLY);
	struct stat st;
	if (fd < 0)
		return NULL;
	if (fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
		close(fd);
		return NULL;
	}
	const size_t n = (size_t)st.st_size;
	char *buf = malloc(n ? n : 1);
	bool ok = buf != NULL;
	/* large files in 8 MB pieces by several threads: copying from the page
	 * cache is what reading the index and the keys costs (one core left to
	 * whatever the caller runs beside, such as starting the GPU runtime) */
	const size_t piece = 8u << 20;
	const size_t np = (n + piece - 1) / piece;
	int nt = omp_get_num_procs() - 1;
	nt = nt < 1 ? 1 : nt > 16 ? 16 : nt;
	nt = np < (size_t)nt ? (int)np : nt;
#pragma omp parallel for schedule(dynamic, 1) if (np > 2) num_threads(nt)
	for (size_t k = 0; k < np; ++k) {
		if (!ok)
			continue;
		size_t at = k * piece;
		const size_t end = at + piece < n ? at + piece : n;
		while (at < end) {
			const ssize_t r = pread(fd, buf + at, end - at, (off_t)at);
			if (r <= 0) {
				ok = false;
				break;
			}
			at += (size_t)r;
		}
	}
	close(fd);
	if (!ok) {
		free(buf);
		return NULL;
	}
	*size = n;
	return buf;
}

static void *read_in_dir(const char *dir, const char *name, size_t *size)
{
	char *p = path_join(dir, name);
	void *b = p ? kssd_read_file(p, size) : NULL;
	free(p);
	return b;
}

int kssd_sketchdir_open(const char *dir, unsigned parts, kssd_sketchdir_t *s, char *err, size_t errlen)
{
	memset(s, 0, sizeof *s);
	size_t n = 0;
	void *stat = read_in_dir(dir, KSSD_STAT_FILE, &n);
	if (!stat || n < sizeof(kssd_stat_t)) {
		snprintf(err, errlen, "%s/%s: cannot read sketch statistics", dir, KSSD_STAT_FILE);
		free(stat);
		return -1;
	}
	memcpy(&s->stat, stat, sizeof s->stat);
	const int num = s->stat.infile_num;
	if (num < 0 || n < sizeof(kssd_stat_t) + (size_t)num * KSSD_PATHLEN) {
		snprintf(err, errlen, "%s/%s: truncated (%zu bytes for %d samples)", dir, KSSD_STAT_FILE, n, num);
		free(stat);
		return -1;
	}
	s->n = num;
	s->names = malloc((size_t)(num ? num : 1) * KSSD_PATHLEN);
	if (!s->names) {
		free(stat);
		return -1;
	}
	memcpy(s->names, (char *)stat + sizeof(kssd_stat_t), (size_t)num * KSSD_PATHLEN);
	for (int i = 0; i < num; ++i)
		s->names[i][KSSD_PATHLEN - 1] = '\0';
	free(stat);
	/* Bitslen, const_comask_init (sketch_rearrange.c:7-39) */
	s->ctx_bits = s->stat.coden_len > 0 ? 4u * (unsigned)s->stat.coden_len : 4u * (unsigned)s->stat.hclen;
	s->obj_bits = 2u * (unsigned)s->stat.klen - s->ctx_bits;
	s->index = read_in_dir(dir, KSSD_INDEX_FILE, &n);
	if (!s->index || n != ((size_t)num + 1) * 8) {
		snprintf(err, errlen, "%s/%s: expected %d offsets", dir, KSSD_INDEX_FILE, num + 1);
		kssd_sketchdir_close(s);
		return -1;
	}
	if (parts & KSSD_SK_KEYS) {
		s->keys = read_in_dir(dir, KSSD_COMB_FILE, &n);
		if (!s->keys || n != s->index[num] * 8) {
			snprintf(err, errlen, "%s/%s: size does not match %s", dir, KSSD_COMB_FILE, KSSD_INDEX_FILE);
			kssd_sketchdir_close(s);
			return -1;
		}
	}
	if ((parts & KSSD_SK_RECORDS) && kssd_file_in_dir(dir, KSSD_SORTED_INDEX_FILE)) {
		s->records = read_in_dir(dir, KSSD_SORTED_INDEX_FILE, &n);
		if (!s->records || n != s->index[num] * sizeof(kssd_ctxgidobj_t)) {
			snprintf(err, errlen, "%s/%s: size does not match the sketch", dir, KSSD_SORTED_INDEX_FILE);
			kssd_sketchdir_close(s);
			return -1;
		}
	}
	if ((parts & KSSD_SK_ANNO) && kssd_file_in_dir(dir, KSSD_ANNO_FILE)) {
		s->anno = read_in_dir(dir, KSSD_ANNO_FILE, &n);
		if (!s->anno || n != (size_t)num * KSSD_PATHLEN) {
			snprintf(err, errlen, "%s/%s has %zu bytes, expected %zu", dir, KSSD_ANNO_FILE, n,
				 (size_t)num * KSSD_PATHLEN);
			kssd_sketchdir_close(s);
			return -1;
		}
	}
	if ((parts & KSSD_SK_META) && kssd_file_in_dir(dir, KSSD_META_FILE)) {
		s->meta = read_in_dir(dir, KSSD_META_FILE, &n);
		if (!s->meta || n != (size_t)num * sizeof(kssd_infile_meta_t)) {
			snprintf(err, errlen, "%s/%s has %zu bytes, expected %zu", dir, KSSD_META_FILE, n,
				 (size_t)num * sizeof(kssd_infile_meta_t));
			kssd_sketchdir_close(s);
			return -1;
		}
	}
	return 0;
}

void kssd_sketchdir_close(kssd_sketchdir_t *s)
{
	free(s->names);
	free(s->index);
	free(s->keys);
	free(s->records);
	free(s->anno);
	free(s->meta);
	memset(s, 0, sizeof *s);
}

static int write_file(const char *dir, const char *name, const void *data, size_t n)
{
	char *p = path_join(dir, name);
	if (!p)
		return -1;
	FILE *fp = fopen(p, "wb");
	free(p);
	if (!fp)
		return -1;
	const int ok = (n == 0 || fwrite(data, 1, n, fp) == n) && fclose(fp) == 0;
	return ok ? 0 : -1;
}

int kssd_sketchdir_write(const char *dir, const kssd_stat_t *stat, char (*names)[KSSD_PATHLEN],
			 const uint64_t *keys, const uint64_t *index)
{
	if (mkdir(dir, 0777) != 0 && errno != EEXIST)
		return -1;
	const size_t nh = sizeof *stat + (size_t)stat->infile_num * KSSD_PATHLEN;
	char *head = calloc(1, nh);
	if (!head)
		return -1;
	memcpy(head, stat, sizeof *stat);
	memcpy(head + sizeof *stat, names, (size_t)stat->infile_num * KSSD_PATHLEN);
	int rc = write_file(dir, KSSD_STAT_FILE, head, nh);
	free(head);
	if (rc)
		return rc;
	rc = write_file(dir, KSSD_COMB_FILE, keys, index[stat->infile_num] * 8);
	if (rc)
		return rc;
	return write_file(dir, KSSD_INDEX_FILE, index, ((size_t)stat->infile_num + 1) * 8);
}

static int cmp_u32(const void *a, const void *b)
{
	const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
	return (x > y) - (x < y);
}

/* occurrences of v in the sorted array s[0..n) (at most 2 are needed) */
static int count_sorted(const uint32_t *s, size_t n, uint32_t v)
{
	size_t lo = 0, hi = n;
	while (lo < hi) {
		const size_t mid = lo + (hi - lo) / 2;
		if (s[mid] < v)
			lo = mid + 1;
		else
			hi = mid;
	}
	int c = 0;
	while (lo < n && s[lo] == v && c < 2) {
		++lo;
		++c;
	}
	return c;
}

int kssd_write_index_files(const char *dir, const kssd_ctxgidobj_t *rec, size_t n, const uint8_t *unique_bits)
{
	if (!unique_bits)
		return kssd_write_sorted_index(dir, rec, n);
	int rc = write_file(dir, KSSD_SORTED_INDEX_FILE, rec, n * sizeof *rec);
	if (rc)
		return rc;
	return write_file(dir, KSSD_SORTED_INDEX_FILE ".unique_bits", unique_bits, (n + 7) / 8);
}

int kssd_write_sorted_index(const char *dir, const kssd_ctxgidobj_t *rec, size_t n)
{
	int rc = write_file(dir, KSSD_SORTED_INDEX_FILE, rec, n * sizeof *rec);
	if (rc)
		return rc;
	/* write_unique_marker_bits_from_sorted, command_sketch.c:855-914: one
	 * bit per record (LSB first), set iff its (ctx, obj) occurs once in its
	 * context group.  A group is ordered by gid, not by object, so large
	 * groups are checked against a sorted copy of their objects. */
	uint8_t *bits = calloc((n + 7) / 8 ? (n + 7) / 8 : 1, 1);
	uint32_t *tmp = NULL;
	size_t tmp_cap = 0;
	if (!bits)
		return -1;
	for (size_t i = 0; i < n;) {
		const uint64_t ctx = rec[i].ctxgid >> 20;
		size_t e = i;
		while (e < n && (rec[e].ctxgid >> 20) == ctx)
			++e;
		const size_t g = e - i;
		if (g <= 16) {
			for (size_t a = i; a < e; ++a) {
				size_t same = 0;
				for (size_t b = i; b < e && same < 2; ++b)
					same += rec[b].obj == rec[a].obj;
				if (same == 1)
					bits[a >> 3] |= (uint8_t)(1u << (a & 7));
			}
		} else {
			if (g > tmp_cap) {
				uint32_t *t = realloc(tmp, g * sizeof *t);
				if (!t) {
					free(tmp);
					free(bits);
					return -1;
				}
				tmp = t;
				tmp_cap = g;
			}
			for (size_t a = 0; a < g; ++a)
				tmp[a] = rec[i + a].obj;
			qsort(tmp, g, sizeof *tmp, cmp_u32);
			for (size_t a = i; a < e; ++a)
				if (count_sorted(tmp, g, rec[a].obj) == 1)
					bits[a >> 3] |= (uint8_t)(1u << (a & 7));
		}
		i = e;
	}
	free(tmp);
	rc = write_file(dir, KSSD_SORTED_INDEX_FILE ".unique_bits", bits, (n + 7) / 8);
	free(bits);
	return rc;
}

/* ------------------------------------------------------------ input metadata */
static bool has_suffix(const char *s, const char *suf)
{
	const size_t a = strlen(s), b = strlen(suf);
	return a >= b && strcmp(s + a - b, suf) == 0;
}

static bool is_compressed(const char *path)
{
	return has_suffix(path, ".gz") || has_suffix(path, ".bz2");
}

int kssd_infile_fmt_from_path(const char *path)
{
	/* the name without a .gz/.bz2 suffix, then the format's extensions */
	char buf[4096];
	if (!path || strlen(path) >= sizeof buf)
		return 0;
	strcpy(buf, path);
	if (has_suffix(buf, ".gz"))
		buf[strlen(buf) - 3] = '\0';
	else if (has_suffix(buf, ".bz2"))
		buf[strlen(buf) - 4] = '\0';
	static const char *const fq[] = {".fq", ".fastq"};
	static const char *const fa[] = {".fasta", ".fna", ".fas", ".fa"};
	for (size_t i = 0; i < sizeof fq / sizeof fq[0]; ++i)
		if (has_suffix(buf, fq[i]))
			return 2;
	for (size_t i = 0; i < sizeof fa / sizeof fa[0]; ++i)
		if (has_suffix(buf, fa[i]))
			return 1;
	return 0;
}

static int cmp_u64(const void *a, const void *b)
{
	const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return (x > y) - (x < y);
}

/* Sequence length of every FASTA record (all characters of its sequence
 * lines; a trailing CR is not counted), empty records skipped
 * (append_asm_lengths_for_path, command_sketch.c:1205-1227). */
int kssd_fasta_record_lengths(const char *path, uint64_t **lens, size_t *n, size_t *cap)
{
	gzFile fp = gzopen(path, "rb");
	if (!fp)
		return -1;
	char *line = malloc(1 << 16);
	if (!line) {
		gzclose(fp);
		return -1;
	}
	uint64_t cur = 0;
	bool in_rec = false, at_line_start = true, in_header = false;
	for (;;) {
		const bool got = gzgets(fp, line, 1 << 16) != NULL;
		size_t l = got ? strlen(line) : 0;
		const bool ends_line = l && line[l - 1] == '\n';
		const bool header = got && !in_header && at_line_start && l && line[0] == '>';
		if ((!got || header) && in_rec && cur) {
			if (*n == *cap) {
				const size_t c2 = *cap ? *cap * 2 : 64;
				uint64_t *w = realloc(*lens, c2 * sizeof *w);
				if (!w)
					abort();
				*lens = w;
				*cap = c2;
			}
			(*lens)[(*n)++] = cur;
		}
		if (!got)
			break;
		if (ends_line)
			--l;
		if (ends_line && l && line[l - 1] == '\r')
			--l;
		if (in_header) { /* the rest of a header longer than the buffer */
			in_header = !ends_line;
		} else if (header) {
			in_header = !ends_line;
			in_rec = true;
			cur = 0;
		} else if (in_rec) {
			cur += l;
		}
		at_line_start = ends_line;
	}
	gzclose(fp);
	free(line);
	return 0;
}

void kssd_infile_meta_from_lengths(uint64_t *len, size_t n, int fmt, int create_type, unsigned flags,
				   kssd_infile_meta_t *out)
{
	memset(out, 0, sizeof *out);
	out->meta_fmt_version = 1;
	out->infile_fmt = (int8_t)fmt;
	out->create_type = (int8_t)create_type;
	out->infile_flags = (uint8_t)flags;
	out->asm_level = fmt == 2 ? 0.0f : -1.0f;
	if (!n)
		return;
	/* running mean and squared deviations in input order, then the median
	 * of the sorted lengths */
	uint64_t total = 0;
	double mean = 0.0, m2 = 0.0;
	for (size_t i = 0; i < n; ++i) {
		total += len[i];
		const double x = (double)len[i];
		const double d = x - mean;
		mean += d / (double)(i + 1);
		m2 += d * (x - mean);
	}
	qsort(len, n, sizeof *len, cmp_u64);
	out->total_length_bp = total;
	out->record_count = n > UINT32_MAX ? UINT32_MAX : (uint32_t)n;
	const uint64_t med = (n & 1) ? len[n / 2] : (len[n / 2 - 1] + len[n / 2]) / 2;
	out->median_length_bp = med > UINT32_MAX ? UINT32_MAX : (uint32_t)med;
	if (mean > 0.0 && n > 1)
		out->length_cv = (float)(sqrt(m2 / (double)n) / mean);
	if (fmt == 1 && total > 0)
		out->asm_level = (float)((double)len[n - 1] / (double)total);
}

unsigned kssd_infile_flags_from_path(const char *path)
{
	return (strcmp(path, "-") == 0 ? 0x02u : 0u) | (is_compressed(path) ? 0x04u : 0u);
}

int kssd_infile_meta_for_path(const char *path, kssd_infile_meta_t *out)
{
	const int fmt = kssd_infile_fmt_from_path(path);
	uint64_t *len = NULL;
	size_t n = 0, cap = 0;
	if (fmt == 1 && kssd_fasta_record_lengths(path, &len, &n, &cap) != 0)
		return -1;
	kssd_infile_meta_from_lengths(len, n, fmt, 0, kssd_infile_flags_from_path(path), out);
	free(len);
	return 0;
}

int kssd_write_dir_file(const char *dir, const char *name, const void *data, size_t n)
{
	return write_file(dir, name, data, n);
}

int kssd_write_infile_meta(const char *dir, const kssd_infile_meta_t *meta, int n)
{
	return write_file(dir, KSSD_META_FILE, meta, (size_t)n * sizeof *meta);
}
