/*
 * Inverted-index builder and per-sample context totals.
 * See kssd_host.h.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "kssd_host.h"

#define GID_NBITS 20
#define BUCKET_BITS 12

static int rec_cmp(const void *pa, const void *pb)
{
	/* kssdlib_sort.c:145-148: (ctxgid, obj) ascending */
	const kssd_ctxgidobj_t *a = pa, *b = pb;
	if (a->ctxgid != b->ctxgid)
		return a->ctxgid < b->ctxgid ? -1 : 1;
	return (a->obj > b->obj) - (a->obj < b->obj);
}

int kssd_build_index(const uint64_t *keys, const uint64_t *index, int n, unsigned obj_bits,
		     int nthreads, kssd_ctxgidobj_t *out)
{
	if (n < 0 || n > (1 << GID_NBITS) || index[n] >= UINT32_MAX || obj_bits < 1 || obj_bits > 32)
		return -1;
	if (nthreads <= 0)
		nthreads = 1;
	const size_t total = (size_t)index[n];
	const uint64_t objmask = obj_bits == 32 ? 0xffffffffULL : ((1ULL << obj_bits) - 1ULL);
	const int nb = 1 << BUCKET_BITS;
	/* per-thread bucket histograms over a static split of the samples */
	size_t *hist = calloc((size_t)nthreads * nb, sizeof *hist);
	if (!hist)
		return -1;
#pragma omp parallel num_threads(nthreads)
	{
#ifdef _OPENMP
		const int t = omp_get_thread_num();
#else
		const int t = 0;
#endif
		size_t *h = hist + (size_t)t * nb;
#pragma omp for schedule(static)
		for (int g = 0; g < n; ++g)
			for (uint64_t i = index[g]; i < index[g + 1]; ++i) {
				const uint64_t cg = (keys[i] >> obj_bits) << GID_NBITS;
				h[cg >> (64 - BUCKET_BITS)]++;
			}
	}
	/* exclusive offsets, bucket-major then thread */
	size_t *start = malloc(((size_t)nb + 1) * sizeof *start);
	if (!start) {
		free(hist);
		return -1;
	}
	size_t sum = 0;
	for (int b = 0; b < nb; ++b) {
		start[b] = sum;
		for (int t = 0; t < nthreads; ++t) {
			const size_t c = hist[(size_t)t * nb + b];
			hist[(size_t)t * nb + b] = sum;
			sum += c;
		}
	}
	start[nb] = sum;
#pragma omp parallel num_threads(nthreads)
	{
#ifdef _OPENMP
		const int t = omp_get_thread_num();
#else
		const int t = 0;
#endif
		size_t *h = hist + (size_t)t * nb;
#pragma omp for schedule(static)
		for (int g = 0; g < n; ++g)
			for (uint64_t i = index[g]; i < index[g + 1]; ++i) {
				kssd_ctxgidobj_t r;
				/* uint64_ctxobj2ctxgidobj96, sketch_rearrange.h:95-100 */
				r.ctxgid = ((keys[i] >> obj_bits) << GID_NBITS) | ((uint64_t)g & ((1ULL << GID_NBITS) - 1));
				r.obj = (uint32_t)(keys[i] & objmask);
				out[h[r.ctxgid >> (64 - BUCKET_BITS)]++] = r;
			}
#pragma omp for schedule(dynamic, 16)
		for (int b = 0; b < nb; ++b)
			if (start[b + 1] - start[b] > 1)
				qsort(out + start[b], start[b + 1] - start[b], sizeof *out, rec_cmp);
	}
	(void)total;
	free(start);
	free(hist);
	return 0;
}

uint32_t kssd_ctx_runs(const uint64_t *a, size_t n, unsigned obj_bits, int ignoreconflict)
{
	/* count_ctx_runs_sorted_ctxobj64_local, command_ani.c:337-349 */
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

void kssd_sample_ctx_counts(const uint64_t *keys, const uint64_t *index, int n, unsigned obj_bits,
			    int conflict, int ignoreconflict, uint32_t *counts)
{
	/* ani_ctx_counts_for_sketch, command_ani.c:1533-1548 */
	for (int i = 0; i < n; ++i) {
		const size_t len = (size_t)(index[i + 1] - index[i]);
		counts[i] = conflict ? kssd_ctx_runs(keys + index[i], len, obj_bits, ignoreconflict)
				     : (uint32_t)len;
	}
}

void kssd_index_ctx_counts(const kssd_ctxgidobj_t *rec, size_t n, const uint64_t *index, int ref_gnum,
			   int ignoreconflict, uint32_t *counts)
{
	/* ref_ctx_counts_from_ctxgidobj, command_ani.c:351-370: entries per gid,
	 * minus the extra objects of each multi-object run (the whole run under
	 * ignoreconflict) */
	for (int g = 0; g < ref_gnum; ++g)
		counts[g] = (uint32_t)(index[g + 1] - index[g]);
	for (size_t i = 0; i < n;) {
		const uint64_t cg = rec[i].ctxgid;
		size_t len = 0;
		do {
			++i;
			++len;
		} while (i < n && rec[i].ctxgid == cg);
		if (len > 1)
			counts[cg & ((1u << 20) - 1)] -= (uint32_t)(ignoreconflict ? len : len - 1);
	}
}
