/*
 * Host-side (CPU) components around the HIP path: the sketcher that turns
 * sequences into context-object keys, a seeded genome simulator used to
 * make synthetic workloads, and the inverted-index builder.
 *
 * Key layout (default coden pattern, NUM_CODENS = 11, klen = 32):
 *   key = ctx44 << 20 | obj20, ctx/obj extracted from the canonical 32-mer
 *   with the mask 0xf3cf3cf3cf3cf3cf (sketch_rearrange.h:109-122,
 *   command_sketch.h:66-72).
 */
#ifndef KSSD_HOST_H
#define KSSD_HOST_H
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include "kssdhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KSSD_CODEN_CTXMASK 0xf3cf3cf3cf3cf3cfULL
#define KSSD_KLEN 32
#define KSSD_OBJ_BITS 20

/* Growable u64 vector owned by the host library. */
typedef struct {
	uint64_t *a;
	size_t n, cap;
} kssd_u64vec_t;
void kssd_u64vec_free(kssd_u64vec_t *v);

/*
 * Append the kept keys of one contiguous sequence of 2-bit base codes
 * (0..3 = A,C,G,T; any value >= 4 is a non-ACGT base that restarts the
 * window) to out.  Restates sketch_read_into_vec, command_sketch.c:3913-3949:
 * canonical strand by the smaller ctx-masked k-mer (ties take the reverse
 * strand), kept iff (u32)mix64(ctx-masked canonical k-mer) <= UINT32_MAX >> drfold.
 */
void kssd_sketch_codes(const uint8_t *codes, size_t len, int drfold, kssd_u64vec_t *out);

/* Sort ascending, drop duplicates and (unless keep_conflict) drop every
 * context that carries more than one object (kssdlib_sort.c:388-407). */
void kssd_finalize_keys(kssd_u64vec_t *v, int keep_conflict);

/* Map ASCII to base codes with the reference's Basemap (global_basic.c:131-142). */
void kssd_ascii_to_codes(const char *s, size_t len, uint8_t *codes);

/* Sketch a FASTA/plain-sequence file as ONE sample (all records pooled,
 * the window restarting at each record).  Returns 0 or -1 (errno set). */
int kssd_sketch_fasta_file(const char *path, int drfold, int keep_conflict, kssd_u64vec_t *out);

/* The keys of every k-mer window of one FASTA/FASTQ[.gz] file appended to
 * out as they occur (unsorted, repeats kept): the input of the counted
 * read-set modes.  Returns 0 or -1 (errno set). */
int kssd_collect_fastx_file(const char *path, int drfold, kssd_u64vec_t *out);
/* kssd_collect_fastx_file for n files at once, out[i] for file i: plain
 * files are mapped and cut into pieces of about 1 MB (FASTA records split
 * with a 31-base overlap, FASTQ reads grouped) sketched by `threads`
 * threads; gzip files go one per thread through the line reader.  The keys
 * are those of the line reader.  0, or -1 with *bad the failing file. */
int kssd_collect_files(char *const *paths, int n, int drfold, int threads, kssd_u64vec_t *out, int *bad,
		       kssd_u64vec_t *fasta_lens, int *lens_done);
/* (fasta_lens, when given: for the plain files read in one pass, the record
 * lengths kssd_fasta_record_lengths would find, lens_done[i] = 1; empty for
 * FASTQ.  Files left at lens_done[i] = 0 were not measured.) */
/* Sort v ascending and merge repeats; *counts (malloc'd, NULL when v is
 * empty) holds each key's number of occurrences (build_kv_from_vec,
 * command_sketch.c:4156-4187).  0 or -1 (allocation). */
int kssd_count_keys(kssd_u64vec_t *v, uint32_t **counts);
/* Drop every context carried by more than one object from sorted keys,
 * keeping counts (may be NULL) aligned (kssdlib_sort.c:358-386). */
void kssd_drop_conflicts(uint64_t *keys, uint32_t *counts, size_t *n);

/* ---- k-mer count filters and read QC (command_sketch.c:56-768) ----------- */
#define KSSD_QC_RANGE_VALID 1u
#define KSSD_QC_RANGE_APPLIED 2u
/* lcofiles.qc record (dim_sketch_qc_stat_t, global_basic.h:156-162) */
typedef struct {
	uint32_t flags, lower, upper, mode;
} kssd_qc_stat_t;
typedef struct {
	int min_count;     /* -n/--LstKmerOcrs (1 = off) */
	double percentile; /* --npercentile (0 = off) */
	int cap;           /* --ncap (0 = off) */
	int reads_qc;      /* --readsQC */
	int record_qc;     /* infer the QC range without applying it (-A) */
} kssd_count_opts_t;
typedef struct {
	uint32_t raw_cutoff, cutoff, upper, mode;
	int used_reads_qc;
	kssd_qc_stat_t qc;
} kssd_count_filter_t;
/* Whether the options need per-key counts (sketch_needs_kmer_counts,
 * command_sketch.c:51-54, with -A given as record_qc). */
int kssd_count_opts_need_counts(const kssd_count_opts_t *o);
/* apply_kmer_count_filters (command_sketch.c:717-767) on sorted keys with
 * their counts: lower cutoff from -n and the weighted count percentile,
 * the inferred read-QC range (recorded with -A, applied with --readsQC),
 * the cap, then keys outside [cutoff, upper] dropped in place. */
kssd_count_filter_t kssd_apply_count_filters(uint64_t *keys, uint32_t *counts, size_t *n,
					     const kssd_count_opts_t *o);
/* The stderr line print_kmer_count_filter_result (command_sketch.c:690-715)
 * writes for --npercentile / --readsQC runs. */
void kssd_print_count_filter(FILE *fp, const kssd_count_opts_t *o, const char *name,
			     const kssd_count_filter_t *r, int leading_newline);

/* The keys of n files pooled as one sample, as `sketch --asone` makes them:
 * every window of every file collected (kssd_collect_files), then, when the
 * count options need counts, counted and filtered by count (the QC range
 * recorded, or applied with reads_qc), else sorted and merged; contexts with
 * several objects dropped unless keep_conflict (--conflict).  *counts
 * (malloc'd, may be NULL when counts are not needed or keys is empty) is
 * aligned with keys; *filter the count filter's result.  0, or -1 with *bad
 * the failing file (or -1 for an allocation). */
int kssd_sketch_pooled(char *const *paths, int n, int drfold, int threads, int keep_conflict,
		       const kssd_count_opts_t *co, kssd_u64vec_t *keys, uint32_t **counts, kssd_count_filter_t *filter,
		       int *bad);

/* GET_SKETCHING_ID(hclen, holen, iolen, drfold, FILTER), command_sketch.h:49-52
 * and command_sketch.c:771.  The coden pattern passes (NUM_CODENS, 0, 0, drfold)
 * (command_sketch_wrapper.c:924). */
uint32_t kssd_sketching_id(int hclen, int holen, int iolen, int drfold);

/* ---- synthetic genomes -------------------------------------------------- */
typedef struct {
	uint64_t anc_seed;  /* seed of the ancestor genome (uniform ACGT) */
	uint64_t anc_len;   /* ancestor length in bases */
	uint64_t mut_seed;  /* seed of the mutation process */
	double sub_rate;    /* i.i.d. substitution rate (new base uniform among the other 3) */
	double indel_frac;  /* bases touched by indels / length (events of 1..50 bp, 50/50 ins/del) */
	uint64_t anc2_seed; /* nonzero: first derive the ancestor from anc_seed with ... */
	double anc2_rate;   /* ... this substitution rate (clade structure) */
} kssd_simspec_t;

/* Generate genome `spec` into codes (malloc'd, *len set). */
int kssd_sim_genome(const kssd_simspec_t *spec, uint8_t **codes, size_t *len);
/* Paired-end reads of a simulated genome (2-bit codes) as two plain FASTQ
 * files: coverage * len / (2 read_len) fragments of N(insert, sd) bp
 * (clipped to [read_len, len - 1]) at uniform starts, read 1 the fragment's
 * first read_len bases and read 2 the reverse complement of its last, the
 * mates swapped for half the fragments (the reverse strand); i.i.d. per-base
 * substitution errors at rate err; a constant quality line.  Deterministic
 * in seed.  Returns the number of pairs, or -1. */
long kssd_sim_fastq_pair(const uint8_t *codes, size_t len, double coverage, uint64_t seed, int read_len,
			 int insert, int sd, double err, const char *path1, const char *path2);

/*
 * Simulate and sketch n genomes in parallel (nthreads OpenMP threads).
 * keys: concatenated finalized sketches (malloc'd, caller frees with free());
 * index: n+1 prefix offsets (caller-allocated).
 */
int kssd_simsketch(const kssd_simspec_t *specs, int n, int drfold, int keep_conflict,
		   int nthreads, uint64_t **keys, uint64_t *index);

/* ---- inverted index ------------------------------------------------------ */
/*
 * ctxobj64_2ctxgidobj (sketch_rearrange.c:65-78) + sort by (ctxgid, obj)
 * (kssdlib_sort.c:145-148) with a parallel radix sort.  out has index[n]
 * records.  Returns 0, or -1 if the input breaks the format's limits
 * (n > 2^20 samples, more than UINT32_MAX entries; command_sketch.c:922-928).
 */
int kssd_build_index(const uint64_t *keys, const uint64_t *index, int n, unsigned obj_bits,
		     int nthreads, kssd_ctxgidobj_t *out);

/* Context totals (AF denominators): command_ani.c:337-401, :1533-1548. */
uint32_t kssd_ctx_runs(const uint64_t *a, size_t n, unsigned obj_bits, int ignoreconflict);
void kssd_sample_ctx_counts(const uint64_t *keys, const uint64_t *index, int n, unsigned obj_bits,
			    int conflict, int ignoreconflict, uint32_t *counts);
/* The same totals for a conflict-keeping reference, read off its inverted
 * index (ref_ctx_counts_from_ctxgidobj, command_ani.c:351-370). */
void kssd_index_ctx_counts(const kssd_ctxgidobj_t *rec, size_t n, const uint64_t *index, int ref_gnum,
			   int ignoreconflict, uint32_t *counts);

#ifdef __cplusplus
}
#endif
#endif
