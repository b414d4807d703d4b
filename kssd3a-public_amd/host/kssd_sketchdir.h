/* Sketch-directory I/O (formats of the reference's sketch directories). */
#ifndef KSSD_SKETCHDIR_H
#define KSSD_SKETCHDIR_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#include "kssd_report.h"
#include "kssdhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KSSD_STAT_FILE "lcofiles.stat"
#define KSSD_COMB_FILE "comblco"
#define KSSD_INDEX_FILE "comblco.index"
#define KSSD_SORTED_INDEX_FILE "sortedcomb_ctxgid64obj32"
#define KSSD_ANNO_FILE "lcofiles.anno"
#define KSSD_META_FILE "lcofiles.infilemeta"
#define KSSD_ABUND_FILE "comblco.a"
#define KSSD_QC_FILE "lcofiles.qc"
#define KSSD_UNIQUE_BITS_FILE "sortedcomb_ctxgid64obj32.unique_bits"

/* dim_sketch_stat_t, global_basic.h:138-149 (32 bytes) */
typedef struct {
	uint32_t hash_id;
	bool koc;
	bool conflict;
	int32_t coden_len, klen, hclen, holen, drfold, infile_num;
} kssd_stat_t;

enum { KSSD_SK_KEYS = 1, KSSD_SK_RECORDS = 2, KSSD_SK_ANNO = 4, KSSD_SK_META = 8 };

typedef struct {
	kssd_stat_t stat;
	int n;
	unsigned ctx_bits, obj_bits;
	char (*names)[KSSD_PATHLEN];
	uint64_t *index;             /* n + 1 */
	uint64_t *keys;              /* index[n] (KSSD_SK_KEYS) */
	kssd_ctxgidobj_t *records;   /* sorted index when present (KSSD_SK_RECORDS) */
	char (*anno)[KSSD_PATHLEN];  /* optional */
	kssd_infile_meta_t *meta;    /* optional */
} kssd_sketchdir_t;

bool kssd_file_in_dir(const char *dir, const char *name);
void *kssd_read_file(const char *path, size_t *size);
/* 0 on success; on failure err holds a message. */
int kssd_sketchdir_open(const char *dir, unsigned parts, kssd_sketchdir_t *s, char *err, size_t errlen);
void kssd_sketchdir_close(kssd_sketchdir_t *s);
/* lcofiles.stat + comblco + comblco.index */
int kssd_sketchdir_write(const char *dir, const kssd_stat_t *stat, char (*names)[KSSD_PATHLEN],
			 const uint64_t *keys, const uint64_t *index);
/* sortedcomb_ctxgid64obj32 + .unique_bits (computed here) */
int kssd_write_sorted_index(const char *dir, const kssd_ctxgidobj_t *rec, size_t n);
/* the same with the unique bits given ((n + 7) / 8 bytes; NULL: computed) */
int kssd_write_index_files(const char *dir, const kssd_ctxgidobj_t *rec, size_t n, const uint8_t *unique_bits);

/* Per-input metadata of a sketched file (sketch_infile_meta_stat_for_path,
 * command_sketch.c:1229-1240 with infile_meta_from_lengths :1134-1174):
 * FASTA record lengths summarised, format and flags from the file name.
 * 0, or -1 when a FASTA file cannot be read. */
int kssd_infile_meta_for_path(const char *path, kssd_infile_meta_t *out);
/* The summary infile_meta_from_lengths (command_sketch.c:1134-1173) makes of
 * record lengths (n may be 0; lens is sorted in place). */
void kssd_infile_meta_from_lengths(uint64_t *lens, size_t n, int fmt, int create_type, unsigned flags,
				   kssd_infile_meta_t *out);
/* Lengths of the non-empty records of a FASTA file appended to *lens
 * (grown with realloc; *n entries, *cap capacity); 0 or -1. */
int kssd_fasta_record_lengths(const char *path, uint64_t **lens, size_t *n, size_t *cap);
/* infile_flags_from_path without a pipe command (command_sketch.c:1106-1116) */
unsigned kssd_infile_flags_from_path(const char *path);
/* Write one whole file of a sketch directory; 0 or -1 (errno set). */
int kssd_write_dir_file(const char *dir, const char *name, const void *data, size_t n);
/* lcofiles.infilemeta */
int kssd_write_infile_meta(const char *dir, const kssd_infile_meta_t *meta, int n);
/* Input file kind by name (isOK_fmt_infile with fasta_fmt / fastq_fmt,
 * global_basic.c:166-189, global_basic.h:214-250): 1 FASTA, 2 FASTQ, 0 other. */
int kssd_infile_fmt_from_path(const char *path);

#ifdef __cplusplus
}
#endif
#endif
