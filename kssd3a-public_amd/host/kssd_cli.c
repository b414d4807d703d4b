/*
 * kssd3a-hip: command-line front end of the comparison path.
 *
 *   kssd3a-hip ani [options]   the reference's `kssd3a ani` for sketch
 *                              directories: options of command_ani_wrapper.c:77-135,
 *                              path choice of run_sketch_ani (:1082-1230),
 *                              every count computed by libkssdhip on the GPU
 *   kssd3a-hip matrix [opts]   the reference's `kssd3a matrix` distance matrices
 *                              (full / triangle, one metric per run)
 *   kssd3a-hip sketch -o DIR F  sketch FASTA/FASTQ files, one sample each, with the
 *                              coden pattern (command_sketch.c:3913-3949, :989-1005,
 *                              :1134-1240 for lcofiles.infilemeta)
 *   kssd3a-hip sketch -i DIR   inverted index of a sketch directory
 *                              (gen_inverted_index4comblco, command_sketch.c:916-938)
 *
 * Output bytes follow the reference path that the same inputs and options
 * select (index, single/multi-query stream, hash, direct or indexed matrix).
 * The reference's auto-sketching inputs (raw FASTA/FASTQ, --reflist,
 * --qrylist, --pair) and the --unique-ctx-report side output are not offered.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <err.h>
#include <errno.h>
#include <getopt.h>
#include <limits.h>
#include <omp.h>
#include <pthread.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "kssd_host.h"
#include "kssd_report.h"
#include "kssd_sketchdir.h"
#include "kssdhip.h"

#define PROG "kssd3a-hip"
#define EX_USAGE 64

/* Every error exit first waits for the thread that may be starting the HIP
 * runtime (hip_start), so the process never exits while it runs. */
static void hip_settle(void);
static void stage(const char *what);
static void __attribute__((noreturn, format(printf, 2, 3))) cli_errx(int code, const char *fmt, ...)
{
	hip_settle();
	va_list ap;
	va_start(ap, fmt);
	verrx(code, fmt, ap);
}
static void __attribute__((noreturn, format(printf, 2, 3))) cli_err(int code, const char *fmt, ...)
{
	const int e = errno;
	hip_settle();
	errno = e;
	va_list ap;
	va_start(ap, fmt);
	verr(code, fmt, ap);
}
#define errx(...) cli_errx(__VA_ARGS__)
#define err(...) cli_err(__VA_ARGS__)

/* ------------------------------------------------------------ the GPU library */
/* libkssdhip.so (beside this binary) is opened on the first GPU call, so the
 * host-only commands (sketch -o, sketch -i --host-index) start without
 * loading the HIP runtime. */
static struct {
	const char *(*last_error)(void);
	int (*init)(int);
	int (*index_open)(const kssd_ctxgidobj_t *, size_t, int, unsigned, kssdhip_index **);
	int (*index_close)(kssdhip_index *);
	int (*build_index)(const uint64_t *, const uint64_t *, int, unsigned, kssd_ctxgidobj_t *, uint8_t *);
	int (*count_block)(kssdhip_index *, const uint64_t *, const uint64_t *, int, int, uint32_t, kssd_ctx_mut2_t *,
			   kssd_obj_section_t *, kssd_ctx_mut2_t *, kssd_obj_section_t *);
	int (*count_votes)(kssdhip_index *, const uint64_t *, const uint64_t *, int, int, uint32_t, uint32_t,
			   uint32_t *);
	int (*select_block)(kssdhip_index *, const uint64_t *, const uint64_t *, int, uint32_t, const uint32_t *,
			    const uint32_t *, const kssdhip_select_opts_t *, kssdhip_ani_row_t *, size_t, size_t *);
	int (*autotune_host)(kssdhip_index *, const uint64_t *, const uint64_t *, int, uint32_t, const char *, int,
			     kssdhip_tune_result_t *);
	int (*tiling_from_cache)(kssdhip_index *, int, uint64_t, uint32_t, const char *, kssdhip_tune_result_t *);
} g_hip;

static void *hip_sym(void *lib, const char *name)
{
	void *f = dlsym(lib, name);
	if (!f)
		errx(EXIT_FAILURE, "libkssdhip.so: missing %s", name);
	return f;
}

static typeof(g_hip) *hip_load(void)
{
	if (g_hip.init)
		return &g_hip;
	char exe[PATH_MAX];
	const ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 32);
	if (n <= 0)
		err(EXIT_FAILURE, "cannot locate %s", PROG);
	exe[n] = '\0';
	char *slash = strrchr(exe, '/');
	strcpy(slash ? slash + 1 : exe, "libkssdhip.so");
	void *lib = dlopen(exe, RTLD_NOW | RTLD_LOCAL);
	if (!lib)
		errx(EXIT_FAILURE, "GPU library: %s", dlerror());
	g_hip.last_error = (const char *(*)(void))hip_sym(lib, "kssdhip_last_error");
	g_hip.index_open = (typeof(g_hip.index_open))hip_sym(lib, "kssdhip_index_open");
	g_hip.index_close = (typeof(g_hip.index_close))hip_sym(lib, "kssdhip_index_close");
	g_hip.build_index = (typeof(g_hip.build_index))hip_sym(lib, "kssdhip_build_index");
	g_hip.count_block = (typeof(g_hip.count_block))hip_sym(lib, "kssdhip_count_block");
	g_hip.count_votes = (typeof(g_hip.count_votes))hip_sym(lib, "kssdhip_count_votes");
	g_hip.select_block = (typeof(g_hip.select_block))hip_sym(lib, "kssdhip_select_block");
	g_hip.autotune_host = (typeof(g_hip.autotune_host))hip_sym(lib, "kssdhip_index_autotune_host");
	g_hip.tiling_from_cache = (typeof(g_hip.tiling_from_cache))hip_sym(lib, "kssdhip_index_tiling_from_cache");
	g_hip.init = (int (*)(int))hip_sym(lib, "kssdhip_init");
	return &g_hip;
}

/* The GPU library loaded and the HIP runtime started on a thread of its own
 * (about 25 ms) while the command reads its inputs; the first GPU call waits
 * for it and selects the device on the calling thread (the library keeps the
 * device and its error text per thread). */
static struct {
	pthread_t thread;
	int device;
	bool pending;
} g_hip_start;

static void *hip_start_main(void *arg)
{
	(void)arg;
	(void)hip_load()->init(g_hip_start.device);
	return NULL;
}

static void hip_start(int device)
{
	g_hip_start.device = device;
	if (!g_hip.init && pthread_create(&g_hip_start.thread, NULL, hip_start_main, NULL) == 0)
		g_hip_start.pending = true;
}

static void hip_settle(void)
{
	if (g_hip_start.pending && !pthread_equal(pthread_self(), g_hip_start.thread)) {
		g_hip_start.pending = false;
		pthread_join(g_hip_start.thread, NULL);
	}
}

static typeof(g_hip) *hip_api(void)
{
	if (g_hip_start.pending) {
		hip_settle();
		if (hip_load()->init(g_hip_start.device) != 0)
			errx(EXIT_FAILURE, "no usable GPU: %s", g_hip.last_error());
		stage("GPU ready");
	}
	return hip_load();
}
#define kssdhip_last_error() (hip_api()->last_error())
#define kssdhip_init(...) (hip_api()->init(__VA_ARGS__))
#define kssdhip_index_open(...) (hip_api()->index_open(__VA_ARGS__))
#define kssdhip_index_close(...) (hip_api()->index_close(__VA_ARGS__))
#define kssdhip_build_index(...) (hip_api()->build_index(__VA_ARGS__))
#define kssdhip_count_block(...) (hip_api()->count_block(__VA_ARGS__))
#define kssdhip_count_votes(...) (hip_api()->count_votes(__VA_ARGS__))
#define kssdhip_select_block(...) (hip_api()->select_block(__VA_ARGS__))

/* ------------------------------------------------------------ small helpers */
static const char *g_cmd = "";
static int g_timing;
static double g_t0;

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void stage(const char *what)
{
	if (g_timing)
		fprintf(stderr, "[%s %s] %8.3f s  %s\n", PROG, g_cmd, now_s() - g_t0, what);
}

static void __attribute__((noreturn, format(printf, 1, 2))) usage_error(const char *fmt, ...)
{
	hip_settle();
	va_list ap;
	fprintf(stderr, "%s %s: ", PROG, g_cmd);
	va_start(ap, fmt);
	vfprintf(stderr, fmt, ap);
	va_end(ap);
	fprintf(stderr, "\nTry `%s %s --help' for more information.\n", PROG, g_cmd);
	exit(EX_USAGE);
}

static long parse_int(const char *name, const char *arg, long lo, long hi)
{
	char *end = NULL;
	errno = 0;
	const long v = strtol(arg, &end, 10);
	if (errno || end == arg || *end || v < lo || v > hi)
		usage_error("%s requires an integer in range %ld..%ld", name, lo, hi);
	return v;
}

static double parse_double(const char *name, const char *arg, double lo, double hi)
{
	char *end = NULL;
	errno = 0;
	const double v = strtod(arg, &end);
	if (errno || end == arg || *end || !(v >= lo && v <= hi))
		usage_error("%s requires a number in range %g..%g", name, lo, hi);
	return v;
}

static void copy_path(char *dst, size_t cap, const char *name, const char *arg)
{
	if (strlen(arg) >= cap)
		usage_error("%s path is too long", name);
	strcpy(dst, arg);
}

static long env_int(const char *name, long dflt, long lo, const char *what)
{
	const char *env = getenv(name);
	if (!env || !env[0])
		return dflt;
	char *end = NULL;
	errno = 0;
	const long v = strtol(env, &end, 10);
	if (errno || end == env || *end || v < lo || v > INT_MAX)
		errx(EXIT_FAILURE, "invalid %s='%s' (use %s)", name, env, what);
	return v;
}

static bool env_flag(const char *name)
{
	const char *env = getenv(name);
	return env && env[0] && strcmp(env, "0") != 0;
}

static bool same_file(const char *a, const char *b)
{
	struct stat sa, sb;
	return a && b && stat(a, &sa) == 0 && stat(b, &sb) == 0 && sa.st_dev == sb.st_dev &&
	       sa.st_ino == sb.st_ino;
}

static void *xmalloc(size_t n)
{
	void *p = malloc(n ? n : 1);
	if (!p)
		err(EXIT_FAILURE, "out of memory (%zu bytes)", n);
	return p;
}

static void *xrealloc(void *p, size_t n)
{
	void *q = realloc(p, n ? n : 1);
	if (!q)
		err(EXIT_FAILURE, "out of memory (%zu bytes)", n);
	return q;
}

static void *xcalloc(size_t n, size_t sz)
{
	void *p = calloc(n ? n : 1, sz ? sz : 1);
	if (!p)
		err(EXIT_FAILURE, "out of memory (%zu x %zu bytes)", n, sz);
	return p;
}

static void open_dir(const char *dir, unsigned parts, kssd_sketchdir_t *s)
{
	char msg[512];
	if (kssd_sketchdir_open(dir, parts, s, msg, sizeof msg) != 0)
		errx(EXIT_FAILURE, "%s", msg);
}

static void hip_check(int rc, const char *what)
{
	if (rc != 0)
		errx(EXIT_FAILURE, "%s: %s", what, kssdhip_last_error());
}

/* The kernel shape for a block of this shape: the tuning cache's entry
 * (kssdhip_index_tiling_from_cache), or with KSSDHIP_TUNE=1 the library's
 * autotuner timed on this block, which caches it for later runs.  Without
 * either the handle keeps the shape it was opened with. */
static void tune_block(kssdhip_index *ix, const uint64_t *keys, const uint64_t *skidx, int nb, uint32_t flags)
{
	kssdhip_tune_result_t r;
	const char *e = getenv("KSSDHIP_TUNE");
	if (nb < 1)
		return;
	if (e && e[0] == '1' && e[1] == '\0')
		hip_check(hip_api()->autotune_host(ix, keys, skidx, nb, flags, NULL, 0, &r), "autotune");
	else if (hip_api()->tiling_from_cache(ix, nb, skidx[nb] - skidx[0], flags, NULL, &r) < 0)
		hip_check(-1, "tuning cache");
}

/* ------------------------------------------------------------ options */
typedef struct {
	kssd_report_opts_t r;
	char refdir[PATH_MAX], qrydir[PATH_MAX], outf[PATH_MAX];
	int p;
	int device;
	int values; /* 0 unset, 1 distance, -1 ani */
	bool afcut_set;
} ani_args_t;

enum {
	OPT_IGNORECONFLICT = 256,
	OPT_METRIC,
	OPT_MAX_PDIST,
	OPT_MAX_SECTION,
	OPT_UNIFIED,
	OPT_FORMAT,
	OPT_VALUES,
	OPT_RAW,
	OPT_UNIQUE,
	OPT_COVERAGE,
	OPT_DEVICE,
	OPT_HELP,
	OPT_NOT_OFFERED,
};

static const char ani_help[] =
	"Usage: " PROG " ani [-r REF_SKETCH] (-q QRY_SKETCH | --qraw RAW_QRY_SKETCH) [options]\n"
	"       " PROG " ani SELF_SKETCH -m1|-m2 [options]\n"
	"ANI between sketch directories; counting runs on the GPU.\n\n"
	"  -r, --ref=DIR            reference sketch directory\n"
	"  -q, --query=DIR          query sketch directory\n"
	"  -x, --qraw=DIR           raw-read query sketch directory (naive ANI mode)\n"
	"  -v, --naive              naive distance/ANI\n"
	"      --ignoreconflict     ignore reference contexts with conflicting objects\n"
	"  -f, --afcut=F            min max(Qry,Ref) align fraction [0.5; 0.2 with --qraw]\n"
	"  -n, --anicut=F           min ANI [0.95]\n"
	"  -t, --ctxcut=N           min shared contexts [3]\n"
	"  -s, --slmetrics=+-1..9   Best(1) Recalibrated(2) CtxMoE(3) Naive(4) MashD(5) AafD(6)\n"
	"                           MashD_if_far(7) AafD_if_far(8) p_dist(9) [1]\n"
	"      --metric=NAME        best, recalibrated, ctx-moe, ctx-naive, mash, aaf,\n"
	"                           mash-if-far, aaf-if-far, p_dist\n"
	"      --max-pdist=F        close-only detail filter on p_dist\n"
	"      --max-diff-section=N close-only detail filter on object-section differences\n"
	"      --unified-metric     honour -s in --qraw mode\n"
	"  -d, --diagonal           print the diagonal in triangle output\n"
	"  -e, --exception=N        missing/low-overlap matrix distance [1]\n"
	"  -m, --outfmt=0|1|2       detail, matrix or triangle [0]\n"
	"      --format=NAME        detail, matrix or triangle\n"
	"      --values=distance|ani  matrix/triangle values\n"
	"  -o, --outfile=FILE       output file [stdout]\n"
	"      --raw-output         skip calibrated/best ANI\n"
	"      --unique-ctx         adaptive reference-unique contexts (indexed -r, --qraw)\n"
	"      --estimate-coverage  coverage columns from the query's comblco.a (--qraw, detail)\n"
	"  -N, --top=N              at most N references per query [all]\n"
	"  -p, --threads=N          host threads for indexing, ranking and formatting [all]\n"
	"      --device=N           HIP device ordinal [0]\n";

static void parse_ani(int argc, char **argv, ani_args_t *a)
{
	static const struct option longopts[] = {
		{"ref", required_argument, NULL, 'r'},
		{"query", required_argument, NULL, 'q'},
		{"qraw", required_argument, NULL, 'x'},
		{"naive", no_argument, NULL, 'v'},
		{"ignoreconflict", no_argument, NULL, OPT_IGNORECONFLICT},
		{"afcut", required_argument, NULL, 'f'},
		{"anicut", required_argument, NULL, 'n'},
		{"metric", required_argument, NULL, OPT_METRIC},
		{"ctxcut", required_argument, NULL, 't'},
		{"slmetrics", required_argument, NULL, 's'},
		{"max-pdist", required_argument, NULL, OPT_MAX_PDIST},
		{"max-diff-section", required_argument, NULL, OPT_MAX_SECTION},
		{"unified-metric", no_argument, NULL, OPT_UNIFIED},
		{"diagonal", no_argument, NULL, 'd'},
		{"exception", required_argument, NULL, 'e'},
		{"outfmt", required_argument, NULL, 'm'},
		{"format", required_argument, NULL, OPT_FORMAT},
		{"values", required_argument, NULL, OPT_VALUES},
		{"outfile", required_argument, NULL, 'o'},
		{"raw-output", no_argument, NULL, OPT_RAW},
		{"unique-ctx", no_argument, NULL, OPT_UNIQUE},
		{"estimate-coverage", no_argument, NULL, OPT_COVERAGE},
		{"top", required_argument, NULL, 'N'},
		{"threads", required_argument, NULL, 'p'},
		{"device", required_argument, NULL, OPT_DEVICE},
		{"help", no_argument, NULL, OPT_HELP},
		/* reference options this front end does not offer */
		{"reflist", required_argument, NULL, OPT_NOT_OFFERED},
		{"qrylist", required_argument, NULL, OPT_NOT_OFFERED},
		{"unique-ctx-report", required_argument, NULL, OPT_NOT_OFFERED},
		{"glist", required_argument, NULL, OPT_NOT_OFFERED},
		{"control", required_argument, NULL, OPT_NOT_OFFERED},
		{"pair", no_argument, NULL, OPT_NOT_OFFERED},
		{"pipecmd", required_argument, NULL, OPT_NOT_OFFERED},
		{NULL, 0, NULL, 0},
	};
	memset(a, 0, sizeof *a);
	kssd_report_opts_t *o = &a->r;
	/* ani_opt defaults, command_ani_wrapper.c:174-221 */
	o->fmt = 0;
	o->s = 1;
	o->max_diff_obj_section = -1;
	o->e = 1;
	o->ntop = -1;
	o->ctxcut = 3;
	o->afcut = 0.5f;
	o->anicut = 0.95f;
	o->drfold = 8;
	a->p = omp_get_max_threads();
	optind = 1;
	int c;
	while ((c = getopt_long(argc, argv, "r:q:x:vf:n:t:s:de:m:o:N:p:", longopts, NULL)) != -1) {
		switch (c) {
		case 'r':
			copy_path(a->refdir, sizeof a->refdir, "-r/--ref", optarg);
			break;
		case 'q':
			copy_path(a->qrydir, sizeof a->qrydir, "-q/--query", optarg);
			break;
		case 'x':
			copy_path(a->qrydir, sizeof a->qrydir, "--qraw", optarg);
			o->unassembled = true;
			o->v = true;
			break;
		case 'v':
			o->v = true;
			break;
		case OPT_IGNORECONFLICT:
			o->ignoreconflict = true;
			break;
		case 'f':
			o->afcut = (float)parse_double("-f/--afcut (fraction, e.g. 0.8, not 80)", optarg, 0.0, 1.0);
			a->afcut_set = true;
			break;
		case 'n':
			o->anicut = (float)parse_double("-n/--anicut (fraction, e.g. 0.95, not 95)", optarg, 0.0, 1.0);
			break;
		case OPT_METRIC: {
			static const char *const names[9] = {"best", "recalibrated", "ctx-moe", "ctx-naive", "mash",
							     "aaf", "mash-if-far", "aaf-if-far", "p_dist"};
			int m = 0;
			for (int i = 0; i < 9; ++i)
				if (strcmp(optarg, names[i]) == 0)
					m = i + 1;
			if (strcmp(optarg, "naive") == 0)
				m = 4;
			if (strcmp(optarg, "pdist") == 0)
				m = 9;
			if (!m)
				usage_error("unknown --metric '%s'; use best, recalibrated, ctx-moe, ctx-naive, p_dist, "
					    "mash, aaf, mash-if-far, or aaf-if-far", optarg);
			o->s = o->s < 0 ? -m : m;
			break;
		}
		case 't':
			o->ctxcut = (int)parse_int("-t/--ctxcut", optarg, 0, INT_MAX);
			break;
		case 's':
			o->s = (int)parse_int("-s/--slmetrics", optarg, -9, 9);
			if (o->s == 0)
				usage_error("-s/--slmetrics accepts -9..-1 or 1..9, but not 0");
			break;
		case OPT_MAX_PDIST:
			o->max_pdist = parse_double("--max-pdist", optarg, 0.0, 1.0);
			o->max_pdist_set = true;
			break;
		case OPT_MAX_SECTION:
			o->max_diff_obj_section = (int)parse_int("--max-diff-section", optarg, 0, INT_MAX);
			break;
		case OPT_UNIFIED:
			o->unified_metric = true;
			break;
		case 'd':
			o->d = true;
			break;
		case 'e':
			o->e = (int)parse_int("-e/--exception", optarg, 1, INT_MAX);
			break;
		case 'm':
			o->fmt = (int)parse_int("-m/--outfmt", optarg, 0, 2);
			break;
		case OPT_FORMAT:
			if (strcmp(optarg, "detail") == 0)
				o->fmt = 0;
			else if (strcmp(optarg, "matrix") == 0)
				o->fmt = 1;
			else if (strcmp(optarg, "triangle") == 0)
				o->fmt = 2;
			else
				usage_error("--format must be detail, matrix, or triangle");
			break;
		case OPT_VALUES:
			if (strcmp(optarg, "distance") == 0)
				a->values = 1;
			else if (strcmp(optarg, "ani") == 0)
				a->values = -1;
			else
				usage_error("--values must be distance or ani");
			break;
		case 'o':
			copy_path(a->outf, sizeof a->outf, "-o/--outfile", optarg);
			break;
		case OPT_RAW:
			o->raw_output = true;
			break;
		case OPT_UNIQUE:
			o->unique_ctx = true;
			break;
		case OPT_COVERAGE:
			o->estimate_coverage = true;
			break;
		case 'N':
			o->ntop = (int)parse_int("-N/--top", optarg, 1, INT_MAX);
			break;
		case 'p':
			a->p = (int)parse_int("-p/--threads", optarg, 1, 65536);
			break;
		case OPT_DEVICE:
			a->device = (int)parse_int("--device", optarg, 0, 1024);
			break;
		case OPT_HELP:
			fputs(ani_help, stdout);
			exit(0);
		case OPT_NOT_OFFERED:
			usage_error("%s is not offered by this front end (sketch directories only)", argv[optind - 1]);
		default:
			fprintf(stderr, "Try `%s %s --help' for more information.\n", PROG, g_cmd);
			exit(EX_USAGE);
		}
	}
	/* end-of-arguments checks, command_ani_wrapper.c:538-620 */
	if (a->values)
		o->s = abs(o->s) * a->values;
	const int npos = argc - optind;
	if (o->unique_ctx) {
		if (!o->unassembled || o->fmt != 0 || !a->refdir[0] || !a->qrydir[0] || npos)
			usage_error("unique-context options require explicit -r/--qraw inputs and detail output");
		if (o->unified_metric || o->raw_output || abs(o->s) > 4 || o->max_pdist_set ||
		    o->max_diff_obj_section >= 0)
			usage_error("--unique-ctx uses calibrated read-mode Naive ANI; incompatible metric, raw-output, "
				    "unified-metric or close-pair filter");
	}
	if (o->estimate_coverage && o->fmt != 0)
		usage_error("--estimate-coverage is supported only with detail output (-m0).");
	if (o->estimate_coverage && !o->unassembled)
		usage_error("--estimate-coverage currently requires --qraw raw/unassembled query mode.");
	if ((o->max_pdist_set || o->max_diff_obj_section >= 0) && o->fmt != 0)
		usage_error("--max-pdist/--max-diff-section are supported only with detail output (-m0).");
	if (!a->refdir[0] && !a->qrydir[0] && npos >= 2)
		usage_error("positional reference/query inputs (auto-sketch) are not offered by this front end");
	else if (!a->refdir[0] && !a->qrydir[0] && npos == 1 && o->fmt != 0)
		copy_path(a->qrydir, sizeof a->qrydir, "sketch", argv[optind]);
	else if (!a->qrydir[0])
		usage_error("missing query: use -q, --qraw, or positional reference and query inputs");
	else if (!a->refdir[0] && o->fmt == 0)
		usage_error("-r/--ref is required for detail output; use -m1 or -m2 for one-sketch self matrices");
	else if (npos)
		usage_error("unexpected argument '%s'", argv[optind]);
}

/* ------------------------------------------------------------ output */
/* print_ani_detail_header, command_ani.c:190-196 */
static void put_detail_header(FILE *fp, const ani_args_t *a)
{
	if (a->r.estimate_coverage)
		fprintf(fp, "%s\t%s\n", kssd_detail_header, kssd_coverage_header);
	else
		fprintf(fp, "%s\n", kssd_detail_header);
}

static FILE *open_out(const ani_args_t *a)
{
	if (!a->outf[0])
		return stdout;
	FILE *fp = fopen(a->outf, "w");
	if (!fp)
		err(errno, "%s", a->outf);
	return fp;
}

static void close_out(FILE *fp)
{
	if (fflush(fp) != 0 || ferror(fp))
		err(EXIT_FAILURE, "Cannot write ANI output");
	if (fp != stdout && fclose(fp) != 0)
		err(EXIT_FAILURE, "Cannot finish ANI output");
}

static void put_buf(FILE *fp, kssd_buf_t *b)
{
	if (b->l && fwrite(b->s, 1, b->l, fp) != b->l)
		err(EXIT_FAILURE, "Cannot write ANI output");
	b->l = 0;
}

static void put_header_names(FILE *fp, const kssd_sketchdir_t *s)
{
	for (int i = 0; i < s->n; ++i)
		fprintf(fp, "\t%s", s->names[i]);
	fputc('\n', fp);
}

/* ------------------------------------------------------------ GPU counting */
/* Cells (query x reference) per counting block: dense counters are 16 B per
 * cell (32 B with the unique-owner copies); the block keeps them near 512 MiB. */
static int block_queries(int nq, int nref, bool unique)
{
	const size_t budget = (size_t)512 << 20;
	const size_t per_q = (size_t)(nref ? nref : 1) * (unique ? 32u : 16u);
	size_t b = budget / per_q;
	if (b < 1)
		b = 1;
	return b < (size_t)nq ? (int)b : nq;
}

/* Queries from q0 (of q_end) in one GPU block: at most blk (the dense rows'
 * budget), the block's keys within 2^28 (2 GiB on the device) and the
 * partitioned join's split table -- one 32-bit offset per query and index
 * partition, at most index_records / 1024 + 2 partitions -- within 512 MiB;
 * at least one query. */
static int block_from(const uint64_t *index, int q0, int q_end, int blk, size_t index_records)
{
	const uint64_t key_cap = (uint64_t)1 << 28;
	const size_t split_cap = ((size_t)512 << 20) / ((index_records / 1024 + 2) * 4);
	int nb = 1;
	while (q0 + nb < q_end && nb < blk && (size_t)nb < split_cap && index[q0 + nb + 1] - index[q0] <= key_cap)
		++nb;
	return nb;
}

static kssdhip_index *gpu_index(const kssd_ctxgidobj_t *rec, size_t n, int gnum, unsigned obj_bits)
{
	kssdhip_index *ix = NULL;
	hip_check(kssdhip_index_open(rec, n, gnum, obj_bits, &ix), "cannot open the index on the GPU");
	return ix;
}

/* The sorted index of a sketch without one (the in-memory index of the
 * reference's indexed paths, comb_sortedsketch64_2sortedcomb_ctxgid64obj32,
 * command_ani.c:733-746), built on the GPU; unique_bits: (n + 7) / 8 bytes
 * or NULL. */
static kssd_ctxgidobj_t *build_records_bits(const kssd_sketchdir_t *s, uint8_t *unique_bits)
{
	const size_t n = s->index[s->n];
	kssd_ctxgidobj_t *rec = xmalloc(n ? n * sizeof *rec : 1);
	hip_check(kssdhip_build_index(s->keys, s->index, s->n, s->obj_bits, rec, unique_bits),
		  "cannot build the index on the GPU");
	return rec;
}

static kssd_ctxgidobj_t *build_records(const kssd_sketchdir_t *s, int threads)
{
	(void)threads;
	return build_records_bits(s, NULL);
}

/* gen_inverted_index4comblco (command_sketch.c:916-938): write the index
 * and its unique-marker bits next to the sketch; on the GPU, or with the
 * host builder (host_index). */
static void write_index_for(const char *dir, int threads, bool host_index, int device)
{
	kssd_sketchdir_t s;
	if (!host_index)
		hip_start(device);
	open_dir(dir, KSSD_SK_KEYS, &s);
	const size_t n = s.index[s.n];
	kssd_ctxgidobj_t *rec;
	uint8_t *bits = NULL;
	if (host_index) {
		rec = xmalloc(n ? n * sizeof *rec : 1);
		if (kssd_build_index(s.keys, s.index, s.n, s.obj_bits, threads, rec) != 0)
			errx(EXIT_FAILURE, "cannot index %d samples with %zu entries (limits: 2^20 samples, 2^32 entries)",
			     s.n, n);
	} else {
		hip_check(kssdhip_init(device), "no usable GPU (--host-index builds the index on the host)");
		bits = xmalloc((n + 7) / 8 + 1);
		rec = build_records_bits(&s, bits);
	}
	if (kssd_write_index_files(dir, rec, n, bits) != 0)
		err(EXIT_FAILURE, "%s/%s", dir, KSSD_SORTED_INDEX_FILE);
	free(bits);
	free(rec);
	kssd_sketchdir_close(&s);
}

/* Keys of `s` without the contexts that carry more than one object (the
 * reference side of --ignoreconflict, command_ani.c:2905-2909). */
static uint64_t *drop_conflict_runs(const kssd_sketchdir_t *s, uint64_t *index_out)
{
	uint64_t *out = xmalloc(s->index[s->n] * sizeof *out);
	size_t w = 0;
	index_out[0] = 0;
	for (int g = 0; g < s->n; ++g) {
		const uint64_t *a = s->keys + s->index[g];
		const size_t len = (size_t)(s->index[g + 1] - s->index[g]);
		for (size_t i = 0; i < len;) {
			const uint64_t c = a[i] >> s->obj_bits;
			const size_t b = i;
			do {
				++i;
			} while (i < len && (a[i] >> s->obj_bits) == c);
			if (i - b == 1)
				out[w++] = a[b];
		}
		index_out[g + 1] = w;
	}
	return out;
}

/*
 * Counts of every (query, reference) pair into ctx/obj [Q][R].  The side
 * with fewer entries becomes the device index and the other side streams
 * through it in blocks; the four counts are symmetric in the two sketches,
 * and the reference-side --ignoreconflict rule is applied either by the
 * kernel flag (reference indexed) or by dropping those contexts from the
 * streamed reference keys.
 */
static void rect_counts(const kssd_sketchdir_t *ref, const kssd_sketchdir_t *qry, bool ignoreconflict,
			int threads, kssd_ctx_mut2_t *ctx, kssd_obj_section_t *obj)
{
	const int R = ref->n, Q = qry->n;
	const bool ignore = ignoreconflict && ref->stat.conflict;
	if (R == 0 || Q == 0)
		return;
	const bool index_ref = ref->index[R] <= qry->index[Q];
	if (index_ref) {
		kssd_ctxgidobj_t *rec = build_records(ref, threads);
		kssdhip_index *ix = gpu_index(rec, ref->index[R], R, ref->obj_bits);
		free(rec);
		const int blk = block_queries(Q, R, false);
		for (int q0 = 0, nb = 0; q0 < Q; q0 += nb) {
			nb = block_from(qry->index, q0, Q, blk, (size_t)ref->index[R]);
			tune_block(ix, qry->keys + qry->index[q0], qry->index + q0, nb, ignore ? KSSDHIP_IGNORECONFLICT : 0);
			hip_check(kssdhip_count_block(ix, qry->keys + qry->index[q0], qry->index + q0, nb, q0,
						      ignore ? KSSDHIP_IGNORECONFLICT : 0, ctx + (size_t)q0 * R,
						      obj + (size_t)q0 * R, NULL, NULL),
				  "GPU counting failed");
		}
		kssdhip_index_close(ix);
		return;
	}
	kssd_ctxgidobj_t *rec = build_records(qry, threads);
	kssdhip_index *ix = gpu_index(rec, qry->index[Q], Q, qry->obj_bits);
	free(rec);
	uint64_t *rindex = NULL, *rkeys = ref->keys;
	if (ignore) {
		rindex = xmalloc(((size_t)R + 1) * sizeof *rindex);
		rkeys = drop_conflict_runs(ref, rindex);
	}
	const uint64_t *ridx = ignore ? rindex : ref->index;
	const int blk = block_queries(R, Q, false);
	kssd_ctx_mut2_t *tc = xmalloc((size_t)blk * Q * sizeof *tc);
	kssd_obj_section_t *to = xmalloc((size_t)blk * Q * sizeof *to);
	for (int r0 = 0, nb = 0; r0 < R; r0 += nb) {
		nb = block_from(ridx, r0, R, blk, (size_t)qry->index[Q]);
		hip_check(kssdhip_count_block(ix, rkeys + ridx[r0], ridx + r0, nb, r0, 0, tc, to, NULL, NULL),
			  "GPU counting failed");
		for (int i = 0; i < nb; ++i)
			for (int q = 0; q < Q; ++q) {
				ctx[(size_t)q * R + r0 + i] = tc[(size_t)i * Q + q];
				obj[(size_t)q * R + r0 + i] = to[(size_t)i * Q + q];
			}
	}
	free(tc);
	free(to);
	if (ignore) {
		free(rkeys);
		free(rindex);
	}
	kssdhip_index_close(ix);
}

/* ------------------------------------------------------------ context totals */
static uint32_t *entry_or_run_totals(const kssd_sketchdir_t *s, bool runs, bool ignoreconflict)
{
	uint32_t *t = xmalloc((size_t)(s->n ? s->n : 1) * sizeof *t);
	kssd_sample_ctx_counts(s->keys, s->index, s->n, s->obj_bits, runs, ignoreconflict, t);
	return t;
}

/* ------------------------------------------------------------ coverage inputs */
/* What --estimate-coverage reads beyond the sketches: the query's read
 * counts and the reference's unique-marker bits. */
typedef struct {
	uint32_t *qcounts; /* NULL: the query has none */
	uint8_t *bits;     /* NULL: markers come from the reported rows */
} cov_inputs_t;

/* The reference's .unique_bits when it and the sorted index are present and
 * sized for the sketch (load_coverage_unique_bits, command_ani.c:2209-2250). */
static uint8_t *load_unique_bits(const char *dir, const kssd_sketchdir_t *ref)
{
	if (!kssd_file_in_dir(dir, KSSD_UNIQUE_BITS_FILE))
		return NULL;
	if (!ref->records) {
		warnx("ignoring %s/%s because %s is missing", dir, KSSD_UNIQUE_BITS_FILE, KSSD_SORTED_INDEX_FILE);
		return NULL;
	}
	char *path = xmalloc(strlen(dir) + sizeof KSSD_UNIQUE_BITS_FILE + 2);
	sprintf(path, "%s/%s", dir, KSSD_UNIQUE_BITS_FILE);
	size_t n = 0;
	uint8_t *bits = kssd_read_file(path, &n);
	free(path);
	const size_t want = (size_t)((ref->index[ref->n] + 7) >> 3);
	if (bits && n != want) {
		warnx("ignoring %s/%s: %zu bytes, expected %zu", dir, KSSD_UNIQUE_BITS_FILE, n, want);
		free(bits);
		bits = NULL;
	}
	return bits;
}

/* The query's comblco.a: read when present in the index path (a size
 * mismatch warns and drops it, command_ani.c:538-554), and when the sketch
 * says it keeps counts in the streamed paths (global_basic.c:1123-1124). */
static uint32_t *load_query_counts(const char *dir, const kssd_sketchdir_t *qry, bool index_path)
{
	if (index_path ? !kssd_file_in_dir(dir, KSSD_ABUND_FILE) : !qry->stat.koc)
		return NULL;
	char *path = xmalloc(strlen(dir) + sizeof KSSD_ABUND_FILE + 2);
	sprintf(path, "%s/%s", dir, KSSD_ABUND_FILE);
	size_t n = 0;
	uint32_t *c = kssd_read_file(path, &n);
	const size_t want = (size_t)qry->index[qry->n] * sizeof *c;
	if (!c && !index_path)
		err(EXIT_FAILURE, "%s", path);
	if (c && n != want) {
		if (!index_path)
			errx(EXIT_FAILURE, "%s: %zu bytes, expected %zu", path, n, want);
		warnx("ignoring %s: %zu bytes, expected %zu", path, n, want);
		free(c);
		c = NULL;
	}
	free(path);
	return c;
}

static void load_coverage(const ani_args_t *a, const kssd_sketchdir_t *ref, const kssd_sketchdir_t *qry,
			  bool index_path, cov_inputs_t *cv)
{
	memset(cv, 0, sizeof *cv);
	if (!a->r.estimate_coverage)
		return;
	cv->bits = load_unique_bits(a->refdir, ref);
	cv->qcounts = load_query_counts(a->qrydir, qry, index_path);
}

/* ------------------------------------------------------------ detail paths */
typedef enum { PATH_INDEX, PATH_STREAM, PATH_MULTISTREAM, PATH_HASH } detail_path_t;

static void emit_rows(FILE *fp, const ani_args_t *a, detail_path_t path, const kssd_sketchdir_t *ref,
		      const kssd_sketchdir_t *qry, int q0, int nq, const kssd_ctx_mut2_t *ctx,
		      const kssd_obj_section_t *obj, const kssd_ctx_mut2_t *uctx, const kssd_obj_section_t *uobj,
		      const uint32_t *qry_ctx, const uint32_t *ref_ctx, const cov_inputs_t *cv)
{
	const int R = ref->n;
	kssd_buf_t *bufs = xcalloc((size_t)nq, sizeof *bufs);
	uint64_t *last_ctx = NULL; /* one-query stream: the last context each reference's walk counts */
	bool *has_last = NULL;
	if (path == PATH_STREAM) {
		last_ctx = xcalloc((size_t)(R ? R : 1), sizeof *last_ctx);
		has_last = xcalloc((size_t)(R ? R : 1), sizeof *has_last);
		const bool ignore = a->r.ignoreconflict && ref->stat.conflict;
		const unsigned ob = ref->obj_bits;
		for (int j = 0; j < R; ++j) {
			size_t e = (size_t)ref->index[j + 1];
			const size_t b = (size_t)ref->index[j];
			while (e > b) {
				const uint64_t c = ref->keys[e - 1] >> ob;
				size_t s = e - 1;
				while (s > b && (ref->keys[s - 1] >> ob) == c)
					--s;
				if (!ignore || e - s == 1) {
					last_ctx[j] = c;
					has_last[j] = true;
					break;
				}
				e = s;
			}
		}
	}
#pragma omp parallel for schedule(dynamic, 1) num_threads(a->p)
	for (int i = 0; i < nq; ++i) {
		const int q = q0 + i;
		/* the hash path computes no coverage (its rows print NA:not_computed) */
		const kssd_coverage_src_t src = {
			.qkeys = qry->keys + qry->index[q],
			.qcounts = cv->qcounts ? cv->qcounts + qry->index[q] : NULL,
			.qn = (size_t)(qry->index[q + 1] - qry->index[q]),
			.ref_keys = ref->keys,
			.ref_index = ref->index,
			.ref_records = cv->bits ? ref->records : NULL,
			.ref_unique_bits = cv->bits,
			.ref_nrec = (size_t)ref->index[R],
			.obj_bits = ref->obj_bits,
		};
		kssd_query_counts_t c = {
			.cov = a->r.estimate_coverage && path != PATH_HASH ? &src : NULL,
			.ctx = ctx + (size_t)i * R,
			.obj = obj + (size_t)i * R,
			.uctx = uctx ? uctx + (size_t)i * R : NULL,
			.uobj = uobj ? uobj + (size_t)i * R : NULL,
			.qry_ctx = qry_ctx[q],
			.ref_ctx = ref_ctx,
			.nref = R,
			.qry_meta = qry->meta ? &qry->meta[q] : NULL,
			.ref_meta = ref->meta,
		};
		uint8_t *last_shared = NULL;
		if (path == PATH_STREAM) {
			const uint64_t *qk = qry->keys + qry->index[q];
			const size_t qn = (size_t)(qry->index[q + 1] - qry->index[q]);
			const unsigned ob = qry->obj_bits;
			last_shared = xcalloc((size_t)(R ? R : 1), 1);
			for (int j = 0; j < R; ++j) {
				if (!has_last[j])
					continue;
				size_t lo = 0, hi = qn; /* first query key with context >= last_ctx[j] */
				while (lo < hi) {
					const size_t mid = lo + (hi - lo) / 2;
					if ((qk[mid] >> ob) < last_ctx[j])
						lo = mid + 1;
					else
						hi = mid;
				}
				last_shared[j] = lo < qn && (qk[lo] >> ob) == last_ctx[j];
			}
			c.last_shared = last_shared;
		}
		switch (path) {
		case PATH_INDEX:
			if (a->r.fmt)
				kssd_emit_index_matrix_row(&bufs[i], &a->r, &c, qry->names[q]);
			else
				kssd_emit_index_detail(&bufs[i], &a->r, &c, qry->names[q], ref->names, ref->anno);
			break;
		case PATH_STREAM:
		case PATH_MULTISTREAM:
			kssd_emit_stream_detail(&bufs[i], &a->r, &c, qry->names[q], ref->names, ref->anno,
						path == PATH_STREAM);
			break;
		case PATH_HASH:
			kssd_emit_hash_detail(&bufs[i], &a->r, &c, qry->names[q], ref->names, ref->anno);
			break;
		}
		free(last_shared);
	}
	free(last_ctx);
	free(has_last);
	for (int i = 0; i < nq; ++i) {
		put_buf(fp, &bufs[i]);
		kssd_buf_free(&bufs[i]);
	}
	free(bufs);
}

/* The index path's detail rows of a block from the pairs that passed the
 * device-side gate (rows ordered by query, then reference): the host ranks
 * and prints each query's candidates exactly as from dense counts. */
static void emit_index_rows(FILE *fp, const ani_args_t *a, const kssd_sketchdir_t *ref, const kssd_sketchdir_t *qry,
			    int q0, int nq, const kssdhip_ani_row_t *rows, size_t n, const uint32_t *qry_ctx,
			    const uint32_t *ref_ctx, const cov_inputs_t *cv)
{
	const int R = ref->n;
	uint32_t *ids = xmalloc((n ? n : 1) * sizeof *ids);
	kssd_ctx_mut2_t *ctx = xmalloc((n ? n : 1) * sizeof *ctx), *uctx = xmalloc((n ? n : 1) * sizeof *uctx);
	kssd_obj_section_t *obj = xmalloc((n ? n : 1) * sizeof *obj), *uobj = xmalloc((n ? n : 1) * sizeof *uobj);
	size_t *off = xcalloc((size_t)nq + 1, sizeof *off);
	for (size_t k = 0; k < n; ++k) {
		ids[k] = rows[k].ref;
		ctx[k] = rows[k].ctx;
		obj[k] = rows[k].obj;
		uctx[k] = rows[k].unique_ctx;
		uobj[k] = rows[k].unique_obj;
		off[rows[k].query + 1]++;
	}
	for (int i = 0; i < nq; ++i)
		off[i + 1] += off[i];
	kssd_buf_t *bufs = xcalloc((size_t)(nq ? nq : 1), sizeof *bufs);
#pragma omp parallel for schedule(dynamic, 16) num_threads(a->p)
	for (int i = 0; i < nq; ++i) {
		const int q = q0 + i;
		const kssd_coverage_src_t src = {
			.qkeys = qry->keys + qry->index[q],
			.qcounts = cv->qcounts ? cv->qcounts + qry->index[q] : NULL,
			.qn = (size_t)(qry->index[q + 1] - qry->index[q]),
			.ref_keys = ref->keys,
			.ref_index = ref->index,
			.ref_records = cv->bits ? ref->records : NULL,
			.ref_unique_bits = cv->bits,
			.ref_nrec = (size_t)ref->index[R],
			.obj_bits = ref->obj_bits,
		};
		const size_t b = off[i];
		kssd_query_counts_t c = {
			.cov = a->r.estimate_coverage ? &src : NULL,
			.ctx = ctx + b,
			.obj = obj + b,
			.uctx = a->r.unique_ctx ? uctx + b : NULL,
			.uobj = a->r.unique_ctx ? uobj + b : NULL,
			.qry_ctx = qry_ctx[q],
			.ref_ctx = ref_ctx,
			.nref = R,
			.qry_meta = qry->meta ? &qry->meta[q] : NULL,
			.ref_meta = ref->meta,
			.ids = ids + b,
			.ncand = (int)(off[i + 1] - b),
		};
		kssd_emit_index_detail(&bufs[i], &a->r, &c, qry->names[q], ref->names, ref->anno);
	}
	for (int i = 0; i < nq; ++i) {
		put_buf(fp, &bufs[i]);
		kssd_buf_free(&bufs[i]);
	}
	free(bufs);
	free(off);
	free(ids);
	free(ctx);
	free(obj);
	free(uctx);
	free(uobj);
}

/* Queries per block of the device-side gate: the block's keys stay below
 * 2^28 (2 GiB on the device) and its queries below 2^16. */
static int select_block_queries(const kssd_sketchdir_t *qry, int q0)
{
	const uint64_t key_cap = (uint64_t)1 << 28;
	int q = q0;
	while (q < qry->n && q - q0 < 65536 &&
	       (q == q0 || qry->index[q + 1] - qry->index[q0] <= key_cap))
		++q;
	return q - q0;
}

/* mem_eff_sorted_ctxgidobj_arrXcomb_sortedsketch64, command_ani.c:494-692 */
static int run_index_path(const ani_args_t *a)
{
	kssd_sketchdir_t ref, qry;
	open_dir(a->refdir,
		 KSSD_SK_RECORDS | KSSD_SK_ANNO | KSSD_SK_META | (a->r.estimate_coverage ? KSSD_SK_KEYS : 0), &ref);
	open_dir(a->qrydir, KSSD_SK_KEYS | KSSD_SK_META, &qry);
	if (!ref.records)
		errx(EXIT_FAILURE, "%s/%s: reference index missing", a->refdir, KSSD_SORTED_INDEX_FILE);
	cov_inputs_t cv;
	load_coverage(a, &ref, &qry, true, &cv);
	if (qry.stat.hash_id != ref.stat.hash_id)
		errx(EXIT_FAILURE, "hash_id mismatch between query and reference sketches");
	stage("sketches read");
	const int R = ref.n, Q = qry.n;
	const size_t nrec = ref.index[R];
	uint32_t *ref_ctx = xmalloc((size_t)(R ? R : 1) * sizeof *ref_ctx);
	if (ref.stat.conflict)
		kssd_index_ctx_counts(ref.records, nrec, ref.index, R, a->r.ignoreconflict, ref_ctx);
	else
		for (int g = 0; g < R; ++g)
			ref_ctx[g] = (uint32_t)(ref.index[g + 1] - ref.index[g]);
	uint32_t *qry_ctx = entry_or_run_totals(&qry, qry.stat.conflict, false);
	kssdhip_index *ix = gpu_index(ref.records, nrec, R, ref.obj_bits);
	stage("index on the GPU");

	FILE *fp = open_out(a);
	if (a->r.fmt)
		put_header_names(fp, &ref);
	else
		put_detail_header(fp, a);
	const bool unique = a->r.unique_ctx;
	const uint32_t flags = (unique ? KSSDHIP_UNIQUE : 0) | (a->r.ignoreconflict ? KSSDHIP_IGNORECONFLICT : 0);
	if (!a->r.fmt && a->r.ctxcut >= 1 && R > 0) {
		/* detail output: the pass gate of command_ani.h:585-606 on the
		 * device, only the passing pairs come back */
		kssdhip_select_opts_t so;
		memset(&so, 0, sizeof so);
		so.ctxcut = (uint32_t)a->r.ctxcut;
		so.afcut = a->r.afcut;
		so.anicut = a->r.anicut;
		so.unique_rank = unique;
		so.value.metric = 3;
		so.value.values_ani = 1;
		so.value.drfold = a->r.drfold;
		so.value.obj_bits = ref.obj_bits;
		so.value.ctx_bits = 44;
		so.value.exception = 1.0;
		size_t cap = 1 << 16;
		kssdhip_ani_row_t *rows = xmalloc(cap * sizeof *rows);
		for (int q0 = 0; q0 < Q;) {
			const int nb = select_block_queries(&qry, q0);
			size_t n = 0;
			tune_block(ix, qry.keys + qry.index[q0], qry.index + q0, nb, flags);
			int rc = kssdhip_select_block(ix, qry.keys + qry.index[q0], qry.index + q0, nb, flags, qry_ctx + q0,
						      ref_ctx, &so, rows, cap, &n);
			if (rc == KSSDHIP_ENOSPC) {
				cap = n;
				rows = xrealloc(rows, cap * sizeof *rows);
				rc = kssdhip_select_block(ix, qry.keys + qry.index[q0], qry.index + q0, nb, flags,
							  qry_ctx + q0, ref_ctx, &so, rows, cap, &n);
			}
			hip_check(rc, "GPU counting failed");
			stage("block counted and gated");
			emit_index_rows(fp, a, &ref, &qry, q0, nb, rows, n, qry_ctx, ref_ctx, &cv);
			stage("block ranked and formatted");
			q0 += nb;
		}
		free(rows);
		close_out(fp);
		kssdhip_index_close(ix);
		stage("index released");
		free(cv.qcounts);
		free(cv.bits);
		free(ref_ctx);
		free(qry_ctx);
		kssd_sketchdir_close(&ref);
		kssd_sketchdir_close(&qry);
		return 0;
	}
	const int blk = block_queries(Q, R, unique);
	const size_t cells = (size_t)blk * (size_t)(R ? R : 1);
	kssd_ctx_mut2_t *ctx = xmalloc(cells * sizeof *ctx);
	kssd_obj_section_t *obj = xmalloc(cells * sizeof *obj);
	kssd_ctx_mut2_t *uctx = unique ? xmalloc(cells * sizeof *uctx) : NULL;
	kssd_obj_section_t *uobj = unique ? xmalloc(cells * sizeof *uobj) : NULL;
	for (int q0 = 0, nb = 0; q0 < Q; q0 += nb) {
		nb = block_from(qry.index, q0, Q, blk, nrec);
		if (R > 0)
			tune_block(ix, qry.keys + qry.index[q0], qry.index + q0, nb, flags);
		if (R > 0)
			hip_check(kssdhip_count_block(ix, qry.keys + qry.index[q0], qry.index + q0, nb, q0, flags, ctx,
						      obj, uctx, uobj),
				  "GPU counting failed");
		stage("block counted");
		emit_rows(fp, a, PATH_INDEX, &ref, &qry, q0, nb, ctx, obj, uctx, uobj, qry_ctx, ref_ctx, &cv);
		stage("block ranked and formatted");
	}
	close_out(fp);
	kssdhip_index_close(ix);
	stage("index released");
	free(cv.qcounts);
	free(cv.bits);
	free(ctx);
	free(obj);
	free(uctx);
	free(uobj);
	free(ref_ctx);
	free(qry_ctx);
	kssd_sketchdir_close(&ref);
	kssd_sketchdir_close(&qry);
	return 0;
}

/*
 * Streamed paths (stream_ref_sketches_one_qraw_lookup, :2490-2666;
 * stream_ref_sketches_multi_qraw_sortedindex, :2792-3015) and the hash path
 * (comb_manysmall_..._filter_and_sort_survivors, :3888-4000): all pairs are
 * counted on the GPU, then ranked per query as the matching path does.
 */
static int run_rect_detail(const ani_args_t *a, detail_path_t path)
{
	kssd_sketchdir_t ref, qry;
	const bool cov = a->r.estimate_coverage && path != PATH_HASH;
	open_dir(a->refdir, KSSD_SK_KEYS | KSSD_SK_ANNO | KSSD_SK_META | (cov ? KSSD_SK_RECORDS : 0), &ref);
	open_dir(a->qrydir, KSSD_SK_KEYS | KSSD_SK_META, &qry);
	if (qry.stat.hash_id != ref.stat.hash_id)
		errx(EXIT_FAILURE, "hash_id mismatch between query and reference sketches");
	cov_inputs_t cv = {0};
	if (cov)
		load_coverage(a, &ref, &qry, false, &cv);
	stage("sketches read");
	const int R = ref.n, Q = qry.n;
	uint32_t *qry_ctx, *ref_ctx;
	if (path == PATH_HASH) {
		qry_ctx = entry_or_run_totals(&qry, qry.stat.conflict, false);
		ref_ctx = entry_or_run_totals(&ref, ref.stat.conflict, a->r.ignoreconflict);
	} else {
		/* the streamed paths count context runs whatever the conflict flag */
		qry_ctx = entry_or_run_totals(&qry, true, false);
		ref_ctx = entry_or_run_totals(&ref, true, a->r.ignoreconflict);
	}
	kssd_ctx_mut2_t *ctx = xcalloc((size_t)Q * R, sizeof *ctx);
	kssd_obj_section_t *obj = xcalloc((size_t)Q * R, sizeof *obj);
	/* the streamed paths skip multi-object reference contexts under
	 * --ignoreconflict whatever the conflict flag; without conflicts there
	 * are none, so the flag-gated rule of rect_counts is the same */
	rect_counts(&ref, &qry, a->r.ignoreconflict, a->p, ctx, obj);
	stage("pairs counted");
	FILE *fp = open_out(a);
	put_detail_header(fp, a);
	emit_rows(fp, a, path, &ref, &qry, 0, Q, ctx, obj, NULL, NULL, qry_ctx, ref_ctx, &cv);
	stage("ranked and formatted");
	close_out(fp);
	free(cv.qcounts);
	free(cv.bits);
	free(ctx);
	free(obj);
	free(qry_ctx);
	free(ref_ctx);
	kssd_sketchdir_close(&ref);
	kssd_sketchdir_close(&qry);
	return 0;
}

/* ------------------------------------------------------------ matrices */
/*
 * Self matrices: comb_sortedsketch64_self_matrix (:3314-3365, "%lf" cells)
 * and comb_sortedsketch64_indexed_self_full / _triangle (:1657-1842; the
 * full matrix prints fixed-6 cells).  Pair (q, r) with q > r is evaluated
 * with q as the query; the upper half mirrors it.
 */
/* Value of pair (q, r), q > r, of a self comparison from its counts. */
typedef double (*pair_value_fn)(const void *arg, int q, int r, const kssd_ctx_mut2_t *c, const kssd_obj_section_t *o);

/* Lower triangle of a self comparison: the pairs (q, r), r < q, counted on
 * the GPU in query blocks with KSSDHIP_LOWER (count_ctx_obj_frm_comb_sketch_
 * section_lower, command_ani.h:635-721) and valued on the host threads;
 * value (q, r) is at q (q - 1) / 2 + r. */
static double *self_lower(const kssd_sketchdir_t *s, int threads, bool ignoreconflict, pair_value_fn fn,
			  const void *arg)
{
	const int n = s->n;
	kssd_ctxgidobj_t *built = s->records ? NULL : build_records(s, threads);
	const kssd_ctxgidobj_t *rec = s->records ? s->records : built;
	kssdhip_index *ix = n ? gpu_index(rec, s->index[n], n, s->obj_bits) : NULL;
	free(built);
	stage("index on the GPU");
	const size_t lower_n = n > 1 ? (size_t)n * (size_t)(n - 1) / 2 : 0;
	double *lower = xmalloc(lower_n * sizeof *lower);
	const uint32_t flags = KSSDHIP_LOWER | (ignoreconflict ? KSSDHIP_IGNORECONFLICT : 0);
	const int blk = block_queries(n, n, false);
	kssd_ctx_mut2_t *ctx = xmalloc((size_t)blk * (n ? n : 1) * sizeof *ctx);
	kssd_obj_section_t *obj = xmalloc((size_t)blk * (n ? n : 1) * sizeof *obj);
	for (int q0 = 0, nb = 0; q0 < n; q0 += nb) {
		nb = block_from(s->index, q0, n, blk, (size_t)s->index[n]);
		tune_block(ix, s->keys + s->index[q0], s->index + q0, nb, flags);
		hip_check(kssdhip_count_block(ix, s->keys + s->index[q0], s->index + q0, nb, q0, flags, ctx, obj, NULL,
					      NULL),
			  "GPU counting failed");
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
		for (int i = 0; i < nb; ++i) {
			const int q = q0 + i;
			for (int r = 0; r < q; ++r)
				lower[(size_t)q * (q - 1) / 2 + r] =
					fn(arg, q, r, &ctx[(size_t)i * n + r], &obj[(size_t)i * n + r]);
		}
	}
	if (ix)
		kssdhip_index_close(ix);
	free(ctx);
	free(obj);
	return lower;
}

struct ani_cell {
	const kssd_report_opts_t *o;
	const uint32_t *qc, *rc;
	const kssd_infile_meta_t *meta;
};

static double ani_cell_value(const void *arg, int q, int r, const kssd_ctx_mut2_t *c, const kssd_obj_section_t *b)
{
	const struct ani_cell *a = arg;
	return kssd_matrix_value(a->o, c->num_ctx, b->diff_obj, b->diff_obj_section, c->num_mut2_ctx, a->qc[q], a->rc[r],
				 a->meta ? &a->meta[q] : NULL, a->meta ? &a->meta[r] : NULL);
}

static int run_self_matrix(const ani_args_t *a, bool fixed6, bool index_from_file)
{
	kssd_sketchdir_t s;
	open_dir(a->qrydir, KSSD_SK_KEYS | KSSD_SK_META | (index_from_file ? KSSD_SK_RECORDS : 0), &s);
	stage("sketch read");
	const int n = s.n;
	const kssd_report_opts_t *o = &a->r;
	uint32_t *qc = entry_or_run_totals(&s, s.stat.conflict, false);
	uint32_t *rc = o->ignoreconflict ? entry_or_run_totals(&s, s.stat.conflict, true) : qc;
	const struct ani_cell cell = {o, qc, rc, s.meta};
	double *lower = self_lower(&s, a->p, o->ignoreconflict, ani_cell_value, &cell);
	stage("pairs counted and evaluated");
	FILE *fp = open_out(a);
	kssd_buf_t b = {0};
	const double diag = kssd_matrix_diagonal(o);
	if (o->fmt == 1) {
		put_header_names(fp, &s);
		for (int q = 0; q < n; ++q) {
			kssd_buf_printf(&b, "%s", s.names[q]);
			for (int r = 0; r < n; ++r) {
				const double v = q == r ? diag
						 : q > r ? lower[(size_t)q * (q - 1) / 2 + r]
							 : lower[(size_t)r * (r - 1) / 2 + q];
				if (fixed6) {
					kssd_buf_printf(&b, "\t");
					kssd_put_fixed6(&b, kssd_scale_fixed6(v));
				} else {
					kssd_buf_printf(&b, "\t%lf", v);
				}
			}
			kssd_buf_printf(&b, "\n");
			put_buf(fp, &b);
		}
	} else {
		for (int q = 0; q < n; ++q) {
			kssd_buf_printf(&b, "%s", s.names[q]);
			for (int r = 0; r < q; ++r)
				kssd_buf_printf(&b, "\t%lf", lower[(size_t)q * (q - 1) / 2 + r]);
			if (o->d)
				kssd_buf_printf(&b, "\t%lf", diag);
			kssd_buf_printf(&b, "\n");
			put_buf(fp, &b);
		}
	}
	kssd_buf_free(&b);
	stage("formatted");
	close_out(fp);
	free(lower);
	if (rc != qc)
		free(rc);
	free(qc);
	kssd_sketchdir_close(&s);
	return 0;
}

/* comb_sortedsketch64Xcomb_sortedsketch64, matrix branch (:3210-3245) */
static int run_rect_matrix(const ani_args_t *a)
{
	kssd_sketchdir_t ref, qry;
	open_dir(a->refdir, KSSD_SK_KEYS | KSSD_SK_META, &ref);
	open_dir(a->qrydir, KSSD_SK_KEYS | KSSD_SK_META, &qry);
	stage("sketches read");
	const int R = ref.n, Q = qry.n;
	const kssd_report_opts_t *o = &a->r;
	uint32_t *qc = entry_or_run_totals(&qry, qry.stat.conflict, false);
	uint32_t *rc = entry_or_run_totals(&ref, ref.stat.conflict, o->ignoreconflict);
	kssd_ctx_mut2_t *ctx = xcalloc((size_t)Q * R, sizeof *ctx);
	kssd_obj_section_t *obj = xcalloc((size_t)Q * R, sizeof *obj);
	rect_counts(&ref, &qry, o->ignoreconflict, a->p, ctx, obj);
	stage("pairs counted");
	FILE *fp = open_out(a);
	put_header_names(fp, &ref);
	kssd_buf_t *bufs = xcalloc((size_t)Q, sizeof *bufs);
#pragma omp parallel for schedule(dynamic, 1) num_threads(a->p)
	for (int q = 0; q < Q; ++q) {
		kssd_buf_printf(&bufs[q], "%s", qry.names[q]);
		for (int r = 0; r < R; ++r) {
			const kssd_ctx_mut2_t *c = &ctx[(size_t)q * R + r];
			const kssd_obj_section_t *b = &obj[(size_t)q * R + r];
			kssd_buf_printf(&bufs[q], "\t%lf",
					kssd_matrix_value(o, c->num_ctx, b->diff_obj, b->diff_obj_section, c->num_mut2_ctx,
							  qc[q], rc[r], qry.meta ? &qry.meta[q] : NULL,
							  ref.meta ? &ref.meta[r] : NULL));
		}
		kssd_buf_printf(&bufs[q], "\n");
	}
	for (int q = 0; q < Q; ++q) {
		put_buf(fp, &bufs[q]);
		kssd_buf_free(&bufs[q]);
	}
	free(bufs);
	stage("formatted");
	close_out(fp);
	free(ctx);
	free(obj);
	free(qc);
	free(rc);
	kssd_sketchdir_close(&ref);
	kssd_sketchdir_close(&qry);
	return 0;
}

/* ------------------------------------------------------------ dispatch */
static bool matrix_should_use_index(int ref_n, int qry_n, int direct_threshold)
{
	/* ani_matrix_should_use_index, command_ani_wrapper.c:699-706 */
	if (ref_n < 1 || qry_n < 1)
		return false;
	if (direct_threshold == 0)
		return true;
	return ref_n > direct_threshold || qry_n > direct_threshold;
}

static int matrix_direct_threshold(void)
{
	const char *name = "KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD";
	if (!getenv(name) || !getenv(name)[0])
		name = "KSSD3A_ANI_M1_DIRECT_THRESHOLD";
	return (int)env_int(name, 1000, 0, "non-negative integer");
}

static void ensure_index(const char *dir, int threads, int device)
{
	if (kssd_file_in_dir(dir, KSSD_SORTED_INDEX_FILE))
		return;
	write_index_for(dir, threads, false, device);
	if (!kssd_file_in_dir(dir, KSSD_SORTED_INDEX_FILE))
		errx(EXIT_FAILURE, "failed to create reference index '%s/%s'", dir, KSSD_SORTED_INDEX_FILE);
}

static void __attribute__((noreturn)) force_missing(const char *dir)
{
	errx(EXIT_FAILURE, "KSSD3A_FORCE_REF_INDEX requires reference index '%s/%s'", dir, KSSD_SORTED_INDEX_FILE);
}

/* run_sketch_ani, command_ani_wrapper.c:1082-1230 */
static int run_ani(ani_args_t *a)
{
	kssd_report_opts_t *o = &a->r;
	if (o->unique_ctx && !kssd_file_in_dir(a->refdir, KSSD_SORTED_INDEX_FILE))
		errx(EXIT_FAILURE, "unique-context options require an existing unfiltered reference inverted index; "
				   "run sketch -i on the reference sketch");
	if (o->unassembled && !a->afcut_set)
		o->afcut = 0.2f;
	kssd_sketchdir_t q;
	open_dir(a->qrydir, 0, &q);
	const int qn = q.n;
	o->drfold = q.stat.drfold;
	o->ctx_bits = q.ctx_bits;
	o->obj_bits = q.obj_bits;
	kssd_sketchdir_close(&q);
	const bool force = env_flag("KSSD3A_FORCE_REF_INDEX") || o->unique_ctx;
	const int direct_n = o->fmt ? matrix_direct_threshold() : 0;

	if (!a->refdir[0] || (o->fmt && same_file(a->refdir, a->qrydir))) {
		if (o->fmt == 0)
			errx(EXIT_FAILURE, "one-sketch ANI detail output is not supported; use -m1 full matrix or -m2 triangle");
		const bool indexed = force || matrix_should_use_index(qn, qn, direct_n);
		const bool have = kssd_file_in_dir(a->qrydir, KSSD_SORTED_INDEX_FILE);
		if (o->fmt == 1 && indexed) {
			if (force && !have)
				force_missing(a->qrydir);
			ensure_index(a->qrydir, a->p, a->device);
			return run_self_matrix(a, true, true);
		}
		if (o->fmt == 2 && indexed) {
			if (force && !have)
				force_missing(a->qrydir);
			return run_self_matrix(a, false, have);
		}
		return run_self_matrix(a, false, false);
	}
	if (o->fmt == 2)
		errx(EXIT_FAILURE, "ANI triangle output requires one sketch or identical -r/-q sketches");
	kssd_sketchdir_t r;
	open_dir(a->refdir, 0, &r);
	const int rn = r.n;
	kssd_sketchdir_close(&r);
	const bool small_query_stream = !force && o->fmt == 0 && rn > 0 && qn < rn;
	if (small_query_stream && qn == 1)
		return run_rect_detail(a, PATH_STREAM);
	if (o->fmt == 0 && qn > 1) {
		bool use_query_index =
			small_query_stream && qn <= env_int("KSSD3A_QRAW_MULTI_THRESHOLD", 20, 1, "positive integer");
		const char *mode = getenv("KSSD3A_QRAW_MULTI");
		if (o->unassembled && !force && mode && mode[0]) {
			if (strcmp(mode, "sortedindex") == 0)
				use_query_index = true;
			else if (strcmp(mode, "legacy") == 0 || strcmp(mode, "refindex") == 0)
				use_query_index = false;
			else
				errx(EXIT_FAILURE, "invalid KSSD3A_QRAW_MULTI='%s' (use sortedindex, refindex, or legacy)",
				     mode);
		}
		if (use_query_index)
			return run_rect_detail(a, PATH_MULTISTREAM);
	}
	const bool ref_index = kssd_file_in_dir(a->refdir, KSSD_SORTED_INDEX_FILE);
	if (o->fmt == 1) {
		if (force) {
			if (!ref_index)
				force_missing(a->refdir);
			return run_index_path(a);
		}
		if (matrix_should_use_index(rn, qn, direct_n)) {
			ensure_index(a->refdir, a->p, a->device);
			return run_index_path(a);
		}
	} else if (ref_index) {
		return run_index_path(a);
	}
	const long auto_n = env_int("KSSD3A_AUTO_REF_INDEX_THRESHOLD", 100, 0, "non-negative integer; 0 disables");
	if (!o->v && !o->unassembled && o->fmt == 0 && auto_n > 0 && rn > auto_n) {
		write_index_for(a->refdir, a->p, false, a->device);
		return run_index_path(a);
	}
	if (force)
		force_missing(a->refdir);
	if (o->fmt)
		return run_rect_matrix(a);
	return run_rect_detail(a, PATH_HASH);
}

static int cmd_ani(int argc, char **argv)
{
	ani_args_t a;
	parse_ani(argc, argv, &a);
	if (kssd_report_init() != 0)
		errx(EXIT_FAILURE, "calibrated ANI model tables are damaged");
	hip_start(a.device);
	return run_ani(&a);
}

/* ------------------------------------------------------------ matrix */
/*
 * The reference's `kssd3a matrix` for its distance-matrix reports
 * (command_matrix_wrapper.c:109-135 defaults, :399-470 checks; compute_triangle
 * and compute_matrix, command_matrix.c:1293-1348): one sketch gives the
 * indexed lower-triangle path (matrix_write_indexed_self_matrix, :439-619)
 * printed as a triangle or a full matrix; -r/-q give the dense report
 * (matrix_compute_dense, :258-315).  --format edges (--sparse) gives the
 * sparse edge report (matrix_compute_edges, :992-1014, and the
 * vote-nominated indexed scan, :1140-1180 with pairwise_graph.c:1148-1316),
 * with metric expressions; --format clusters gives the connected components
 * of those edges with their best representative, and --format dedup-plan the
 * keep/remove plan of a greedy or full-linkage grouping of them, with the
 * kept samples' distance matrix (matrix_write_component_report, :1055-1118;
 * pairwise_graph.c:792-973).
 */
typedef struct {
	char refdir[PATH_MAX], qrydir[PATH_MAX], outf[PATH_MAX], edge_outf[PATH_MAX];
	char keep_outf[PATH_MAX], remove_outf[PATH_MAX], keep_matrix_outf[PATH_MAX], keep_idmap_outf[PATH_MAX];
	int metric;
	kssd_metric_expr_t expr;
	int format; /* 0 auto, 1 full, 2 triangle, 3 edges, 4 clusters, 5 dedup-plan */
	int full_linkage, keep_phylip;
	double e, diag, cut, max_afcut;
	bool d, cut_set, max_afcut_set, strategy_set;
	uint32_t ctxcut, max_ctx_freq, min_votes, sample_step;
	int p, device;
} matrix_args_t;

struct matrix_cell {
	int metric;
	double e;
	const uint32_t *tot;
	int drfold;
	unsigned obj_bits, ctx_bits;
};

static double matrix_cell_value(const void *arg, int q, int r, const kssd_ctx_mut2_t *c, const kssd_obj_section_t *b)
{
	const struct matrix_cell *m = arg;
	return kssd_matrix_context_value(m->metric, m->e, c->num_ctx, b->diff_obj, b->diff_obj_section, c->num_mut2_ctx,
					 m->tot[q], m->tot[r], m->drfold, m->obj_bits, m->ctx_bits);
}

static void parse_matrix(int argc, char **argv, matrix_args_t *a)
{
	enum { M_FORMAT = 300, M_DIAGV, M_DEVICE, M_HELP, M_NOT_OFFERED, M_SPARSE, M_CUT, M_MAXAF, M_CTXCUT,
	       M_MAXFREQ, M_MINVOTES, M_STEP, M_EDGEOUT, M_KEEPOUT, M_REMOVEOUT, M_KMOUT, M_KMFMT, M_KMIDMAP,
	       M_STRATEGY };
	static const struct option longopts[] = {
		{"ref", required_argument, NULL, 'r'},
		{"query", required_argument, NULL, 'q'},
		{"metric", required_argument, NULL, 'm'},
		{"format", required_argument, NULL, M_FORMAT},
		{"exception", required_argument, NULL, 'e'},
		{"diagonal", no_argument, NULL, 'd'},
		{"diagonal-value", required_argument, NULL, M_DIAGV},
		{"outfile", required_argument, NULL, 'o'},
		{"threads", required_argument, NULL, 'p'},
		{"device", required_argument, NULL, M_DEVICE},
		{"help", no_argument, NULL, M_HELP},
		{"sparse", no_argument, NULL, M_SPARSE},
		{"cut", required_argument, NULL, M_CUT},
		{"max-afcut", required_argument, NULL, M_MAXAF},
		{"ctxcut", required_argument, NULL, M_CTXCUT},
		{"index-max-ctx-freq", required_argument, NULL, M_MAXFREQ},
		{"index-min-votes", required_argument, NULL, M_MINVOTES},
		{"index-sample-step", required_argument, NULL, M_STEP},
		{"edge-out", required_argument, NULL, M_EDGEOUT},
		{"control", required_argument, NULL, M_CUT},
		{"keep-out", required_argument, NULL, M_KEEPOUT},
		{"remove-out", required_argument, NULL, M_REMOVEOUT},
		{"keep-matrix-out", required_argument, NULL, M_KMOUT},
		{"keep-matrix-format", required_argument, NULL, M_KMFMT},
		{"keep-matrix-idmap", required_argument, NULL, M_KMIDMAP},
		{"dedup-strategy", required_argument, NULL, M_STRATEGY},
		{"query-sketch-list", required_argument, NULL, M_NOT_OFFERED},
		{"matrix-format", required_argument, NULL, M_NOT_OFFERED},
		{"glist", required_argument, NULL, M_NOT_OFFERED},
		{"ani", no_argument, NULL, M_NOT_OFFERED},
		{NULL, 0, NULL, 0},
	};
	memset(a, 0, sizeof *a);
	a->metric = KSSD_MM_NAIVE;
	a->expr = kssd_metric_expr_single(KSSD_MM_NAIVE);
	a->e = 1.0;
	a->p = omp_get_max_threads();
	/* matrix_opt defaults, command_matrix_wrapper.c:119-130 */
	a->max_ctx_freq = 256;
	a->min_votes = 1;
	a->sample_step = 1;
	optind = 1;
	int c;
	while ((c = getopt_long(argc, argv, "r:q:m:e:do:p:g:c:", longopts, NULL)) != -1) {
		switch (c) {
		case 'r':
			copy_path(a->refdir, sizeof a->refdir, "-r/--ref", optarg);
			break;
		case 'q':
			copy_path(a->qrydir, sizeof a->qrydir, "-q/--query", optarg);
			break;
		case 'm':
			if (kssd_metric_expr_parse(optarg, &a->expr) != 0)
				usage_error("-m/--metric must be one metric such as ctx-naive, p_dist, mash, or aaf, "
					    "or an expression like 'ctx-naive&aaf' or 'ctx-moe|mash'");
			a->metric = a->expr.count > 1 ? -2 : a->expr.metrics[0];
			break;
		case M_FORMAT:
			if (!strcmp(optarg, "full") || !strcmp(optarg, "matrix"))
				a->format = 1;
			else if (!strcmp(optarg, "triangle") || !strcmp(optarg, "tri"))
				a->format = 2;
			else if (!strcmp(optarg, "edges") || !strcmp(optarg, "edge") || !strcmp(optarg, "sparse"))
				a->format = 3;
			else if (!strcmp(optarg, "clusters") || !strcmp(optarg, "cluster") || !strcmp(optarg, "components"))
				a->format = 4;
			else if (!strcmp(optarg, "dedup-plan") || !strcmp(optarg, "dedup"))
				a->format = 5;
			else
				usage_error("--format must be one of full, triangle, edges/sparse, clusters, or dedup-plan");
			break;
		case 'e':
			a->e = parse_double("-e/--exception", optarg, 0.0, 1e300);
			break;
		case 'd':
			a->d = true;
			break;
		case M_DIAGV:
			a->diag = parse_double("--diagonal-value", optarg, 0.0, 1e300);
			a->d = true;
			break;
		case 'o':
			copy_path(a->outf, sizeof a->outf, "-o/--outfile", optarg);
			break;
		case 'p':
			a->p = (int)parse_int("-p/--threads", optarg, 1, 65536);
			break;
		case M_DEVICE:
			a->device = (int)parse_int("--device", optarg, 0, 1024);
			break;
		case M_SPARSE:
			a->format = 3;
			break;
		case 'c':
		case M_CUT:
			a->cut = parse_double("--cut", optarg, 0.0, 1e300);
			a->cut_set = true;
			break;
		case M_MAXAF:
			a->max_afcut = parse_double("--max-afcut", optarg, 0.0, 1.0);
			a->max_afcut_set = true;
			break;
		case M_CTXCUT:
			a->ctxcut = (uint32_t)parse_int("--ctxcut", optarg, 0, INT32_MAX);
			break;
		case M_MAXFREQ:
			a->max_ctx_freq = (uint32_t)parse_int("--index-max-ctx-freq", optarg, 0, INT32_MAX);
			break;
		case M_MINVOTES:
			a->min_votes = (uint32_t)parse_int("--index-min-votes", optarg, 1, INT32_MAX);
			break;
		case M_STEP:
			a->sample_step = (uint32_t)parse_int("--index-sample-step", optarg, 1, INT32_MAX);
			break;
		case M_EDGEOUT:
			copy_path(a->edge_outf, sizeof a->edge_outf, "--edge-out", optarg);
			break;
		case M_KEEPOUT:
			copy_path(a->keep_outf, sizeof a->keep_outf, "--keep-out", optarg);
			break;
		case M_REMOVEOUT:
			copy_path(a->remove_outf, sizeof a->remove_outf, "--remove-out", optarg);
			break;
		case M_KMOUT:
			copy_path(a->keep_matrix_outf, sizeof a->keep_matrix_outf, "--keep-matrix-out", optarg);
			break;
		case M_KMIDMAP:
			copy_path(a->keep_idmap_outf, sizeof a->keep_idmap_outf, "--keep-matrix-idmap", optarg);
			break;
		case M_KMFMT: /* parse_keep_matrix_format, command_matrix_wrapper.c:202-212 */
			if (!strcmp(optarg, "tsv") || !strcmp(optarg, "table"))
				a->keep_phylip = 0;
			else if (!strcmp(optarg, "phylip") || !strcmp(optarg, "phy") || !strcmp(optarg, "fillip"))
				a->keep_phylip = 1;
			else
				usage_error("--keep-matrix-format must be one of tsv or phylip");
			break;
		case M_STRATEGY: /* pairwise_dedup_strategy_from_string, pairwise_graph.c:118-141 */
			a->strategy_set = true;
			if (!strcmp(optarg, "greedy") || !strcmp(optarg, "rep") || !strcmp(optarg, "representative") ||
			    !strcmp(optarg, "rep-centric") || !strcmp(optarg, "representative-centric"))
				a->full_linkage = 0;
			else if (!strcmp(optarg, "full") || !strcmp(optarg, "full-linkage") ||
				 !strcmp(optarg, "full_linkage") || !strcmp(optarg, "complete") ||
				 !strcmp(optarg, "complete-linkage") || !strcmp(optarg, "complete_linkage") ||
				 !strcmp(optarg, "clique"))
				a->full_linkage = 1;
			else
				usage_error("--dedup-strategy must be one of greedy, full-linkage, complete-linkage, or clique");
			break;
		case M_HELP:
			printf("Usage: %s matrix [-m METRIC] [--format full|triangle] [-e E] [-d | --diagonal-value V] "
			       "[-o FILE] (SKETCH | -q SKETCH | -r REF -q QRY)\n"
			       "       %s matrix --format edges|clusters|dedup-plan --cut D [-m METRIC|EXPR] [--ctxcut N]\n"
			       "             [--max-afcut F] [--index-max-ctx-freq N] [--index-min-votes N]\n"
			       "             [--index-sample-step N] [--edge-out FILE] [-o FILE] (SKETCH | -r REF -q QRY)\n"
			       "       dedup-plan: [--dedup-strategy greedy|full-linkage] [--keep-out F] [--remove-out F]\n"
			       "             [--keep-matrix-out F [--keep-matrix-format tsv|phylip [--keep-matrix-idmap F]]]\n"
			       "METRIC: ctx-naive (default), ctx-moe, p_dist, mash, aaf; EXPR: metrics joined by & or |\n"
			       "(edges, clusters and dedup plans).\n",
			       PROG, PROG);
			exit(0);
		case 'g':
		case M_NOT_OFFERED:
			usage_error("%s is not offered by this front end (distance matrices only)", argv[optind - 1]);
		default:
			fprintf(stderr, "Try `%s %s --help' for more information.\n", PROG, g_cmd);
			exit(EX_USAGE);
		}
	}
	const int npos = argc - optind;
	if (!a->qrydir[0] && npos == 1)
		copy_path(a->qrydir, sizeof a->qrydir, "sketch", argv[optind]);
	else if (npos > 0)
		usage_error("matrix accepts at most one positional sketch argument; use -r/-q for rectangular reports");
	if (!a->qrydir[0])
		usage_error("missing sketch input; provide one sketch argument, -q/--query, or --query-sketch-list");
	if (a->refdir[0] && a->format == 2)
		usage_error("--format triangle is only valid for one sketch");
	if (a->metric == -2 && a->format != 3 && a->format != 4 && a->format != 5)
		usage_error("combined --metric expressions are only valid for --sparse/--format edges, clusters, or "
			    "dedup-plan");
	if (a->format >= 3 && !a->cut_set)
		usage_error("--sparse/--format edges, clusters, or dedup-plan requires --cut");
	if ((a->format == 4 || a->format == 5) && a->refdir[0])
		usage_error("--format clusters/dedup-plan requires one sketch, not -r/-q");
	/* the rest of the end-of-arguments checks, command_matrix_wrapper.c:433-509 */
	if ((a->keep_outf[0] || a->remove_outf[0] || a->keep_matrix_outf[0] || a->keep_idmap_outf[0]) && a->format != 5)
		usage_error("--keep-out/--remove-out/--keep-matrix-out/--keep-matrix-idmap require --format dedup-plan");
	if (a->strategy_set && a->format != 5)
		usage_error("--dedup-strategy requires --format dedup-plan");
	if (a->keep_phylip && !a->keep_matrix_outf[0])
		usage_error("--keep-matrix-format requires --keep-matrix-out");
	if (a->keep_idmap_outf[0] && !a->keep_matrix_outf[0])
		usage_error("--keep-matrix-idmap requires --keep-matrix-out");
	if (a->keep_idmap_outf[0] && !a->keep_phylip)
		usage_error("--keep-matrix-idmap requires --keep-matrix-format phylip");
	static const struct {
		size_t a, b;
		const char *msg;
	} differ[] = {
		{offsetof(matrix_args_t, keep_outf), offsetof(matrix_args_t, remove_outf),
		 "--keep-out and --remove-out must be different files"},
		{offsetof(matrix_args_t, keep_outf), offsetof(matrix_args_t, outf), "--keep-out must differ from --outfile"},
		{offsetof(matrix_args_t, remove_outf), offsetof(matrix_args_t, outf),
		 "--remove-out must differ from --outfile"},
		{offsetof(matrix_args_t, keep_matrix_outf), offsetof(matrix_args_t, outf),
		 "--keep-matrix-out must differ from --outfile"},
		{offsetof(matrix_args_t, edge_outf), offsetof(matrix_args_t, outf), "--edge-out must differ from --outfile"},
		{offsetof(matrix_args_t, keep_outf), offsetof(matrix_args_t, edge_outf),
		 "--keep-out must differ from --edge-out"},
		{offsetof(matrix_args_t, remove_outf), offsetof(matrix_args_t, edge_outf),
		 "--remove-out must differ from --edge-out"},
		{offsetof(matrix_args_t, keep_matrix_outf), offsetof(matrix_args_t, edge_outf),
		 "--keep-matrix-out must differ from --edge-out"},
		{offsetof(matrix_args_t, keep_matrix_outf), offsetof(matrix_args_t, keep_outf),
		 "--keep-matrix-out must differ from --keep-out"},
		{offsetof(matrix_args_t, keep_matrix_outf), offsetof(matrix_args_t, remove_outf),
		 "--keep-matrix-out must differ from --remove-out"},
		{offsetof(matrix_args_t, keep_idmap_outf), offsetof(matrix_args_t, outf),
		 "--keep-matrix-idmap must differ from --outfile"},
		{offsetof(matrix_args_t, keep_idmap_outf), offsetof(matrix_args_t, edge_outf),
		 "--keep-matrix-idmap must differ from --edge-out"},
		{offsetof(matrix_args_t, keep_idmap_outf), offsetof(matrix_args_t, keep_outf),
		 "--keep-matrix-idmap must differ from --keep-out"},
		{offsetof(matrix_args_t, keep_idmap_outf), offsetof(matrix_args_t, remove_outf),
		 "--keep-matrix-idmap must differ from --remove-out"},
		{offsetof(matrix_args_t, keep_idmap_outf), offsetof(matrix_args_t, keep_matrix_outf),
		 "--keep-matrix-idmap must differ from --keep-matrix-out"},
	};
	for (size_t i = 0; i < sizeof differ / sizeof differ[0]; ++i) {
		const char *x = (const char *)a + differ[i].a, *y = (const char *)a + differ[i].b;
		if (x[0] && y[0] && strcmp(x, y) == 0)
			usage_error("%s", differ[i].msg);
	}
}

static int matrix_self(const matrix_args_t *a)
{
	kssd_sketchdir_t s;
	const bool have = kssd_file_in_dir(a->qrydir, KSSD_SORTED_INDEX_FILE);
	open_dir(a->qrydir, KSSD_SK_KEYS | (have ? KSSD_SK_RECORDS : 0), &s);
	stage("sketch read");
	const int n = s.n;
	uint32_t *tot = entry_or_run_totals(&s, s.stat.conflict, false);
	const struct matrix_cell cell = {a->metric, a->e, tot, s.stat.drfold, s.obj_bits, s.ctx_bits};
	double *lower = self_lower(&s, a->p, false, matrix_cell_value, &cell);
	stage("pairs counted and evaluated");
	FILE *fp = a->outf[0] ? fopen(a->outf, "w") : stdout;
	if (!fp)
		err(errno, "%s", a->outf);
	kssd_buf_t b = {0};
	if (a->format == 1) {
		put_header_names(fp, &s);
		const uint32_t dg = kssd_matrix_scale_fixed6(a->diag);
		for (int q = 0; q < n; ++q) {
			kssd_buf_printf(&b, "%s", s.names[q]);
			for (int r = 0; r < n; ++r) {
				const uint32_t v = q == r ? dg
						 : kssd_matrix_scale_fixed6(q > r ? lower[(size_t)q * (q - 1) / 2 + r]
										  : lower[(size_t)r * (r - 1) / 2 + q]);
				kssd_buf_printf(&b, "\t");
				kssd_matrix_put_scaled6(&b, v);
			}
			kssd_buf_printf(&b, "\n");
			put_buf(fp, &b);
		}
	} else {
		for (int q = 0; q < n; ++q) {
			kssd_buf_printf(&b, "%s", s.names[q]);
			for (int r = 0; r < q; ++r) {
				kssd_buf_printf(&b, "\t");
				kssd_matrix_put_fixed6(&b, lower[(size_t)q * (q - 1) / 2 + r]);
			}
			if (a->d) {
				kssd_buf_printf(&b, "\t");
				kssd_matrix_put_fixed6(&b, a->diag);
			}
			kssd_buf_printf(&b, "\n");
			put_buf(fp, &b);
		}
	}
	kssd_buf_free(&b);
	close_out(fp);
	stage("formatted");
	free(lower);
	free(tot);
	kssd_sketchdir_close(&s);
	return 0;
}

static int matrix_rect(const matrix_args_t *a)
{
	kssd_sketchdir_t ref, qry;
	open_dir(a->refdir, KSSD_SK_KEYS, &ref);
	open_dir(a->qrydir, KSSD_SK_KEYS, &qry);
	/* pairwise_check_compatible, pairwise_graph.c:262-273 */
	if (ref.stat.hash_id != qry.stat.hash_id)
		errx(EXIT_FAILURE, "pairwise_check_compatible(): ref hash_id %u != qry %u", ref.stat.hash_id,
		     qry.stat.hash_id);
	stage("sketches read");
	const int R = ref.n, Q = qry.n;
	const bool same = strcmp(a->refdir, a->qrydir) == 0;
	uint32_t *qc = entry_or_run_totals(&qry, qry.stat.conflict, false);
	uint32_t *rc = entry_or_run_totals(&ref, ref.stat.conflict, false);
	kssd_ctx_mut2_t *ctx = xcalloc((size_t)Q * R, sizeof *ctx);
	kssd_obj_section_t *obj = xcalloc((size_t)Q * R, sizeof *obj);
	rect_counts(&ref, &qry, false, a->p, ctx, obj);
	stage("pairs counted");
	FILE *fp = a->outf[0] ? fopen(a->outf, "w") : stdout;
	if (!fp)
		err(errno, "%s", a->outf);
	put_header_names(fp, &ref);
	kssd_buf_t *bufs = xcalloc((size_t)Q, sizeof *bufs);
#pragma omp parallel for schedule(dynamic, 1) num_threads(a->p)
	for (int q = 0; q < Q; ++q) {
		kssd_buf_printf(&bufs[q], "%s", qry.names[q]);
		const size_t qlen = (size_t)(qry.index[q + 1] - qry.index[q]);
		for (int r = 0; r < R; ++r) {
			double v = a->diag;
			if (!(same && q == r)) {
				const kssd_ctx_mut2_t *c = &ctx[(size_t)q * R + r];
				const kssd_obj_section_t *o = &obj[(size_t)q * R + r];
				v = kssd_pairwise_value(a->metric, a->e, qlen, (size_t)(ref.index[r + 1] - ref.index[r]),
							c->num_ctx, o->diff_obj, o->diff_obj_section, c->num_mut2_ctx, qc[q],
							rc[r], ref.stat.drfold, ref.obj_bits, ref.ctx_bits);
			}
			kssd_buf_printf(&bufs[q], "\t%lf", v);
		}
		kssd_buf_printf(&bufs[q], "\n");
	}
	for (int q = 0; q < Q; ++q) {
		put_buf(fp, &bufs[q]);
		kssd_buf_free(&bufs[q]);
	}
	free(bufs);
	close_out(fp);
	stage("formatted");
	free(ctx);
	free(obj);
	free(qc);
	free(rc);
	kssd_sketchdir_close(&ref);
	kssd_sketchdir_close(&qry);
	return 0;
}

/* ---- sparse edge reports ---- */
typedef struct {
	int q, r;
	uint64_t order; /* indexed scan: the context where r got its first vote from q */
	kssd_pair_eval_t ev;
} edge_t;

typedef struct {
	edge_t *v;
	size_t n, cap;
} edge_list_t;

static void edge_push(edge_list_t *l, const edge_t *e)
{
	if (l->n == l->cap) {
		l->cap = l->cap ? 2 * l->cap : 16;
		l->v = realloc(l->v, l->cap * sizeof *l->v);
		if (!l->v)
			err(EXIT_FAILURE, "edges");
	}
	l->v[l->n++] = *e;
}

static int edge_by_order(const void *pa, const void *pb)
{
	const edge_t *a = pa, *b = pb;
	if (a->order != b->order)
		return a->order < b->order ? -1 : 1;
	return (a->r > b->r) - (a->r < b->r);
}

static int edge_by_ref(const void *pa, const void *pb)
{
	const edge_t *a = pa, *b = pb;
	if (a->r != b->r)
		return (a->r > b->r) - (a->r < b->r);
	return (a->q > b->q) - (a->q < b->q);
}

/* Number of index records of context c (records sorted by ctxgid). */
static size_t ctx_group_size(const kssd_ctxgidobj_t *rec, size_t n, uint64_t c)
{
	size_t lo = 0, hi = n;
	while (lo < hi) {
		const size_t mid = lo + (hi - lo) / 2;
		if ((rec[mid].ctxgid >> 20) < c)
			lo = mid + 1;
		else
			hi = mid;
	}
	size_t e = lo;
	while (e < n && (rec[e].ctxgid >> 20) == c)
		++e;
	return e - lo;
}

/* The context where genome r gets its first vote from query q in the
 * indexed scan (pairwise_graph.c:1200-1235 walks q's sampled keys in
 * order and, for each, the context's records in gid order): the first of
 * q's every step-th key whose context r carries and whose index group has
 * at most max_freq records (0: any). */
static uint64_t first_vote_ctx(const kssd_sketchdir_t *s, const kssd_ctxgidobj_t *rec, size_t nrec, int q, int r,
			       uint32_t max_freq, uint32_t step)
{
	const unsigned ob = s->obj_bits;
	const uint64_t *a = s->keys + s->index[q], *b = s->keys + s->index[r];
	const size_t na = (size_t)(s->index[q + 1] - s->index[q]), nb = (size_t)(s->index[r + 1] - s->index[r]);
	size_t j = 0;
	for (size_t i = 0; i < na; i += step) {
		const uint64_t c = a[i] >> ob;
		while (j < nb && (b[j] >> ob) < c)
			++j;
		if (j < nb && (b[j] >> ob) == c && (!max_freq || ctx_group_size(rec, nrec, c) <= max_freq))
			return c;
	}
	return UINT64_MAX;
}

static void put_edges(FILE *fp, const kssd_sketchdir_t *qs, const kssd_sketchdir_t *rs, const edge_t *v, size_t n,
		      const char *label)
{
	kssd_buf_t b = {0};
	for (size_t i = 0; i < n; ++i) {
		kssd_put_edge(&b, qs->names[v[i].q], rs->names[v[i].r], &v[i].ev, label);
		if (b.l > (1u << 20))
			put_buf(fp, &b);
	}
	put_buf(fp, &b);
	kssd_buf_free(&b);
}

/* One sketch: the vote-nominated indexed scan when the sketch is indexed
 * and keeps no conflicts (matrix_can_use_sorted_index, command_matrix.c:
 * 1045-1053; pairwise_indexed_self_scan, pairwise_graph.c:1148-1316),
 * else every pair (matrix_compute_edges, command_matrix.c:992-1014).  The
 * GPU counts the shared-context votes and the pair counts; rows come out in
 * query order, each row's edges in first-vote order (the reference's order
 * with one thread; with more its rows interleave). */
static edge_list_t *self_edges(const matrix_args_t *a, const kssd_sketchdir_t *s, const uint32_t *tot,
				bool indexed)
{
	const int n = s->n;
	kssd_ctxgidobj_t *built = s->records ? NULL : build_records(s, a->p);
	const kssd_ctxgidobj_t *rec = s->records ? s->records : built;
	const size_t nrec = s->index[n];
	kssdhip_index *ix = n ? gpu_index(rec, nrec, n, s->obj_bits) : NULL;
	edge_list_t *rows = xcalloc((size_t)(n ? n : 1), sizeof *rows);
	const int blk = block_queries(n, n, false);
	kssd_ctx_mut2_t *ctx = xmalloc((size_t)blk * (n ? n : 1) * sizeof *ctx);
	kssd_obj_section_t *obj = xmalloc((size_t)blk * (n ? n : 1) * sizeof *obj);
	uint32_t *votes = indexed ? xmalloc((size_t)blk * (n ? n : 1) * sizeof *votes) : NULL;
	for (int q0 = 0, nb = 0; q0 < n; q0 += nb) {
		nb = block_from(s->index, q0, n, blk, nrec);
		hip_check(kssdhip_count_block(ix, s->keys + s->index[q0], s->index + q0, nb, q0, KSSDHIP_LOWER, ctx, obj,
					      NULL, NULL),
			  "GPU counting failed");
		if (indexed)
			hip_check(kssdhip_count_votes(ix, s->keys + s->index[q0], s->index + q0, nb, q0, a->max_ctx_freq,
						      a->sample_step, votes),
				  "GPU vote counting failed");
#pragma omp parallel for schedule(dynamic, 1) num_threads(a->p)
		for (int i = 0; i < nb; ++i) {
			const int q = q0 + i;
			const size_t qlen = (size_t)(s->index[q + 1] - s->index[q]);
			for (int r = 0; r < q; ++r) {
				const size_t c = (size_t)i * n + r;
				if (indexed && (votes[c] == 0 || votes[c] < a->min_votes))
					continue;
				edge_t e = {q, r, 0, {0}};
				e.ev = kssd_pair_eval(&a->expr, qlen, (size_t)(s->index[r + 1] - s->index[r]), ctx[c].num_ctx,
						      obj[c].diff_obj, obj[c].diff_obj_section, ctx[c].num_mut2_ctx, tot[q],
						      tot[r], s->stat.drfold, s->obj_bits, s->ctx_bits);
				if (!kssd_pair_passes_edge(&e.ev, a->cut, a->ctxcut, a->max_afcut))
					continue;
				if (indexed)
					e.order = first_vote_ctx(s, rec, nrec, q, r, a->max_ctx_freq, a->sample_step);
				edge_push(&rows[q], &e);
			}
		}
	}
	stage("pairs counted and evaluated");
	if (ix)
		kssdhip_index_close(ix);
	free(votes);
	free(ctx);
	free(obj);
	free(built);
	return rows;
}

/* Every accepted edge of `rows` in the order the reference reports them:
 * the indexed scan's (query rows, first-vote order inside a row) or
 * matrix_compute_edges' (ref-major). */
static edge_list_t ordered_edges(edge_list_t *rows, int n, bool indexed)
{
	edge_list_t all = {0};
	for (int q = 0; q < n; ++q) {
		if (indexed)
			qsort(rows[q].v, rows[q].n, sizeof *rows[q].v, edge_by_order);
		for (size_t k = 0; k < rows[q].n; ++k)
			edge_push(&all, &rows[q].v[k]);
	}
	if (!indexed && all.n)
		qsort(all.v, all.n, sizeof *all.v, edge_by_ref);
	return all;
}

static FILE *open_or_stdout(const char *path)
{
	FILE *fp = path[0] ? fopen(path, "w") : stdout;
	if (!fp)
		err(errno, "%s", path);
	return fp;
}

static void write_edge_file(const char *path, const kssd_sketchdir_t *s, const edge_list_t *all, const char *label)
{
	FILE *fp = open_or_stdout(path);
	kssd_buf_t b = {0};
	kssd_put_edge_header(&b);
	put_buf(fp, &b);
	kssd_buf_free(&b);
	put_edges(fp, s, s, all->v, all->n, label);
	close_out(fp);
}

/* One sketch: the vote-nominated indexed scan when the sketch is indexed
 * and keeps no conflicts (matrix_can_use_sorted_index, command_matrix.c:
 * 1045-1053; pairwise_indexed_self_scan, pairwise_graph.c:1148-1316),
 * else every pair (matrix_compute_edges, command_matrix.c:992-1014).  The
 * GPU counts the shared-context votes and the pair counts; rows come out in
 * query order, each row's edges in first-vote order (the reference's order
 * with one thread; with more its rows interleave). */
static int matrix_edges_self(const matrix_args_t *a)
{
	kssd_sketchdir_t s;
	const bool have = kssd_file_in_dir(a->qrydir, KSSD_SORTED_INDEX_FILE);
	open_dir(a->qrydir, KSSD_SK_KEYS | (have ? KSSD_SK_RECORDS : 0), &s);
	stage("sketch read");
	const bool indexed = have && !s.stat.conflict;
	uint32_t *tot = entry_or_run_totals(&s, s.stat.conflict, false);
	edge_list_t *rows = self_edges(a, &s, tot, indexed);
	edge_list_t all = ordered_edges(rows, s.n, indexed);
	write_edge_file(a->outf, &s, &all, a->expr.label);
	stage("formatted");
	for (int q = 0; q < s.n; ++q)
		free(rows[q].v);
	free(rows);
	free(all.v);
	free(tot);
	kssd_sketchdir_close(&s);
	return 0;
}

static int uf_find(int *parent, int x)
{
	while (parent[x] != x) {
		parent[x] = parent[parent[x]];
		x = parent[x];
	}
	return x;
}

/* Samples best first (pairwise_quality_order, pairwise_graph.c:719-790): a
 * top-down merge sort whose merge takes the left head when it is better by
 * kssd_quality_compare, ties going to the smaller index. */
static bool quality_better(const kssd_sketchdir_t *s, int a, int b)
{
	const int c = kssd_quality_compare(s->meta, s->index, a, b);
	return c > 0 || (c == 0 && a < b);
}

static void quality_sort(int *idx, int *tmp, int lo, int hi, const kssd_sketchdir_t *s)
{
	if (hi - lo <= 1)
		return;
	const int mid = lo + (hi - lo) / 2;
	quality_sort(idx, tmp, lo, mid, s);
	quality_sort(idx, tmp, mid, hi, s);
	int i = lo, j = mid, k = lo;
	while (i < mid && j < hi)
		tmp[k++] = quality_better(s, idx[i], idx[j]) ? idx[i++] : idx[j++];
	while (i < mid)
		tmp[k++] = idx[i++];
	while (j < hi)
		tmp[k++] = idx[j++];
	memcpy(idx + lo, tmp + lo, (size_t)(hi - lo) * sizeof *idx);
}

static int cmp_u64_key(const void *a, const void *b)
{
	const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return (x > y) - (x < y);
}

/* Dedup groups.  Greedy (pairwise_greedy_dedup_from_edges, pairwise_graph.c:
 * 792-860): in quality order each sample not yet removed keeps itself and
 * removes every neighbour not yet removed.  Full linkage (pairwise_complete_
 * linkage_dedup_from_edges, :895-973): each unassigned sample in quality order
 * opens a group and takes, again in quality order, every unassigned sample
 * linked to all members so far.  group[i] is the representative of i. */
static void dedup_groups(int n, const edge_list_t *all, const int *order, bool full_linkage, int *group,
			 int *gsize)
{
	for (int i = 0; i < n; ++i) {
		group[i] = -1;
		gsize[i] = 0;
	}
	if (!full_linkage) {
		size_t *off = xcalloc((size_t)n + 1, sizeof *off);
		for (size_t k = 0; k < all->n; ++k) {
			off[all->v[k].q + 1]++;
			off[all->v[k].r + 1]++;
		}
		for (int i = 0; i < n; ++i)
			off[i + 1] += off[i];
		int *adj = xmalloc((off[n] ? off[n] : 1) * sizeof *adj);
		size_t *cur = xmalloc(((size_t)n + 1) * sizeof *cur);
		memcpy(cur, off, ((size_t)n + 1) * sizeof *cur);
		for (size_t k = 0; k < all->n; ++k) {
			adj[cur[all->v[k].q]++] = all->v[k].r;
			adj[cur[all->v[k].r]++] = all->v[k].q;
		}
		for (int oi = 0; oi < n; ++oi) {
			const int rep = order[oi];
			if (group[rep] >= 0)
				continue;
			group[rep] = rep;
			gsize[rep] = 1;
			for (size_t t = off[rep]; t < off[rep + 1]; ++t) {
				const int x = adj[t];
				if (group[x] >= 0)
					continue;
				group[x] = rep;
				gsize[rep]++;
			}
		}
		free(adj);
		free(cur);
		free(off);
		return;
	}
	/* edges as sorted a * n + b keys, both directions */
	uint64_t *key = xmalloc((2 * all->n + 1) * sizeof *key);
	for (size_t k = 0; k < all->n; ++k) {
		key[2 * k] = (uint64_t)all->v[k].q * (uint64_t)n + (uint64_t)all->v[k].r;
		key[2 * k + 1] = (uint64_t)all->v[k].r * (uint64_t)n + (uint64_t)all->v[k].q;
	}
	qsort(key, 2 * all->n, sizeof *key, cmp_u64_key);
	int *members = xmalloc((size_t)(n ? n : 1) * sizeof *members);
	for (int oi = 0; oi < n; ++oi) {
		const int rep = order[oi];
		if (group[rep] >= 0)
			continue;
		group[rep] = rep;
		gsize[rep] = 1;
		members[0] = rep;
		int nm = 1;
		for (int ci = 0; ci < n; ++ci) {
			const int x = order[ci];
			if (group[x] >= 0)
				continue;
			bool linked = true;
			for (int m = 0; m < nm && linked; ++m) {
				const uint64_t want = (uint64_t)x * (uint64_t)n + (uint64_t)members[m];
				linked = bsearch(&want, key, 2 * all->n, sizeof *key, cmp_u64_key) != NULL;
			}
			if (!linked)
				continue;
			group[x] = rep;
			gsize[rep]++;
			members[nm++] = x;
		}
	}
	free(members);
	free(key);
}

/* matrix_meta_valid (command_matrix.c:155-162) */
static bool meta_valid(const kssd_sketchdir_t *s, int i)
{
	return s->meta && s->meta[i].meta_fmt_version == 1 && s->meta[i].total_length_bp > 0;
}

/* Why rep was kept over sample (matrix_rep_rule, command_matrix.c:182-231). */
static const char *rep_rule(const kssd_sketchdir_t *s, int sample, int rep)
{
	if (sample == rep)
		return "self";
	const bool sv = meta_valid(s, sample), rv = meta_valid(s, rep);
	if (sv != rv)
		return rv ? "valid_metadata" : "sample_valid_metadata";
	if (sv) {
		const kssd_infile_meta_t *a = &s->meta[sample], *b = &s->meta[rep];
		const bool af = a->infile_fmt == 1, bf = b->infile_fmt == 1;
		if (af != bf)
			return bf ? "fasta_input" : "sample_fasta_input";
		if (a->asm_level != b->asm_level)
			return b->asm_level > a->asm_level ? "higher_asm_level" : "sample_higher_asm_level";
		if (a->total_length_bp != b->total_length_bp)
			return b->total_length_bp > a->total_length_bp ? "longer_total_length" : "sample_longer_total_length";
		if (a->record_count != b->record_count) {
			if (a->record_count == 0)
				return "sample_missing_record_count";
			if (b->record_count == 0)
				return "rep_missing_record_count";
			return b->record_count < a->record_count ? "fewer_records" : "sample_fewer_records";
		}
		if (a->median_length_bp != b->median_length_bp)
			return b->median_length_bp > a->median_length_bp ? "longer_median_length"
									: "sample_longer_median_length";
		if (a->length_cv != b->length_cv)
			return b->length_cv < a->length_cv ? "lower_length_cv" : "sample_lower_length_cv";
	}
	const uint64_t se = s->index[sample + 1] - s->index[sample], re = s->index[rep + 1] - s->index[rep];
	if (se != re)
		return re > se ? "larger_sketch_entries" : "sample_larger_sketch_entries";
	return rep < sample ? "earlier_sample_order" : "sample_earlier_sample_order";
}

/* matrix_print_meta_field (command_matrix.c:233-244) */
static void put_meta_fields(kssd_buf_t *b, const kssd_sketchdir_t *s, int i)
{
	if (!meta_valid(s, i)) {
		kssd_buf_printf(b, "\tNA\tNA\tNA\tNA\tNA\tNA");
		return;
	}
	const kssd_infile_meta_t *m = &s->meta[i];
	kssd_buf_printf(b, "\t%s\t%.6g\t%llu\t%u\t%u\t%.6g",
			m->infile_fmt == 1 ? "FASTA" : (m->infile_fmt == 2 ? "FASTQ" : "NA"), (double)m->asm_level,
			(unsigned long long)m->total_length_bp, m->record_count, m->median_length_bp,
			(double)m->length_cv);
}

/* The kept samples' dense distance matrix (matrix_write_keep_matrix and its
 * indexed form, command_matrix.c:785-960): rows and columns in sample order,
 * each row counted on the GPU against every sample, cells the metric
 * expression's value (the diagonal value on the diagonal); TSV with "%lf"
 * cells or PHYLIP with S%09d ids and " %.12g" cells, and the id map. */
static void write_keep_matrix(const matrix_args_t *a, const kssd_sketchdir_t *s, const kssd_ctxgidobj_t *rec,
			      const uint32_t *tot, const int *kept, int nk)
{
	if (a->keep_idmap_outf[0]) {
		FILE *fp = open_or_stdout(a->keep_idmap_outf);
		fprintf(fp, "id\tsample\n");
		for (int k = 0; k < nk; ++k)
			fprintf(fp, "S%09d\t%s\n", k + 1, s->names[kept[k]]);
		close_out(fp);
	}
	FILE *fp = open_or_stdout(a->keep_matrix_outf);
	kssd_buf_t b = {0};
	if (a->keep_phylip) {
		kssd_buf_printf(&b, "%d\n", nk);
	} else {
		for (int k = 0; k < nk; ++k)
			kssd_buf_printf(&b, "\t%s", s->names[kept[k]]);
		kssd_buf_printf(&b, "\n");
	}
	put_buf(fp, &b);
	const int n = s->n;
	if (nk > 0) {
		kssdhip_index *ix = gpu_index(rec, s->index[n], n, s->obj_bits);
		const int blk = block_queries(nk, n, false);
		kssd_ctx_mut2_t *ctx = xmalloc((size_t)blk * n * sizeof *ctx);
		kssd_obj_section_t *obj = xmalloc((size_t)blk * n * sizeof *obj);
		kssd_buf_t *rows = xcalloc((size_t)blk, sizeof *rows);
		uint64_t *skidx = xmalloc(((size_t)blk + 1) * sizeof *skidx);
		for (int k0 = 0; k0 < nk; k0 += blk) {
			const int nb = nk - k0 < blk ? nk - k0 : blk;
			/* the block's kept samples gathered into one contiguous run of keys */
			skidx[0] = 0;
			for (int i = 0; i < nb; ++i)
				skidx[i + 1] = skidx[i] + (s->index[kept[k0 + i] + 1] - s->index[kept[k0 + i]]);
			uint64_t *keys = xmalloc((size_t)(skidx[nb] ? skidx[nb] : 1) * sizeof *keys);
			for (int i = 0; i < nb; ++i)
				memcpy(keys + skidx[i], s->keys + s->index[kept[k0 + i]],
				       (size_t)(skidx[i + 1] - skidx[i]) * sizeof *keys);
			hip_check(kssdhip_count_block(ix, keys, skidx, nb, 0, 0, ctx, obj, NULL, NULL),
				  "GPU counting failed");
			free(keys);
#pragma omp parallel for schedule(dynamic, 1) num_threads(a->p)
			for (int i = 0; i < nb; ++i) {
				const int q = kept[k0 + i];
				if (a->keep_phylip)
					kssd_buf_printf(&rows[i], "S%09d", k0 + i + 1); /* "%-10s" of a 10-character id */
				else
					kssd_buf_printf(&rows[i], "%s", s->names[q]);
				for (int k = 0; k < nk; ++k) {
					const int r = kept[k];
					const size_t c = (size_t)i * n + r;
					const double v = q == r ? a->diag
							: kssd_matrix_context_expr_value(&a->expr, a->e, ctx[c].num_ctx, obj[c].diff_obj,
											 obj[c].diff_obj_section, ctx[c].num_mut2_ctx,
											 tot[q], tot[r], s->stat.drfold, s->obj_bits,
											 s->ctx_bits);
					kssd_buf_printf(&rows[i], a->keep_phylip ? " %.12g" : "\t%lf", v);
				}
				kssd_buf_printf(&rows[i], "\n");
			}
			for (int i = 0; i < nb; ++i) {
				put_buf(fp, &rows[i]);
				kssd_buf_free(&rows[i]);
			}
		}
		free(skidx);
		free(rows);
		free(ctx);
		free(obj);
		kssdhip_index_close(ix);
	}
	kssd_buf_free(&b);
	close_out(fp);
}

/* Clusters (format 4): the connected components of the accepted edges
 * (pairwise_build_components, pairwise_graph.c:975-1030, or the indexed
 * scan's unions), each with its best member by kssd_quality_compare
 * (pairwise_component_result_finalize, :682-712).  Dedup plans (format 5):
 * the greedy or full-linkage groups of the same edges, one keep/remove row
 * per sample with the rule that chose its representative, --keep-out /
 * --remove-out name lists and the kept samples' matrix.  Cluster ids number
 * the groups in order of their first member (matrix_write_component_report,
 * command_matrix.c:1055-1118).  --edge-out: the accepted edges, as the scan
 * (indexed) or the pair loop (i < j, queried as i against j) reports them. */
static int matrix_components(const matrix_args_t *a)
{
	const bool dedup = a->format == 5;
	kssd_sketchdir_t s;
	const bool have = kssd_file_in_dir(a->qrydir, KSSD_SORTED_INDEX_FILE);
	open_dir(a->qrydir, KSSD_SK_KEYS | KSSD_SK_META | (have ? KSSD_SK_RECORDS : 0), &s);
	stage("sketch read");
	const int n = s.n;
	const bool indexed = have && !s.stat.conflict;
	uint32_t *tot = entry_or_run_totals(&s, s.stat.conflict, false);
	matrix_args_t eff = *a;
	if (dedup && !a->max_afcut_set)
		eff.max_afcut = 0.8; /* compute_triangle, command_matrix.c:1314-1316 */
	edge_list_t *rows = self_edges(&eff, &s, tot, indexed);
	edge_list_t all = ordered_edges(rows, n, indexed);
	int *group = xmalloc((size_t)(n ? n : 1) * sizeof *group);
	int *size = xcalloc((size_t)(n ? n : 1), sizeof *size);
	int *rep = xmalloc((size_t)(n ? n : 1) * sizeof *rep);
	int *cid = xcalloc((size_t)(n ? n : 1), sizeof *cid);
	if (dedup) {
		int *order = xmalloc((size_t)(n ? n : 1) * sizeof *order);
		int *tmp = xmalloc((size_t)(n ? n : 1) * sizeof *tmp);
		for (int i = 0; i < n; ++i)
			order[i] = i;
		quality_sort(order, tmp, 0, n, &s);
		dedup_groups(n, &all, order, a->full_linkage, group, size);
		for (int i = 0; i < n; ++i)
			rep[i] = group[i];
		free(order);
		free(tmp);
	} else {
		for (int i = 0; i < n; ++i) {
			group[i] = i;
			rep[i] = -1;
		}
		for (size_t k = 0; k < all.n; ++k) {
			const int x = uf_find(group, all.v[k].q), y = uf_find(group, all.v[k].r);
			if (x != y)
				group[x > y ? x : y] = x < y ? x : y;
		}
		for (int i = 0; i < n; ++i)
			group[i] = uf_find(group, i);
		for (int i = 0; i < n; ++i)
			++size[group[i]];
		for (int i = 0; i < n; ++i)
			if (rep[group[i]] < 0 || kssd_quality_compare(s.meta, s.index, i, rep[group[i]]) > 0)
				rep[group[i]] = i;
	}
	int next = 1;
	for (int i = 0; i < n; ++i)
		if (!cid[group[i]])
			cid[group[i]] = next++;
	if (a->edge_outf[0]) {
		if (!indexed) /* the pair loop queries the smaller index against the larger */
			for (size_t k = 0; k < all.n; ++k) {
				edge_t *e = &all.v[k];
				const int i = e->r, j = e->q;
				e->ev = kssd_pair_eval(&a->expr, (size_t)(s.index[i + 1] - s.index[i]),
						       (size_t)(s.index[j + 1] - s.index[j]), e->ev.xny, e->ev.dobj, e->ev.dsec,
						       e->ev.mut2, tot[i], tot[j], s.stat.drfold, s.obj_bits, s.ctx_bits);
				e->q = i;
				e->r = j;
			}
		write_edge_file(a->edge_outf, &s, &all, a->expr.label);
	}
	FILE *fp = open_or_stdout(a->outf);
	FILE *keep_fp = dedup && a->keep_outf[0] ? open_or_stdout(a->keep_outf) : NULL;
	FILE *remove_fp = dedup && a->remove_outf[0] ? open_or_stdout(a->remove_outf) : NULL;
	kssd_buf_t b = {0};
	if (dedup)
		kssd_buf_printf(&b, "sample\tcomponent_id\taction\trepresentative\treason\trep_rule\tcomponent_size\t"
				    "sample_sketch_entries\trep_sketch_entries\tsample_fmt\tsample_asm_level\t"
				    "sample_total_length_bp\tsample_record_count\tsample_median_length_bp\t"
				    "sample_length_cv\trep_fmt\trep_asm_level\trep_total_length_bp\trep_record_count\t"
				    "rep_median_length_bp\trep_length_cv\n");
	else
		kssd_buf_printf(&b, "sample\tcluster_id\tcomponent_size\trepresentative\n");
	int *kept = xmalloc((size_t)(n ? n : 1) * sizeof *kept);
	int nk = 0;
	for (int i = 0; i < n; ++i) {
		const int root = group[i], r = rep[root];
		if (dedup) {
			const bool keep = i == r;
			const char *reason = size[root] == 1 ? "singleton" : (keep ? "best_representative" : "duplicate");
			if (keep)
				kept[nk++] = i;
			if (keep && keep_fp)
				fprintf(keep_fp, "%s\n", s.names[i]);
			if (!keep && remove_fp)
				fprintf(remove_fp, "%s\n", s.names[i]);
			kssd_buf_printf(&b, "%s\tcluster_%06d\t%s\t%s\t%s\t%s\t%d\t%llu\t%llu", s.names[i], cid[root],
					keep ? "keep" : "remove", s.names[r], reason, rep_rule(&s, i, r), size[root],
					(unsigned long long)(s.index[i + 1] - s.index[i]),
					(unsigned long long)(s.index[r + 1] - s.index[r]));
			put_meta_fields(&b, &s, i);
			put_meta_fields(&b, &s, r);
			kssd_buf_printf(&b, "\n");
		} else {
			kssd_buf_printf(&b, "%s\tcluster_%06d\t%d\t%s\n", s.names[i], cid[root], size[root], s.names[r]);
		}
		if (b.l > (1u << 20))
			put_buf(fp, &b);
	}
	put_buf(fp, &b);
	kssd_buf_free(&b);
	close_out(fp);
	if (keep_fp)
		close_out(keep_fp);
	if (remove_fp)
		close_out(remove_fp);
	if (dedup && a->keep_matrix_outf[0]) {
		kssd_ctxgidobj_t *built = s.records || !nk ? NULL : build_records(&s, a->p);
		write_keep_matrix(a, &s, s.records ? s.records : built, tot, kept, nk);
		free(built);
	}
	stage("formatted");
	for (int q = 0; q < n; ++q)
		free(rows[q].v);
	free(rows);
	free(all.v);
	free(kept);
	free(group);
	free(size);
	free(rep);
	free(cid);
	free(tot);
	kssd_sketchdir_close(&s);
	return 0;
}

/* -r/-q: every (ref, query) pair, refs outer (matrix_compute_edges,
 * command_matrix.c:992-1014); with the same sketch on both sides only
 * queries after the ref. */
static int matrix_edges_rect(const matrix_args_t *a)
{
	kssd_sketchdir_t ref, qry;
	open_dir(a->refdir, KSSD_SK_KEYS, &ref);
	open_dir(a->qrydir, KSSD_SK_KEYS, &qry);
	if (ref.stat.hash_id != qry.stat.hash_id)
		errx(EXIT_FAILURE, "pairwise_check_compatible(): ref hash_id %u != qry %u", ref.stat.hash_id,
		     qry.stat.hash_id);
	stage("sketches read");
	const int R = ref.n, Q = qry.n;
	const bool same = strcmp(a->refdir, a->qrydir) == 0;
	uint32_t *qc = entry_or_run_totals(&qry, qry.stat.conflict, false);
	uint32_t *rc = entry_or_run_totals(&ref, ref.stat.conflict, false);
	kssd_ctx_mut2_t *ctx = xcalloc((size_t)Q * R + 1, sizeof *ctx);
	kssd_obj_section_t *obj = xcalloc((size_t)Q * R + 1, sizeof *obj);
	rect_counts(&ref, &qry, false, a->p, ctx, obj);
	stage("pairs counted");
	edge_list_t *rows = xcalloc((size_t)(R ? R : 1), sizeof *rows);
#pragma omp parallel for schedule(dynamic, 1) num_threads(a->p)
	for (int r = 0; r < R; ++r) {
		const size_t rlen = (size_t)(ref.index[r + 1] - ref.index[r]);
		for (int q = same ? r + 1 : 0; q < Q; ++q) {
			const size_t c = (size_t)q * R + r;
			edge_t e = {q, r, 0, {0}};
			e.ev = kssd_pair_eval(&a->expr, (size_t)(qry.index[q + 1] - qry.index[q]), rlen, ctx[c].num_ctx,
					      obj[c].diff_obj, obj[c].diff_obj_section, ctx[c].num_mut2_ctx, qc[q], rc[r],
					      ref.stat.drfold, ref.obj_bits, ref.ctx_bits);
			if (kssd_pair_passes_edge(&e.ev, a->cut, a->ctxcut, a->max_afcut))
				edge_push(&rows[r], &e);
		}
	}
	FILE *fp = a->outf[0] ? fopen(a->outf, "w") : stdout;
	if (!fp)
		err(errno, "%s", a->outf);
	kssd_buf_t b = {0};
	kssd_put_edge_header(&b);
	put_buf(fp, &b);
	kssd_buf_free(&b);
	for (int r = 0; r < R; ++r) {
		put_edges(fp, &qry, &ref, rows[r].v, rows[r].n, a->expr.label);
		free(rows[r].v);
	}
	close_out(fp);
	stage("formatted");
	free(rows);
	free(ctx);
	free(obj);
	free(qc);
	free(rc);
	kssd_sketchdir_close(&ref);
	kssd_sketchdir_close(&qry);
	return 0;
}

static int cmd_matrix(int argc, char **argv)
{
	matrix_args_t a;
	parse_matrix(argc, argv, &a);
	hip_start(a.device);
	if (a.format == 3)
		return a.refdir[0] ? matrix_edges_rect(&a) : matrix_edges_self(&a);
	if (a.format == 4 || a.format == 5)
		return matrix_components(&a);
	if (a.refdir[0]) {
		if (a.format == 0)
			a.format = 1;
		return matrix_rect(&a);
	}
	if (a.format == 0)
		a.format = 2;
	return matrix_self(&a);
}

/* ------------------------------------------------------------ sketch */
/* Input paths of `sketch`: positional files, then the lines of -l lists
 * (the order the reference keeps them in, global_basic.c:240-420). */
static char **sketch_inputs(int npos, char **pos, const char *list, int *n_out)
{
	int cap = 16, n = 0;
	char **v = xmalloc((size_t)cap * sizeof *v);
	for (int i = 0; i < npos; ++i) {
		if (n == cap)
			v = realloc(v, (size_t)(cap *= 2) * sizeof *v);
		v[n++] = strdup(pos[i]);
	}
	if (list) {
		FILE *fp = fopen(list, "r");
		if (!fp)
			err(errno, "%s", list);
		char *line = NULL;
		size_t m = 0;
		ssize_t l;
		while ((l = getline(&line, &m, fp)) >= 0) {
			while (l && (line[l - 1] == '\n' || line[l - 1] == '\r' || line[l - 1] == ' '))
				line[--l] = '\0';
			if (!l)
				continue;
			if (n == cap)
				v = realloc(v, (size_t)(cap *= 2) * sizeof *v);
			v[n++] = strdup(line);
		}
		free(line);
		fclose(fp);
	}
	*n_out = n;
	return v;
}

typedef struct {
	int drfold, threads;
	bool conflict, meta, abundance, asone;
	const char *sample_name;
	kssd_count_opts_t counts;
} sketch_opts_t;

/* One sample's keys: collected, counted when a count option asks for it,
 * filtered by count, then cleared of conflicting contexts
 * (finalize_filter_conflict_and_write, command_sketch.c:5194-5214; the
 * per-file loop of sketch_many_files_in_parallel, :4455-4510). */
typedef struct {
	kssd_u64vec_t keys;
	uint32_t *counts; /* kept with -A */
	kssd_count_filter_t filter;
} sample_keys_t;

static void finish_sample(sample_keys_t *sm, const sketch_opts_t *o)
{
	if (!kssd_count_opts_need_counts(&o->counts)) {
		kssd_finalize_keys(&sm->keys, o->conflict);
		return;
	}
	if (kssd_count_keys(&sm->keys, &sm->counts) != 0)
		errx(EXIT_FAILURE, "out of memory counting k-mers");
	sm->filter = kssd_apply_count_filters(sm->keys.a, sm->counts, &sm->keys.n, &o->counts);
	if (!o->conflict)
		kssd_drop_conflicts(sm->keys.a, sm->counts, &sm->keys.n);
	if (!o->abundance) {
		free(sm->counts);
		sm->counts = NULL;
	}
}

static void check_input(const char *path)
{
	struct stat st;
	if (stat(path, &st) != 0 || !S_ISREG(st.st_mode))
		errx(EXIT_FAILURE, "%s: not a readable file (directories are not scanned here; list files with -l)", path);
	if (!kssd_infile_fmt_from_path(path))
		errx(EXIT_FAILURE, "%s: not a FASTA/FASTQ name (.fasta .fna .fas .fa .fq .fastq, optionally .gz)", path);
	if (strlen(path) >= KSSD_PATHLEN)
		errx(EXIT_FAILURE, "%s: path longer than %d bytes", path, KSSD_PATHLEN - 1);
}

/* --asone: every input pooled into one sample (sketch_few_files_with_
 * intrafile_parallel, command_sketch.c:5435-5497): the format shared by all
 * inputs (else unknown, flagged mixed), FASTA record lengths of all inputs
 * in order, the first path (or --sample-name) as the name. */
static void sketch_asone(char **files, int n, const sketch_opts_t *o, sample_keys_t *sm, kssd_infile_meta_t *mt)
{
	int bad = -1;
	if (kssd_sketch_pooled(files, n, o->drfold, o->threads, o->conflict, &o->counts, &sm->keys,
			       o->abundance ? &sm->counts : NULL, &sm->filter, &bad) != 0) {
		if (bad < 0)
			errx(EXIT_FAILURE, "out of memory sketching the pooled inputs");
		err(EXIT_FAILURE, "%s", files[bad]);
	}
	if (!o->meta)
		return;
	int fmt = kssd_infile_fmt_from_path(files[0]);
	unsigned flags = 0x10u; /* KSSD_INFILE_FLAG_MIXED_FORMAT */
	for (int i = 0; i < n; ++i) {
		flags |= kssd_infile_flags_from_path(files[i]);
		if (kssd_infile_fmt_from_path(files[i]) != fmt)
			fmt = 0;
	}
	if (fmt)
		flags &= ~0x10u;
	uint64_t *len = NULL;
	size_t nl = 0, cap = 0;
	if (fmt == 1)
		for (int i = 0; i < n; ++i)
			if (kssd_fasta_record_lengths(files[i], &len, &nl, &cap) != 0)
				err(EXIT_FAILURE, "%s", files[i]);
	kssd_infile_meta_from_lengths(len, nl, fmt, 1 /* KSSD_CREATE_ASONE */, flags, mt);
	free(len);
}

static int sketch_files(const char *outdir, char **files, int n, const sketch_opts_t *o)
{
	for (int i = 0; i < n; ++i)
		check_input(files[i]);
	const int ns = o->asone ? 1 : n;
	sample_keys_t *sm = xcalloc((size_t)ns, sizeof *sm);
	kssd_infile_meta_t *mt = xcalloc((size_t)ns, sizeof *mt);
	const bool report = o->counts.percentile > 0.0 || o->counts.reads_qc;
	if (o->asone) {
		sketch_asone(files, n, o, &sm[0], &mt[0]);
		if (report)
			kssd_print_count_filter(stderr, &o->counts, "as-one sketch", &sm[0].filter, 0);
	} else {
		int bad = -1;
		kssd_u64vec_t *raw = xcalloc((size_t)n, sizeof *raw);
		kssd_u64vec_t *lens = o->meta ? xcalloc((size_t)n, sizeof *lens) : NULL;
		int *lens_done = xcalloc((size_t)n, sizeof *lens_done);
		if (kssd_collect_files(files, n, o->drfold, o->threads, raw, &bad, lens, lens_done) != 0)
			err(EXIT_FAILURE, "%s", bad >= 0 ? files[bad] : "inputs");
#pragma omp parallel for schedule(dynamic, 1) num_threads(o->threads)
		for (int i = 0; i < n; ++i) {
			sm[i].keys = raw[i];
			if (o->meta) {
				/* record lengths from the sketching pass when it had them */
				const int fmt = kssd_infile_fmt_from_path(files[i]);
				if (lens_done[i])
					kssd_infile_meta_from_lengths(lens[i].a, fmt == 1 ? lens[i].n : 0, fmt, 0,
								      kssd_infile_flags_from_path(files[i]), &mt[i]);
				else if (kssd_infile_meta_for_path(files[i], &mt[i]) != 0) {
#pragma omp critical
					bad = i;
					continue;
				}
				kssd_u64vec_free(&lens[i]);
			}
			finish_sample(&sm[i], o);
		}
		free(raw);
		free(lens);
		free(lens_done);
		if (bad >= 0)
			err(EXIT_FAILURE, "%s", files[bad]);
		if (report)
			for (int i = 0; i < n; ++i)
				kssd_print_count_filter(stderr, &o->counts, files[i], &sm[i].filter, 1);
	}
	stage("inputs sketched");
	uint64_t *index = xmalloc(((size_t)ns + 1) * sizeof *index);
	index[0] = 0;
	for (int i = 0; i < ns; ++i)
		index[i + 1] = index[i] + sm[i].keys.n;
	uint64_t *keys = xmalloc((size_t)(index[ns] ? index[ns] : 1) * sizeof *keys);
	uint32_t *ab = o->abundance ? xmalloc((size_t)(index[ns] ? index[ns] : 1) * sizeof *ab) : NULL;
	kssd_qc_stat_t *qc = xcalloc((size_t)ns, sizeof *qc);
	for (int i = 0; i < ns; ++i) {
		if (sm[i].keys.n) {
			memcpy(keys + index[i], sm[i].keys.a, sm[i].keys.n * sizeof *keys);
			if (ab)
				memcpy(ab + index[i], sm[i].counts, sm[i].keys.n * sizeof *ab);
		}
		qc[i] = sm[i].filter.qc;
		kssd_u64vec_free(&sm[i].keys);
		free(sm[i].counts);
	}
	char (*names)[KSSD_PATHLEN] = xcalloc((size_t)ns, KSSD_PATHLEN);
	for (int i = 0; i < ns; ++i)
		strcpy(names[i], files[i]);
	if (o->sample_name)
		snprintf(names[0], KSSD_PATHLEN, "%s", o->sample_name); /* name_asone_sample */
	/* lcofiles.stat of the coden pattern (command_sketch.c:989-1005,
	 * command_sketch_wrapper.c:1186-1209) */
	kssd_stat_t st = {0};
	st.hash_id = kssd_sketching_id(11, 0, 0, o->drfold);
	st.koc = o->abundance;
	st.conflict = o->conflict;
	st.coden_len = 11;
	st.klen = 32;
	st.hclen = 0;
	st.holen = 0;
	st.drfold = o->drfold;
	st.infile_num = ns;
	const size_t nk = (size_t)index[ns];
	if (kssd_sketchdir_write(outdir, &st, names, keys, index) != 0 ||
	    (o->meta && kssd_write_infile_meta(outdir, mt, ns) != 0) ||
	    (ab && kssd_write_dir_file(outdir, KSSD_ABUND_FILE, ab, nk * sizeof *ab) != 0) ||
	    ((o->abundance || o->counts.reads_qc) &&
	     kssd_write_dir_file(outdir, KSSD_QC_FILE, qc, (size_t)ns * sizeof *qc) != 0))
		err(EXIT_FAILURE, "%s", outdir);
	stage("sketch written");
	free(sm);
	free(mt);
	free(index);
	free(keys);
	free(ab);
	free(qc);
	free(names);
	return 0;
}

static int cmd_sketch(int argc, char **argv)
{
	enum { S_CONFLICT = 400, S_NOMETA, S_HELP, S_NOT_OFFERED, S_HOST_INDEX, S_DEVICE, S_NPCT, S_NCAP,
	       S_READSQC, S_SAMPLE_NAME };
	static const struct option longopts[] = {
		{"idx", required_argument, NULL, 'i'},
		{"index", required_argument, NULL, 'i'},
		{"outdir", required_argument, NULL, 'o'},
		{"list", required_argument, NULL, 'l'},
		{"DimRdcFold", required_argument, NULL, 'f'},
		{"conflict", no_argument, NULL, S_CONFLICT},
		{"nocomputemeta", no_argument, NULL, S_NOMETA},
		{"use_coden_ctxobj", no_argument, NULL, 'T'},
		{"threads", required_argument, NULL, 'p'},
		{"host-index", no_argument, NULL, S_HOST_INDEX},
		{"device", required_argument, NULL, S_DEVICE},
		{"help", no_argument, NULL, S_HELP},
		{"abundance", no_argument, NULL, 'A'},
		{"asone", no_argument, NULL, 'a'},
		{"sample-name", required_argument, NULL, S_SAMPLE_NAME},
		{"LstKmerOcrs", required_argument, NULL, 'n'},
		{"npercentile", required_argument, NULL, S_NPCT},
		{"ncap", required_argument, NULL, S_NCAP},
		{"readsQC", no_argument, NULL, S_READSQC},
		{"splitmfa", no_argument, NULL, S_NOT_OFFERED},
		{"anno", no_argument, NULL, S_NOT_OFFERED},
		{"separate", no_argument, NULL, S_NOT_OFFERED},
		{"position", no_argument, NULL, S_NOT_OFFERED},
		{"pipecmd", required_argument, NULL, S_NOT_OFFERED},
		{NULL, 0, NULL, 0},
	};
	const char *dir = NULL, *outdir = NULL, *list = NULL;
	int threads = omp_get_max_threads(), device = 0;
	bool host_index = false;
	sketch_opts_t so = {.drfold = 8, .meta = true, .counts = {.min_count = 1}};
	optind = 1;
	int c;
	while ((c = getopt_long(argc, argv, "i:o:l:f:Tp:aAn:", longopts, NULL)) != -1) {
		switch (c) {
		case 'i':
			dir = optarg;
			break;
		case 'o':
			outdir = optarg;
			break;
		case 'l':
			list = optarg;
			break;
		case 'f':
			so.drfold = (int)parse_int("-f/--DimRdcFold", optarg, 0, 24);
			break;
		case 'T':
			break;
		case S_CONFLICT:
			so.conflict = true;
			break;
		case S_NOMETA:
			so.meta = false;
			break;
		case 'A':
			so.abundance = true;
			so.counts.record_qc = 1;
			break;
		case 'a':
			so.asone = true;
			break;
		case S_SAMPLE_NAME:
			for (const char *q = optarg; *q; ++q)
				if ((unsigned char)*q < 0x20 || *q == 0x7f)
					usage_error("--sample-name must not contain control characters");
			so.sample_name = optarg;
			break;
		case 'n':
			so.counts.min_count = (int)parse_int("-n/--LstKmerOcrs", optarg, 1, 65536);
			break;
		case S_NPCT:
			so.counts.percentile = parse_double("--npercentile", optarg, 0.0, 1.0);
			break;
		case S_NCAP:
			so.counts.cap = (int)parse_int("--ncap", optarg, 0, 65536);
			break;
		case S_READSQC:
			so.counts.reads_qc = 1; /* implies a lower cutoff of at least 2 */
			if (so.counts.min_count < 2)
				so.counts.min_count = 2;
			break;
		case 'p':
			threads = (int)parse_int("-p/--threads", optarg, 1, 65536);
			break;
		case S_HOST_INDEX:
			host_index = true;
			break;
		case S_DEVICE:
			device = (int)parse_int("--device", optarg, 0, 1024);
			break;
		case S_HELP:
			printf("Usage: %s sketch -o OUTDIR [-f DRFOLD] [--conflict] [--nocomputemeta] [-l LIST]\n"
			       "              [-A] [--asone [--sample-name NAME]] [-n MIN] [--npercentile P]\n"
			       "              [--ncap N] [--readsQC] FILE...\n"
			       "       %s sketch -i SKETCH_DIR [--device N | --host-index [-p THREADS]]\n"
			       "Sketch FASTA/FASTQ[.gz] files with the coden context-object pattern (one sample per\n"
			       "file, or all pooled with --asone; -A keeps per-k-mer counts in comblco.a), or build\n"
			       "SKETCH_DIR/%s and its .unique_bits (on the GPU; --host-index: on the host).\n",
			       PROG, PROG, KSSD_SORTED_INDEX_FILE);
			return 0;
		case S_NOT_OFFERED:
			usage_error("%s is not offered by this front end", argv[optind - 1]);
		default:
			fprintf(stderr, "Try `%s %s --help' for more information.\n", PROG, g_cmd);
			exit(EX_USAGE);
		}
	}
	if (so.sample_name && (!so.asone || dir))
		usage_error("--sample-name requires raw-input --asone mode without --splitmfa or --separate");
	if (dir) {
		if (outdir || optind != argc)
			usage_error("-i builds the index of an existing sketch; it takes no inputs");
		write_index_for(dir, threads, host_index, device);
		stage("index written");
		return 0;
	}
	if (!outdir)
		usage_error("give -o OUTDIR and input files, or -i SKETCH_DIR");
	int n = 0;
	char **files = sketch_inputs(argc - optind, argv + optind, list, &n);
	if (n == 0)
		usage_error("no input files");
	if (!so.asone && n >= (1 << 20))
		errx(EXIT_FAILURE, "%d inputs exceed the %d genomes a sketch can index", n, 1 << 20);
	so.threads = threads;
	const int rc = sketch_files(outdir, files, n, &so);
	for (int i = 0; i < n; ++i)
		free(files[i]);
	free(files);
	return rc;
}

int main(int argc, char **argv)
{
	g_t0 = now_s();
	g_timing = env_flag("KSSDHIP_TIMING");
	if (argc < 2 || strcmp(argv[1], "--help") == 0 || strcmp(argv[1], "-h") == 0) {
		printf("Usage: %s ani|matrix|sketch [options]\n", PROG);
		return argc < 2 ? EX_USAGE : 0;
	}
	g_cmd = argv[1];
	int rc;
	if (strcmp(g_cmd, "ani") == 0) {
		rc = cmd_ani(argc - 1, argv + 1);
	} else if (strcmp(g_cmd, "matrix") == 0) {
		rc = cmd_matrix(argc - 1, argv + 1);
	} else if (strcmp(g_cmd, "sketch") == 0) {
		rc = cmd_sketch(argc - 1, argv + 1);
	} else {
		fprintf(stderr, "%s: unknown command '%s' (ani, matrix, sketch)\n", PROG, g_cmd);
		return EX_USAGE;
	}
	hip_settle();
	return rc;
}
