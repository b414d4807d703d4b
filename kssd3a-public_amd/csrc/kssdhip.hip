// kssdhip -- sketch-comparison kernels for MI355X (gfx950), C ABI in
// include/kssdhip.h.
//
// Layout in HBM for a resident index of N records (sorted by ctxgid, obj):
//   ctxgid : u64[N]   ctx << 20 | gid          (8 B/record)
//   obj    : u32[N]   object bits              (4 B/record)
//   bucket : u32[2^k + 1]  bucket[t] = first record whose context has
//            top-k bits (of its actual bit width) >= t
// Dense per-pair counters are u64 cells laid out exactly like the
// reference's ctx_mut2_t / obj_section_t arrays (low word first), so a single
// 64-bit atomic add updates both fields of a cell.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <new>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_internal.h"

// ---------------------------------------------------------------- errors
static thread_local char g_err[1024];
static thread_local int g_device = 0;

int kssdhip_fail(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof g_err, fmt, ap);
	va_end(ap);
	return code;
}

extern "C" const char *kssdhip_last_error(void) { return g_err; }

#define HIP_TRY(expr)                                                                  \
	do {                                                                               \
		hipError_t e_ = (expr);                                                        \
		if (e_ != hipSuccess)                                                          \
			return kssdhip_fail(KSSDHIP_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
	} while (0)

// The first stream made on a device costs about 12 ms. kssdhip_init makes
// one per device ahead of time (once per process), and the first index
// opened on that device takes it over, so a caller that runs kssdhip_init
// on a thread of its own while it reads its inputs gets the cost off its
// critical path.
static std::mutex g_spare_mu;
static hipStream_t g_spare[64];
static bool g_spare_made[64];

static hipStream_t take_spare_stream(int device)
{
	std::lock_guard<std::mutex> lock(g_spare_mu);
	if (device < 0 || device >= 64)
		return nullptr;
	hipStream_t st = g_spare[device];
	g_spare[device] = nullptr;
	return st;
}

extern "C" int kssdhip_init(int device)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
		return kssdhip_fail(KSSDHIP_ENODEV, "no HIP device available");
	if (device < 0 || device >= n)
		return kssdhip_fail(KSSDHIP_EINVAL, "device %d out of range (0..%d)", device, n - 1);
	HIP_TRY(hipSetDevice(device));
	g_device = device;
	g_err[0] = '\0';
	int rc = kssd_noalloc_setup(device);
	if (rc)
		return rc;
	if (device < 64) {
		std::lock_guard<std::mutex> lock(g_spare_mu);
		if (!g_spare_made[device]) {
			g_spare_made[device] = true;
			if (hipStreamCreateWithFlags(&g_spare[device], hipStreamNonBlocking) != hipSuccess)
				g_spare[device] = nullptr; // index_open makes its own
		}
	}
	return 0;
}

// ---------------------------------------------------------------- device helpers
namespace kssd {

__device__ __forceinline__ int dna_popcount(uint32_t x) { return dna_popcount_dev(x); }

// out[i] = low 32 bits of in[i] (the shared-context count of a counter cell)
__global__ void k_low_words(const unsigned long long *__restrict__ in, uint64_t n, uint32_t *out)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
		out[i] = (uint32_t)in[i];
}

__global__ void k_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
	     i += (uint64_t)gridDim.x * blockDim.x)
		p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
	     i += (uint64_t)gridDim.x * blockDim.x)
		p[i] = v;
}

// 12-byte packed records -> SoA.  A record starts at a 4-byte aligned offset.
__global__ void k_unpack_records(const uint32_t *raw, uint64_t n, uint64_t *ctxgid, uint32_t *obj)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
	     i += (uint64_t)gridDim.x * blockDim.x) {
		const uint32_t *r = raw + 3 * i;
		ctxgid[i] = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
		obj[i] = r[2];
	}
}

// Flags records whose gid is out of range or that break the (ctxgid, obj)
// order, so no later kernel can index a counter row out of bounds.
__global__ void k_validate(const uint64_t *ctxgid, const uint32_t *obj, uint64_t n, uint32_t ref_gnum,
			   unsigned obj_bits, uint32_t *bad)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
	     i += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t c = ctxgid[i];
		bool ok = (c & GID_MASK) < ref_gnum && (obj_bits >= 32 || (obj[i] >> obj_bits) == 0);
		if (i) {
			const uint64_t p = ctxgid[i - 1];
			ok = ok && (p < c || (p == c && obj[i - 1] <= obj[i]));
		}
		if (!ok)
			atomicOr(bad, 1u);
	}
}

// bucket[t] = first record with (ctx >> shift) >= t, bucket[2^k] = N.
__global__ void k_buckets(const uint64_t *ctxgid, uint64_t n, int k, int shift, uint32_t *bucket)
{
	const uint64_t nb = 1ull << k;
	for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t <= nb;
	     t += (uint64_t)gridDim.x * blockDim.x) {
		if (t == nb) {
			bucket[t] = (uint32_t)n;
			continue;
		}
		uint64_t lo = 0, hi = n;
		while (lo < hi) {
			const uint64_t mid = lo + ((hi - lo) >> 1);
			if ((ctxgid[mid] >> (GID_NBITS + shift)) < t)
				lo = mid + 1;
			else
				hi = mid;
		}
		bucket[t] = (uint32_t)lo;
	}
}

// First record with context == qctx, or n.  qctx must be <= ctx_max.
__device__ __forceinline__ uint64_t lookup_ctx(const uint64_t *__restrict__ ctxgid,
					       const uint32_t *__restrict__ bucket, int shift,
					       uint64_t qctx)
{
	const uint32_t t = (uint32_t)(qctx >> shift);
	uint64_t lo = bucket[t], hi = bucket[t + 1];
	const uint64_t target = qctx << GID_NBITS;
	const uint64_t end = hi;
	while (lo < hi) {
		const uint64_t mid = lo + ((hi - lo) >> 1);
		if (ctxgid[mid] < target)
			lo = mid + 1;
		else
			hi = mid;
	}
	if (lo < end && (ctxgid[lo] >> GID_NBITS) == qctx)
		return lo;
	return ~0ull;
}

// Index-join counting, one thread per query key; the first key of each
// query context group does the work for the group (command_ani.h:506-575,
// :668-716).
__global__ __launch_bounds__(256) void k_join(JoinArgs a)
{
	const uint64_t objmask = a.obj_bits >= 64 ? ~0ull : ((1ull << a.obj_bits) - 1ull);
	const bool lower = (a.flags & KSSDHIP_LOWER) != 0;
	const bool unique = !lower && (a.flags & KSSDHIP_UNIQUE) != 0;
	const bool ignore = (a.flags & KSSDHIP_IGNORECONFLICT) != 0;
	for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < a.total;
	     e += (uint64_t)gridDim.x * blockDim.x) {
		const int q = owner_query(a.skidx, a.nq, e);
		const uint64_t qb = a.skidx[q], qe = a.skidx[q + 1];
		if (a.vote_step > 1 && (e - qb) % a.vote_step)
			continue; // not a sampled key
		const uint64_t qctx = a.keys[e] >> a.obj_bits;
		if (e > qb && (a.keys[e - 1] >> a.obj_bits) == qctx)
			continue;
		if (qctx > a.ctx_max)
			continue;
		uint64_t s_end = e + 1;
		while (s_end < qe && (a.keys[s_end] >> a.obj_bits) == qctx)
			++s_end;
		uint64_t d = lookup_ctx(a.ctxgid, a.bucket, a.shift, qctx);
		if (d == ~0ull)
			continue;
		if (a.vote_max_freq) { // skip contexts carried by more index records
			uint64_t x = d;
			while (x < a.n && x - d <= a.vote_max_freq && (a.ctxgid[x] >> GID_NBITS) == qctx)
				++x;
			if (x - d > a.vote_max_freq)
				continue;
		}
		const int qgid = a.qry_gid_offset + q;
		int owners = 0, sole_diff = 0;
		uint32_t sole_gid = 0;
		bool sole_allowed = false;
		const uint64_t rowbase = (uint64_t)q * (uint64_t)a.ref_gnum;
		while (d < a.n) {
			const uint64_t cg = a.ctxgid[d];
			if ((cg >> GID_NBITS) != qctx)
				break;
			const uint32_t gid = (uint32_t)(cg & GID_MASK);
			const uint64_t rb = d;
			do {
				++d;
			} while (d < a.n && a.ctxgid[d] == cg);
			const uint64_t re = d;
			if (lower && (int)gid >= qgid) {
				if ((int)gid > qgid)
					break;
				continue;
			}
			if (unique) {
				++owners;
				sole_gid = gid;
			}
			if (ignore && re - rb > 1)
				continue;
			int m = MIN_INIT;
			for (uint64_t x = e; x < s_end && m; ++x) {
				const uint32_t oq = (uint32_t)(a.keys[x] & objmask);
				for (uint64_t r = rb; r < re; ++r) {
					const uint32_t diff = oq ^ a.robj[r];
					if (diff == 0) {
						m = 0;
						break;
					}
					const int c = dna_popcount(diff);
					m = c < m ? c : m;
				}
			}
			const uint64_t cell = rowbase + gid;
			atomicAdd(&a.ctx[cell], 1ull | ((unsigned long long)(m > 1) << 32));
			if (m)
				atomicAdd(&a.objc[cell], 1ull | ((unsigned long long)m << 32));
			if (unique) {
				sole_diff = m;
				sole_allowed = true;
			}
		}
		if (unique && owners == 1 && sole_allowed) {
			const uint64_t cell = rowbase + sole_gid;
			atomicAdd(&a.uctx[cell], 1ull | ((unsigned long long)(sole_diff > 1) << 32));
			if (sole_diff)
				atomicAdd(&a.uobj[cell], 1ull | ((unsigned long long)sole_diff << 32));
		}
	}
}

// Leftmost matching index record per query key (or ~0).
__global__ void k_find_first(const uint64_t *ctxgid, const uint32_t *bucket, int shift,
			     uint64_t ctx_max, const uint64_t *keys, uint64_t n, unsigned obj_bits,
			     uint64_t *idx)
{
	for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n;
	     e += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t qctx = keys[e] >> obj_bits;
		idx[e] = qctx > ctx_max ? ~0ull : lookup_ctx(ctxgid, bucket, shift, qctx);
	}
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off, 64);
	return v;
}

// Plain kernel: merge-join of one pair (command_ani.h:303-346) in ONE launch.
// Threads stride over a's keys; the first key of each context group
// binary-searches b.  Each block writes its partial sums to the workspace;
// the last block to finish adds them up, writes *out and re-arms the
// workspace counter, so no separate zeroing launch is needed.
__global__ __launch_bounds__(256) void k_pair(const uint64_t *__restrict__ A, uint64_t n,
					      const uint64_t *__restrict__ B, uint64_t m,
					      unsigned obj_bits, int ignore_b,
					      kssd_ani_features_t *out, PairWorkspace *ws)
{
	const uint64_t objmask = obj_bits >= 64 ? ~0ull : ((1ull << obj_bits) - 1ull);
	uint32_t xny = 0, dobj = 0, dsec = 0, mut2 = 0;
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
	     i += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t ca = A[i] >> obj_bits;
		if (i && (A[i - 1] >> obj_bits) == ca)
			continue;
		uint64_t ae = i + 1;
		while (ae < n && (A[ae] >> obj_bits) == ca)
			++ae;
		uint64_t lo = 0, hi = m;
		while (lo < hi) {
			const uint64_t mid = lo + ((hi - lo) >> 1);
			if ((B[mid] >> obj_bits) < ca)
				lo = mid + 1;
			else
				hi = mid;
		}
		if (lo >= m || (B[lo] >> obj_bits) != ca)
			continue;
		uint64_t be = lo + 1;
		while (be < m && (B[be] >> obj_bits) == ca)
			++be;
		if (ignore_b && be - lo > 1)
			continue;
		int mn = MIN_INIT;
		for (uint64_t x = i; x < ae && mn; ++x) {
			const uint32_t oa = (uint32_t)(A[x] & objmask);
			for (uint64_t y = lo; y < be; ++y) {
				const uint32_t d = oa ^ (uint32_t)(B[y] & objmask);
				if (d == 0) {
					mn = 0;
					break;
				}
				const int c = dna_popcount(d);
				mn = c < mn ? c : mn;
			}
		}
		xny += 1;
		if (mn > 0) {
			dobj += 1;
			dsec += (uint32_t)mn;
			mut2 += mn > 1;
		}
	}
	xny = wave_sum(xny);
	dobj = wave_sum(dobj);
	dsec = wave_sum(dsec);
	mut2 = wave_sum(mut2);
	__shared__ uint4 part[4];
	__shared__ bool last;
	const int wave = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0)
		part[wave] = make_uint4(xny, dobj, dsec, mut2);
	__syncthreads();
	if (threadIdx.x == 0) {
		uint4 t = part[0];
		for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
			t.x += part[w].x;
			t.y += part[w].y;
			t.z += part[w].z;
			t.w += part[w].w;
		}
		ws->partial[blockIdx.x] = t;
		__threadfence();
		last = atomicAdd(&ws->done, 1u) == gridDim.x - 1;
	}
	__syncthreads();
	if (!last)
		return;
	// last block: reduce the per-block partials
	__threadfence();
	uint4 t = make_uint4(0, 0, 0, 0);
	for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
		const uint4 p = ws->partial[b];
		t.x += p.x;
		t.y += p.y;
		t.z += p.z;
		t.w += p.w;
	}
	t.x = wave_sum(t.x);
	t.y = wave_sum(t.y);
	t.z = wave_sum(t.z);
	t.w = wave_sum(t.w);
	__syncthreads();
	if ((threadIdx.x & 63) == 0)
		part[wave] = t;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint4 r = part[0];
		for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
			r.x += part[w].x;
			r.y += part[w].y;
			r.z += part[w].z;
			r.w += part[w].w;
		}
		out->XnY_ctx = r.x;
		out->X_ctx = 0;
		out->N_diff_obj = r.y;
		out->N_diff_obj_section = r.z;
		out->N_mut2_ctx = r.w;
		ws->done = 0;
	}
}

__global__ void k_pair_ws_init(PairWorkspace *ws)
{
	if (threadIdx.x == 0)
		ws->done = 0;
}

static unsigned grid_for(uint64_t n, unsigned block = 256, unsigned cap = 1u << 20)
{
	uint64_t g = (n + block - 1) / block;
	if (g == 0)
		g = 1;
	if (g > cap)
		g = cap;
	return (unsigned)g;
}

} // namespace kssd

using namespace kssd;

// ---------------------------------------------------------------- index handle
int DevBuf::reserve(size_t want)
{
	if (want <= bytes)
		return 0;
	if (ptr)
		(void)hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
	zero_bytes = 0;
	HIP_TRY(hipMalloc(&ptr, want));
	bytes = want;
	return 0;
}

int kssd_reserve_for_stream(DevBuf *b, size_t want, hipStream_t s, const char *what)
{
	if (want <= b->bytes)
		return 0;
	hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
	if (s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
		return kssdhip_fail(KSSDHIP_EINVAL,
				    "%s: %zu bytes needed while the stream is being captured (the handle's workspace "
				    "holds %zu); run the call once outside the capture first",
				    what, want, b->bytes);
	return b->reserve(want);
}

void DevBuf::release()
{
	if (ptr)
		(void)hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
	zero_bytes = 0;
}

static int bit_width_u64(uint64_t v)
{
	return v ? 64 - __builtin_clzll(v) : 0;
}

// KSSDHIP_TIMING set: the phases of kssdhip_index_open on stderr, each
// ended by a stream synchronisation of its own
static void open_phase(hipStream_t s, const char *what)
{
	static const bool on = getenv("KSSDHIP_TIMING") != nullptr;
	static double last = 0;
	if (!on)
		return;
	if (s)
		(void)hipStreamSynchronize(s);
	timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	const double t = ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
	if (last > 0)
		fprintf(stderr, "[kssd3a-hip index_open] %7.2f ms  %s\n", t - last, what);
	last = t;
}

extern "C" int kssdhip_index_open(const kssd_ctxgidobj_t *records, size_t n, int ref_gnum,
				  unsigned obj_bits, kssdhip_index **out)
{
	open_phase(nullptr, "start");
	if (!out)
		return kssdhip_fail(KSSDHIP_EINVAL, "index_open: null output handle");
	*out = nullptr;
	if (n && !records)
		return kssdhip_fail(KSSDHIP_EINVAL, "index_open: null records");
	if (n >= 0xffffffffull)
		return kssdhip_fail(KSSDHIP_EINVAL, "index_open: %zu records exceed 2^32-1", n);
	if (ref_gnum <= 0 || ref_gnum > (1 << GID_NBITS))
		return kssdhip_fail(KSSDHIP_EINVAL, "index_open: ref_gnum %d out of range", ref_gnum);
	if (obj_bits < 1 || obj_bits > 32)
		return kssdhip_fail(KSSDHIP_EINVAL, "index_open: obj_bits %u out of range", obj_bits);
	for (size_t probe = 0; probe < n && probe < 2; ++probe) {
		const size_t at = probe ? n - 1 : 0;
		if ((records[at].ctxgid & GID_MASK) >= (uint64_t)ref_gnum)
			return kssdhip_fail(KSSDHIP_EINVAL, "index_open: gid beyond ref_gnum");
	}

	kssdhip_index *ix = new (std::nothrow) kssdhip_index();
	if (!ix)
		return kssdhip_fail(KSSDHIP_ENOMEM, "index_open: out of host memory");
	ix->device = g_device;
	ix->n = n;
	ix->ref_gnum = ref_gnum;
	ix->obj_bits = obj_bits;
	ix->ctx_max = n ? (records[n - 1].ctxgid >> GID_NBITS) : 0;
	const int bw = bit_width_u64(ix->ctx_max);
	int k = n > 4 ? bit_width_u64(n) - 2 : 0; // ~2-4 records per bucket
	if (k > 26)
		k = 26;
	if (k > bw)
		k = bw;
	ix->k = k;
	ix->shift = bw - k;

	int rc = 0;
	void *d_raw = nullptr;
	const size_t nb = ((size_t)1 << k) + 1;
	do {
		if (hipSetDevice(ix->device) != hipSuccess ||
		    (!(ix->stream = take_spare_stream(ix->device)) &&
		     hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking) != hipSuccess)) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "index_open: stream creation failed");
			break;
		}
		open_phase(nullptr, "stream created");
		// padded so 16-byte vector loads of aligned 4-record groups that
		// straddle the end stay inside the allocation
		const size_t nalloc = n + INDEX_PAD;
		if (hipMalloc(&ix->d_ctxgid, nalloc * sizeof(uint64_t)) != hipSuccess ||
		    hipMalloc(&ix->d_obj, nalloc * sizeof(uint32_t)) != hipSuccess ||
		    hipMalloc(&ix->d_bucket, nb * sizeof(uint32_t)) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_ENOMEM, "index_open: device allocation of %zu records failed", n);
			break;
		}
		open_phase(nullptr, "index arrays allocated");
		if (n) {
			if (hipMalloc(&d_raw, n * sizeof(kssd_ctxgidobj_t)) != hipSuccess) {
				rc = kssdhip_fail(KSSDHIP_ENOMEM, "index_open: staging allocation failed");
				break;
			}
			if (copy_h2d(d_raw, records, n * sizeof(kssd_ctxgidobj_t), ix->stream) != hipSuccess) {
				rc = kssdhip_fail(KSSDHIP_EHIP, "index_open: upload failed");
				break;
			}
			open_phase(ix->stream, "records uploaded");
			hipLaunchKernelGGL(k_unpack_records, dim3(grid_for(n)), dim3(256), 0, ix->stream,
					   (const uint32_t *)d_raw, (uint64_t)n, ix->d_ctxgid, ix->d_obj);
			uint32_t *d_bad = (uint32_t *)d_raw; // staging no longer needed after unpack
			uint32_t bad = 0;
			hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ix->stream, d_bad, (uint64_t)1, 0u);
			hipLaunchKernelGGL(k_validate, dim3(grid_for(n)), dim3(256), 0, ix->stream,
					   ix->d_ctxgid, ix->d_obj, (uint64_t)n, (uint32_t)ref_gnum, obj_bits, d_bad);
			if (hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, ix->stream) != hipSuccess ||
			    hipStreamSynchronize(ix->stream) != hipSuccess) {
				rc = kssdhip_fail(KSSDHIP_EHIP, "index_open: validation failed to run");
				break;
			}
			if (bad) {
				rc = kssdhip_fail(KSSDHIP_EINVAL,
						  "index_open: records not sorted by (ctxgid, obj), gid >= ref_gnum or an "
						  "object wider than obj_bits");
				break;
			}
		}
		open_phase(nullptr, "unpacked and validated");
		hipLaunchKernelGGL(k_buckets, dim3(grid_for(nb)), dim3(256), 0, ix->stream,
				   ix->d_ctxgid, (uint64_t)n, k, ix->shift, ix->d_bucket);
		if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ix->stream) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "index_open: bucket build failed");
			break;
		}
		int dev_cus = 0;
		if (hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ix->device) == hipSuccess)
			ix->cus = dev_cus;
		kssdhip_tiling_t t0;
		kssd_default_tiling(ix, &t0);
		ix->mode = 0;
		ix->block = t0.block_threads;
		ix->group = t0.group_partitions;
		ix->auto_min_keys = 512.0;
		open_phase(nullptr, "bucket table");
		if ((rc = kssd_build_partitions(ix, t0.chunk_records)))
			break;
		open_phase(nullptr, "partition table");
		if ((rc = kssd_qjoin_build(ix)))
			break;
		open_phase(nullptr, "query-major tables");
		kssdhip_tiling_t tc;
		if (kssd_cached_index_tiling(ix, &tc) && kssdhip_index_set_tiling(ix, &tc) != 0)
			(void)kssdhip_index_set_tiling(ix, &t0); // an unusable cache entry: the default shape
		open_phase(nullptr, "tiling");
	} while (0);
	if (d_raw)
		(void)hipFree(d_raw);
	if (rc) {
		kssdhip_index_close(ix);
		return rc;
	}
	*out = ix;
	return 0;
}

extern "C" int kssdhip_index_close(kssdhip_index *ix)
{
	if (!ix)
		return 0;
	(void)hipSetDevice(ix->device);
	if (ix->stream)
		(void)hipStreamSynchronize(ix->stream);
	if (ix->d_ctxgid)
		(void)hipFree(ix->d_ctxgid);
	if (ix->d_obj)
		(void)hipFree(ix->d_obj);
	if (ix->d_bucket)
		(void)hipFree(ix->d_bucket);
	kssd_free_partitions(ix);
	kssd_qjoin_free(ix);
	DevBuf *bufs[] = {&ix->keys,  &ix->skidx, &ix->ctx,  &ix->obj,  &ix->uctx,  &ix->uobj,  &ix->idx,
			  &ix->split, &ix->qws,   &ix->rows, &ix->qmeta, &ix->qctxb, &ix->rctxb, &ix->sdense};
	for (DevBuf *b : bufs)
		b->release();
	if (ix->stream)
		(void)hipStreamDestroy(ix->stream);
	delete ix;
	return 0;
}

extern "C" size_t kssdhip_index_device_bytes(const kssdhip_index *ix)
{
	if (!ix)
		return 0;
	size_t b = ((size_t)ix->n + INDEX_PAD) * 12 + ((((size_t)1) << ix->k) + 1) * 4;
	b += ix->keys.bytes + ix->skidx.bytes + ix->ctx.bytes + ix->obj.bytes + ix->uctx.bytes +
	     ix->uobj.bytes + ix->idx.bytes + ix->split.bytes + ((size_t)ix->P + 1) * 12;
	b += (size_t)ix->br_lines * 128 + ix->qws.bytes + ix->sdense.bytes + ix->rows.bytes +
	     ix->qmeta.bytes + ix->qctxb.bytes + ix->rctxb.bytes;
	return b;
}

int kssdhip_index_info(const kssdhip_index *ix, IndexView *v)
{
	if (!ix || !v)
		return kssdhip_fail(KSSDHIP_EINVAL, "null index");
	v->ctxgid = ix->d_ctxgid;
	v->obj = ix->d_obj;
	v->bucket = ix->d_bucket;
	v->n = ix->n;
	v->ref_gnum = ix->ref_gnum;
	v->obj_bits = ix->obj_bits;
	v->shift = ix->shift;
	v->k = ix->k;
	v->ctx_max = ix->ctx_max;
	v->device = ix->device;
	return 0;
}

// ---------------------------------------------------------------- counting
int kssd_count_dev(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx, int nq, uint64_t total,
		   int qry_gid_offset, uint32_t flags, kssd_ctx_mut2_t *d_ctx, kssd_obj_section_t *d_obj,
		   kssd_ctx_mut2_t *d_uctx, kssd_obj_section_t *d_uobj, hipStream_t s, DevBuf *split, DevBuf *qws,
		   uint32_t vote_max_freq, uint32_t vote_step)
{
	if (!ix)
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: null index");
	if (nq < 0 || (nq > 0 && (!d_skidx || !d_ctx || !d_obj)) || (total && !d_keys))
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: bad arguments");
	const bool unique = (flags & KSSDHIP_UNIQUE) && !(flags & KSSDHIP_LOWER);
	if (unique && (!d_uctx || !d_uobj))
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: UNIQUE needs unique outputs");
	if (flags & ~(KSSDHIP_LOWER | KSSDHIP_UNIQUE | KSSDHIP_IGNORECONFLICT | KSSDHIP_ACCUMULATE))
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: unknown flags 0x%x", flags);
	const uint64_t cells = (uint64_t)nq * (uint64_t)ix->ref_gnum;
	// the query-major kernel writes every cell of its rows itself (or adds
	// into them under ACCUMULATE), so it needs no zeroed outputs
	const bool work = total && nq && ix->n;
	const bool voting = vote_max_freq || vote_step > 1;
	const bool qjoin = work && qws && !voting && ix->d_brec &&
			   (ix->mode == 3 || (ix->mode == 0 && kssd_qjoin_auto(ix, nq, total)));
	if (cells && !(flags & KSSDHIP_ACCUMULATE) && !qjoin) {
		hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(cells)), dim3(256), 0, s,
				   (unsigned long long *)d_ctx, cells, 0ull);
		hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(cells)), dim3(256), 0, s,
				   (unsigned long long *)d_obj, cells, 0ull);
		if (unique) {
			hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(cells)), dim3(256), 0, s,
					   (unsigned long long *)d_uctx, cells, 0ull);
			hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(cells)), dim3(256), 0, s,
					   (unsigned long long *)d_uobj, cells, 0ull);
		}
	}
	if (work) {
		JoinArgs a;
		a.ctxgid = ix->d_ctxgid;
		a.robj = ix->d_obj;
		a.bucket = ix->d_bucket;
		a.n = ix->n;
		a.shift = ix->shift;
		a.ctx_max = ix->ctx_max;
		a.ref_gnum = ix->ref_gnum;
		a.keys = d_keys;
		a.skidx = d_skidx;
		a.nq = nq;
		a.total = total;
		a.qry_gid_offset = qry_gid_offset;
		a.obj_bits = ix->obj_bits;
		a.flags = flags;
		a.vote_max_freq = vote_max_freq;
		a.vote_step = vote_step ? vote_step : 1;
		a.ctx = (unsigned long long *)d_ctx;
		a.objc = (unsigned long long *)d_obj;
		a.uctx = (unsigned long long *)d_uctx;
		a.uobj = (unsigned long long *)d_uobj;
		if (qjoin) {
			int rc = kssd_launch_qjoin_dense(ix, a, s, qws);
			if (rc)
				return rc;
			HIP_TRY(hipGetLastError());
			return 0;
		}
		bool used = false;
		int rc = ix->mode == 3 ? 0 : kssd_launch_pjoin(ix, a, s, split, &used);
		if (rc)
			return rc;
		if (!used)
			hipLaunchKernelGGL(k_join, dim3(grid_for(total)), dim3(256), 0, s, a);
	}
	HIP_TRY(hipGetLastError());
	return 0;
}

extern "C" int kssdhip_count_block_dev(kssdhip_index *ix, const uint64_t *d_keys,
				       const uint64_t *d_skidx, int nq, uint64_t total,
				       int qry_gid_offset, uint32_t flags, kssd_ctx_mut2_t *d_ctx,
				       kssd_obj_section_t *d_obj, kssd_ctx_mut2_t *d_uctx,
				       kssd_obj_section_t *d_uobj, void *stream)
{
	if (!ix)
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: null index");
	return kssd_count_dev(ix, d_keys, d_skidx, nq, total, qry_gid_offset, flags, d_ctx, d_obj, d_uctx, d_uobj,
			      (hipStream_t)stream, &ix->split, &ix->qws);
}

extern "C" int kssdhip_count_block(kssdhip_index *ix, const uint64_t *section_sk,
				   const uint64_t *section_skidx, int nq, int qry_gid_offset,
				   uint32_t flags, kssd_ctx_mut2_t *ctx, kssd_obj_section_t *obj,
				   kssd_ctx_mut2_t *uctx, kssd_obj_section_t *uobj)
{
	if (!ix)
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: null index");
	if (nq < 0 || (nq > 0 && (!section_skidx || !ctx || !obj)))
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: bad arguments");
	if (nq == 0)
		return 0;
	const bool unique = (flags & KSSDHIP_UNIQUE) && !(flags & KSSDHIP_LOWER);
	if (unique && (!uctx || !uobj))
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: UNIQUE needs unique outputs");
	const uint64_t base = section_skidx[0];
	std::vector<uint64_t> rel((size_t)nq + 1);
	for (int i = 0; i <= nq; ++i) {
		if (section_skidx[i] < base || (i && section_skidx[i] < section_skidx[i - 1]))
			return kssdhip_fail(KSSDHIP_EINVAL, "count_block: section_skidx not ascending");
		rel[i] = section_skidx[i] - base;
	}
	const uint64_t total = rel[nq];
	if (total && !section_sk)
		return kssdhip_fail(KSSDHIP_EINVAL, "count_block: null keys");
	const size_t cells = (size_t)nq * (size_t)ix->ref_gnum;
	HIP_TRY(hipSetDevice(ix->device));
	int rc;
	if ((rc = ix->keys.reserve((total ? total : 1) * 8)) || (rc = ix->skidx.reserve(rel.size() * 8)) ||
	    (rc = ix->ctx.reserve(cells * 8)) || (rc = ix->obj.reserve(cells * 8)))
		return rc;
	if (unique && ((rc = ix->uctx.reserve(cells * 8)) || (rc = ix->uobj.reserve(cells * 8))))
		return rc;
	hipStream_t s = ix->stream;
	if (total)
		HIP_TRY(copy_h2d(ix->keys.ptr, section_sk, total * 8, s));
	HIP_TRY(hipMemcpyAsync(ix->skidx.ptr, rel.data(), rel.size() * 8, hipMemcpyHostToDevice, s));
	rc = kssdhip_count_block_dev(ix, (const uint64_t *)ix->keys.ptr, (const uint64_t *)ix->skidx.ptr,
				     nq, total, qry_gid_offset, flags, (kssd_ctx_mut2_t *)ix->ctx.ptr,
				     (kssd_obj_section_t *)ix->obj.ptr,
				     unique ? (kssd_ctx_mut2_t *)ix->uctx.ptr : nullptr,
				     unique ? (kssd_obj_section_t *)ix->uobj.ptr : nullptr, s);
	if (rc)
		return rc;
	HIP_TRY(copy_d2h(ctx, ix->ctx.ptr, cells * 8, s));
	HIP_TRY(copy_d2h(obj, ix->obj.ptr, cells * 8, s));
	if (unique) {
		HIP_TRY(copy_d2h(uctx, ix->uctx.ptr, cells * 8, s));
		HIP_TRY(copy_d2h(uobj, ix->uobj.ptr, cells * 8, s));
	}
	HIP_TRY(hipStreamSynchronize(s));
	return 0;
}

extern "C" int kssdhip_count_votes(kssdhip_index *ix, const uint64_t *section_sk, const uint64_t *section_skidx,
				   int nq, int qry_gid_offset, uint32_t max_ctx_freq, uint32_t sample_step,
				   uint32_t *votes)
{
	if (!ix)
		return kssdhip_fail(KSSDHIP_EINVAL, "count_votes: null index");
	if (nq < 0 || (nq > 0 && (!section_skidx || !votes)) || sample_step < 1)
		return kssdhip_fail(KSSDHIP_EINVAL, "count_votes: bad arguments");
	if (nq == 0)
		return 0;
	const uint64_t base = section_skidx[0];
	std::vector<uint64_t> rel((size_t)nq + 1);
	for (int i = 0; i <= nq; ++i) {
		if (section_skidx[i] < base || (i && section_skidx[i] < section_skidx[i - 1]))
			return kssdhip_fail(KSSDHIP_EINVAL, "count_votes: section_skidx not ascending");
		rel[i] = section_skidx[i] - base;
	}
	const uint64_t total = rel[nq];
	if (total && !section_sk)
		return kssdhip_fail(KSSDHIP_EINVAL, "count_votes: null keys");
	const size_t cells = (size_t)nq * (size_t)ix->ref_gnum;
	HIP_TRY(hipSetDevice(ix->device));
	int rc;
	if ((rc = ix->keys.reserve((total ? total : 1) * 8)) || (rc = ix->skidx.reserve(rel.size() * 8)) ||
	    (rc = ix->ctx.reserve(cells * 8)) || (rc = ix->obj.reserve(cells * 8)))
		return rc;
	hipStream_t s = ix->stream;
	if (total)
		HIP_TRY(copy_h2d(ix->keys.ptr, section_sk, total * 8, s));
	HIP_TRY(hipMemcpyAsync(ix->skidx.ptr, rel.data(), rel.size() * 8, hipMemcpyHostToDevice, s));
	rc = kssd_count_dev(ix, (const uint64_t *)ix->keys.ptr, (const uint64_t *)ix->skidx.ptr, nq, total,
			    qry_gid_offset, KSSDHIP_LOWER, (kssd_ctx_mut2_t *)ix->ctx.ptr,
			    (kssd_obj_section_t *)ix->obj.ptr, nullptr, nullptr, s, &ix->split, nullptr, max_ctx_freq,
			    sample_step);
	if (rc)
		return rc;
	if (cells)
		hipLaunchKernelGGL(k_low_words, dim3(grid_for(cells)), dim3(256), 0, s,
				   (const unsigned long long *)ix->ctx.ptr, cells, (uint32_t *)ix->obj.ptr);
	HIP_TRY(hipGetLastError());
	HIP_TRY(copy_d2h(votes, ix->obj.ptr, cells * 4, s));
	HIP_TRY(hipStreamSynchronize(s));
	return 0;
}

extern "C" int kssdhip_find_first(kssdhip_index *ix, const uint64_t *a, size_t n, uint64_t *idx)
{
	if (!ix || (n && (!a || !idx)))
		return kssdhip_fail(KSSDHIP_EINVAL, "find_first: bad arguments");
	if (!n)
		return 0;
	HIP_TRY(hipSetDevice(ix->device));
	int rc;
	if ((rc = ix->keys.reserve(n * 8)) || (rc = ix->idx.reserve(n * 8)))
		return rc;
	hipStream_t s = ix->stream;
	HIP_TRY(hipMemcpyAsync(ix->keys.ptr, a, n * 8, hipMemcpyHostToDevice, s));
	if (ix->n)
		hipLaunchKernelGGL(k_find_first, dim3(grid_for(n)), dim3(256), 0, s, ix->d_ctxgid,
				   ix->d_bucket, ix->shift, ix->ctx_max, (const uint64_t *)ix->keys.ptr,
				   (uint64_t)n, ix->obj_bits, (uint64_t *)ix->idx.ptr);
	else
		hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(256), 0, s,
				   (unsigned long long *)ix->idx.ptr, (uint64_t)n, ~0ull);
	HIP_TRY(hipGetLastError());
	HIP_TRY(hipMemcpyAsync(idx, ix->idx.ptr, n * 8, hipMemcpyDeviceToHost, s));
	HIP_TRY(hipStreamSynchronize(s));
	return 0;
}

// ---------------------------------------------------------------- plain kernel
static_assert(sizeof(PairWorkspace) <= KSSDHIP_PAIR_WORKSPACE_BYTES, "pair workspace size");

extern "C" int kssdhip_pair_workspace_init_dev(void *d_workspace, void *stream)
{
	if (!d_workspace)
		return kssdhip_fail(KSSDHIP_EINVAL, "pair workspace: null");
	hipLaunchKernelGGL(k_pair_ws_init, dim3(1), dim3(64), 0, (hipStream_t)stream,
			   (PairWorkspace *)d_workspace);
	HIP_TRY(hipGetLastError());
	return 0;
}

extern "C" int kssdhip_pair_features_dev(const uint64_t *d_a, size_t n, const uint64_t *d_b, size_t m,
					 unsigned obj_bits, int ignore_b_conflict,
					 kssd_ani_features_t *d_out, void *d_workspace, void *stream)
{
	if (!d_out || !d_workspace || (n && !d_a) || (m && !d_b))
		return kssdhip_fail(KSSDHIP_EINVAL, "pair_features: bad arguments");
	if (obj_bits < 1 || obj_bits > 32)
		return kssdhip_fail(KSSDHIP_EINVAL, "pair_features: obj_bits %u out of range", obj_bits);
	uint64_t g = (n + 255) / 256;
	if (g < 1)
		g = 1;
	if (g > PAIR_MAX_BLOCKS)
		g = PAIR_MAX_BLOCKS;
	hipLaunchKernelGGL(k_pair, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, d_a,
			   (uint64_t)(m ? n : 0), d_b, (uint64_t)m, obj_bits, ignore_b_conflict, d_out,
			   (PairWorkspace *)d_workspace);
	HIP_TRY(hipGetLastError());
	return 0;
}

namespace {
struct PairScratch {
	int device = -1;
	hipStream_t stream = nullptr;
	DevBuf a, b, out, ws;
};
thread_local PairScratch g_pair;
} // namespace

extern "C" int kssdhip_pair_features(const uint64_t *a, size_t n, const uint64_t *b, size_t m,
				     unsigned obj_bits, int ignore_b_conflict,
				     kssd_ani_features_t *out)
{
	if (!out || (n && !a) || (m && !b))
		return kssdhip_fail(KSSDHIP_EINVAL, "pair_features: bad arguments");
	HIP_TRY(hipSetDevice(g_device));
	PairScratch &ps = g_pair;
	if (ps.device != g_device) {
		ps.a.release();
		ps.b.release();
		ps.out.release();
		ps.ws.release();
		if (ps.stream)
			(void)hipStreamDestroy(ps.stream);
		HIP_TRY(hipStreamCreateWithFlags(&ps.stream, hipStreamNonBlocking));
		ps.device = g_device;
	}
	int rc;
	if ((rc = ps.a.reserve((n ? n : 1) * 8)) || (rc = ps.b.reserve((m ? m : 1) * 8)) ||
	    (rc = ps.out.reserve(sizeof(kssd_ani_features_t))))
		return rc;
	if (!ps.ws.ptr) {
		if ((rc = ps.ws.reserve(KSSDHIP_PAIR_WORKSPACE_BYTES)) ||
		    (rc = kssdhip_pair_workspace_init_dev(ps.ws.ptr, ps.stream)))
			return rc;
	}
	if (n)
		HIP_TRY(hipMemcpyAsync(ps.a.ptr, a, n * 8, hipMemcpyHostToDevice, ps.stream));
	if (m)
		HIP_TRY(hipMemcpyAsync(ps.b.ptr, b, m * 8, hipMemcpyHostToDevice, ps.stream));
	rc = kssdhip_pair_features_dev((const uint64_t *)ps.a.ptr, n, (const uint64_t *)ps.b.ptr, m,
				       obj_bits, ignore_b_conflict, (kssd_ani_features_t *)ps.out.ptr,
				       ps.ws.ptr, ps.stream);
	if (rc)
		return rc;
	HIP_TRY(hipMemcpyAsync(out, ps.out.ptr, sizeof *out, hipMemcpyDeviceToHost, ps.stream));
	HIP_TRY(hipStreamSynchronize(ps.stream));
	return 0;
}
