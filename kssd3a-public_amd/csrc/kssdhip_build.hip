// Inverted-index builder on the GPU: the sorted record array of a sketch
// (sortedcomb_ctxgid64obj32) and its unique-marker bitset.
//
// The reference builds the index on the host (gen_inverted_index4comblco,
// command_sketch.c:916-938): every key of sample g becomes the record
// {ctx << 20 | g, obj} (uint64_ctxobj2ctxgidobj96, sketch_rearrange.h:95-100),
// the records are sorted by (ctxgid, obj) (ctxgidobj_sort_array,
// kssdlib_sort.c:215-219), and a bit per record is set iff its object occurs
// once among the records of its context (write_unique_marker_bits_from_sorted,
// command_sketch.c:855-914).
//
// Here: the records are made in input order, which is gid-major with each
// sample's keys ascending, so a STABLE sort by context alone yields the
// (ctx, gid, obj) order.  The stable sort is an LSD radix sort over the 44
// context bits (6 passes of 8-bit digits over bits 20..63 of ctxgid, object
// carried along): per 4096-record tile a digit histogram, one exclusive scan
// of the digit-major histogram table, then a scatter that ranks each tile's
// records with wave ballots (stable), stages the tile in LDS in digit order
// and writes each digit's run contiguously.  Unique bits: a thread per record
// compares its object with the rest of a context group of up to 64 records;
// larger groups get a workgroup each, which sorts (object, position) pairs in
// LDS (or, past the LDS, with the same radix sort in HBM) and marks the
// objects that occur once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_internal.h"

using namespace kssd;

namespace {

constexpr int RS_THREADS = 256;                 // 4 waves
constexpr int RS_ROUNDS = 16;                   // 64-record rounds per wave
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS; // 4096 records per tile
constexpr int RS_BINS = 256;
constexpr uint32_t SMALL_GROUP = 64;            // context groups checked by one thread per record
constexpr uint32_t LDS_GROUP = 4096;            // larger groups sorted in LDS up to this size

// ---------------------------------------------------------------- records
// One workgroup per sample: record i of sample g = {ctx << 20 | g, obj}.
// bad[0] is set when a sample's keys are not ascending or a context does
// not fit 44 bits.
__global__ void k_make_records(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ skidx, int nsamples,
			       unsigned obj_bits, uint64_t *ctxgid, uint32_t *obj, unsigned *bad)
{
	const uint64_t objmask = obj_bits >= 64 ? ~0ull : ((1ull << obj_bits) - 1ull);
	for (int g = blockIdx.x; g < nsamples; g += gridDim.x) {
		const uint64_t b = skidx[g], e = skidx[g + 1];
		for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
			const uint64_t k = keys[i];
			const uint64_t c = k >> obj_bits;
			if ((c >> (64 - GID_NBITS)) != 0 || (i > b && keys[i - 1] > k))
				atomicOr(bad, 1u);
			ctxgid[i] = (c << GID_NBITS) | (uint64_t)g;
			obj[i] = (uint32_t)(k & objmask);
		}
	}
}

// ---------------------------------------------------------------- radix sort
// hist[d * ntiles + t] = number of records of tile t whose digit is d.
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ key, uint64_t n, int shift,
							uint32_t ntiles, uint32_t *hist)
{
	__shared__ uint32_t cnt[RS_BINS];
	const uint32_t t = blockIdx.x;
	cnt[threadIdx.x] = 0;
	__syncthreads();
	const uint64_t base = (uint64_t)t * RS_TILE;
	for (int r = 0; r < RS_ROUNDS; ++r) {
		const uint64_t i = base + (uint64_t)r * RS_THREADS + threadIdx.x;
		if (i < n)
			atomicAdd(&cnt[(key[i] >> shift) & (RS_BINS - 1)], 1u);
	}
	__syncthreads();
	hist[(uint64_t)threadIdx.x * ntiles + t] = cnt[threadIdx.x];
}

// Exclusive scan of a u32 array in three launches: per-block sums, a scan
// of the sums by one block, per-block scans with the block's offset.
constexpr int SCAN_THREADS = 256, SCAN_PER = 16, SCAN_BLOCK = SCAN_THREADS * SCAN_PER;

__device__ uint32_t block_scan_excl(uint32_t v, uint32_t *tmp, uint32_t *total)
{
	// tmp: LDS scratch of blockDim/64 + 1 words
	const uint32_t lane = threadIdx.x & 63, w = threadIdx.x / 64, nw = blockDim.x / 64;
	uint32_t x = v;
	for (int d = 1; d < 64; d <<= 1) {
		const uint32_t y = __shfl_up(x, d);
		if (lane >= (uint32_t)d)
			x += y;
	}
	if (lane == 63)
		tmp[w] = x;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t run = 0;
		for (uint32_t k = 0; k < nw; ++k) {
			const uint32_t c = tmp[k];
			tmp[k] = run;
			run += c;
		}
		tmp[nw] = run;
	}
	__syncthreads();
	const uint32_t r = tmp[w] + x - v;
	*total = tmp[nw];
	__syncthreads();
	return r;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(const uint32_t *__restrict__ a, uint64_t n, uint32_t *sums)
{
	__shared__ uint32_t tmp[SCAN_THREADS / 64 + 1];
	const uint64_t base = (uint64_t)blockIdx.x * SCAN_BLOCK + (uint64_t)threadIdx.x * SCAN_PER;
	uint32_t s = 0;
	for (int k = 0; k < SCAN_PER; ++k)
		if (base + k < n)
			s += a[base + k];
	uint32_t total;
	(void)block_scan_excl(s, tmp, &total);
	if (threadIdx.x == 0)
		sums[blockIdx.x] = total;
}

// One block: exclusive scan of sums[0, m) in place.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(uint32_t *sums, uint32_t m)
{
	__shared__ uint32_t tmp[SCAN_THREADS / 64 + 1];
	const uint32_t per = (m + SCAN_THREADS - 1) / SCAN_THREADS;
	const uint32_t b0 = threadIdx.x * per < m ? threadIdx.x * per : m, b1 = b0 + per < m ? b0 + per : m;
	uint32_t s = 0;
	for (uint32_t k = b0; k < b1; ++k)
		s += sums[k];
	uint32_t total;
	uint32_t off = block_scan_excl(s, tmp, &total);
	for (uint32_t k = b0; k < b1; ++k) {
		const uint32_t c = sums[k];
		sums[k] = off;
		off += c;
	}
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(uint32_t *a, uint64_t n, const uint32_t *__restrict__ sums)
{
	__shared__ uint32_t tmp[SCAN_THREADS / 64 + 1];
	const uint64_t base = (uint64_t)blockIdx.x * SCAN_BLOCK + (uint64_t)threadIdx.x * SCAN_PER;
	uint32_t v[SCAN_PER];
	uint32_t s = 0;
#pragma unroll
	for (int k = 0; k < SCAN_PER; ++k) {
		v[k] = base + k < n ? a[base + k] : 0;
		s += v[k];
	}
	uint32_t total;
	uint32_t off = sums[blockIdx.x] + block_scan_excl(s, tmp, &total);
#pragma unroll
	for (int k = 0; k < SCAN_PER; ++k)
		if (base + k < n) {
			a[base + k] = off;
			off += v[k];
		}
}

// Stable scatter of one radix pass.  Record i of tile t sits at wave
// w = (i % 4096) / 1024, round r, lane l; its rank among the tile's records
// of its digit counts earlier waves, earlier rounds of its wave and earlier
// lanes of its round (ballots), so equal digits keep their input order.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ kin,
							   const uint32_t *__restrict__ vin, uint64_t n, int shift,
							   uint32_t ntiles, const uint32_t *__restrict__ hist,
							   uint64_t *kout, uint32_t *vout)
{
	__shared__ uint32_t wcnt[RS_WAVES][RS_BINS]; // per-wave digit counts, then prefixes
	__shared__ uint32_t dstart[RS_BINS];         // first tile slot of each digit
	__shared__ uint32_t gbase[RS_BINS];          // first output slot of each digit
	__shared__ uint32_t tmp[RS_THREADS / 64 + 1];
	__shared__ uint64_t skey[RS_TILE];
	__shared__ uint32_t sval[RS_TILE];
	const uint32_t t = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x / 64;
	for (int k = threadIdx.x; k < RS_WAVES * RS_BINS; k += RS_THREADS)
		(&wcnt[0][0])[k] = 0;
	__syncthreads();
	const uint64_t base = (uint64_t)t * RS_TILE + (uint64_t)w * (64 * RS_ROUNDS);
	const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull; // lanes below this one
	uint64_t kk[RS_ROUNDS];
	uint32_t vv[RS_ROUNDS];
	uint32_t rk[RS_ROUNDS];
#pragma unroll
	for (int r = 0; r < RS_ROUNDS; ++r) {
		const uint64_t i = base + (uint64_t)r * 64 + lane;
		const bool valid = i < n;
		kk[r] = valid ? kin[i] : 0;
		vv[r] = valid ? vin[i] : 0;
		const uint32_t d = (uint32_t)(kk[r] >> shift) & (RS_BINS - 1);
		uint64_t m = __ballot(valid);
#pragma unroll
		for (int bit = 0; bit < 8; ++bit) {
			const uint64_t bb = __ballot((d >> bit) & 1u);
			m &= ((d >> bit) & 1u) ? bb : ~bb;
		}
		// m: lanes of this round with digit d (valid lanes only)
		const uint32_t before = (uint32_t)__popcll(m & lt);
		const uint32_t old = valid ? wcnt[w][d] : 0;
		if (valid && before == 0)
			wcnt[w][d] = old + (uint32_t)__popcll(m);
		rk[r] = old + before;
	}
	__syncthreads();
	// thread = digit: prefixes over the waves, tile total, start in the tile
	{
		const uint32_t d = threadIdx.x;
		uint32_t run = 0;
		for (int x = 0; x < RS_WAVES; ++x) {
			const uint32_t c = wcnt[x][d];
			wcnt[x][d] = run;
			run += c;
		}
		uint32_t total;
		dstart[d] = block_scan_excl(run, tmp, &total);
		gbase[d] = hist[(uint64_t)d * ntiles + t];
	}
	__syncthreads();
#pragma unroll
	for (int r = 0; r < RS_ROUNDS; ++r) {
		const uint64_t i = base + (uint64_t)r * 64 + lane;
		if (i < n) {
			const uint32_t d = (uint32_t)(kk[r] >> shift) & (RS_BINS - 1);
			const uint32_t slot = dstart[d] + wcnt[w][d] + rk[r];
			skey[slot] = kk[r];
			sval[slot] = vv[r];
		}
	}
	__syncthreads();
	const uint64_t tbase = (uint64_t)t * RS_TILE;
	const uint32_t cnt = n - tbase < (uint64_t)RS_TILE ? (uint32_t)(n - tbase) : (uint32_t)RS_TILE;
	for (uint32_t s = threadIdx.x; s < cnt; s += RS_THREADS) {
		const uint64_t k = skey[s];
		const uint32_t d = (uint32_t)(k >> shift) & (RS_BINS - 1);
		const uint64_t o = (uint64_t)gbase[d] + (s - dstart[d]);
		kout[o] = k;
		vout[o] = sval[s];
	}
}

// ---------------------------------------------------------------- unique bits
__device__ __forceinline__ void set_bit(uint32_t *bits, uint64_t i)
{
	atomicOr(&bits[i >> 5], 1u << (i & 31));
}

// A thread per record, 256 records per workgroup, whose contexts and
// objects (with SMALL_GROUP records on each side) are staged in LDS with
// coalesced loads.  Groups of <= SMALL_GROUP records: the record's object is
// compared with the group's others.  A larger group is listed (by its first
// record) for k_unique_large.  Each wave owns the 64 bits of its 64 records
// and stores them as one word (bits of large groups stay 0 here).
constexpr int UQ_THREADS = 256;
constexpr int UQ_SPAN = UQ_THREADS + 2 * SMALL_GROUP + 1;
__global__ __launch_bounds__(UQ_THREADS) void k_unique_small(const uint64_t *__restrict__ ctxgid,
							     const uint32_t *__restrict__ obj, uint64_t n, uint32_t *bits,
							     uint64_t *large, unsigned *nlarge, unsigned large_cap)
{
	__shared__ uint64_t sc[UQ_SPAN]; // contexts of records [i0 - SMALL_GROUP - 1, i0 + 256 + SMALL_GROUP)
	__shared__ uint32_t so[UQ_SPAN];
	const uint64_t NONE = ~0ull; // past either end of the array: no context
	for (uint64_t i0 = (uint64_t)blockIdx.x * UQ_THREADS; i0 < n; i0 += (uint64_t)gridDim.x * UQ_THREADS) {
		const int64_t lo = (int64_t)i0 - (int64_t)SMALL_GROUP - 1;
		for (int k = threadIdx.x; k < UQ_SPAN; k += UQ_THREADS) {
			const int64_t x = lo + k;
			const bool in = x >= 0 && (uint64_t)x < n;
			sc[k] = in ? ctxgid[x] >> GID_NBITS : NONE;
			so[k] = in ? obj[x] : 0;
		}
		__syncthreads();
		const uint64_t i = i0 + threadIdx.x;
		bool once = false;
		if (i < n) {
			const int s = (int)SMALL_GROUP + 1 + threadIdx.x; // slot of record i
			const uint64_t c = sc[s];
			int b = s, e = s + 1;
#pragma unroll 1
			while (s - b < (int)SMALL_GROUP && sc[b - 1] == c)
				--b;
#pragma unroll 1
			while (e - s <= (int)SMALL_GROUP && sc[e] == c)
				++e;
			const bool head = sc[s - 1] != c;
			if (e - b > (int)SMALL_GROUP) {
				// too large for this kernel: its first record lists it
				if (head) {
					const unsigned k = atomicAdd(nlarge, 1u);
					if (k < large_cap)
						large[k] = i;
				}
			} else {
				const uint32_t o = so[s];
				once = true;
#pragma unroll 1
				for (int x = b; x < e && once; ++x)
					once = x == s || so[x] != o;
			}
		}
		const uint64_t word = __ballot(once);
		if ((threadIdx.x & 63) == 0 && i < n)
			((uint64_t *)bits)[i >> 6] = word; // i is a multiple of 64 here
		__syncthreads();
	}
}

// One workgroup per large group with at most LDS_GROUP records: sorts the
// group's (object << 32 | position) pairs in LDS (bitonic) and marks the
// records whose object occurs once.  Groups past LDS_GROUP are marked in
// `over` for the HBM path.
__global__ __launch_bounds__(1024) void k_unique_large(const uint64_t *__restrict__ ctxgid,
						       const uint32_t *__restrict__ obj, uint64_t n,
						       const uint64_t *__restrict__ large, unsigned nlarge,
						       uint32_t *bits, uint64_t *over, unsigned *nover)
{
	__shared__ uint64_t pr[LDS_GROUP];
	__shared__ uint32_t gsize;
	for (unsigned gi = blockIdx.x; gi < nlarge; gi += gridDim.x) {
		const uint64_t b = large[gi];
		const uint64_t c = ctxgid[b] >> GID_NBITS;
		if (threadIdx.x == 0) {
			// group length, capped: a stride-doubling search for its end
			uint64_t lo = b + 1, step = 1, hi = b + 2;
			while (hi <= n && hi - b <= LDS_GROUP + 1 && (ctxgid[hi - 1] >> GID_NBITS) == c) {
				lo = hi;
				step <<= 1;
				hi = b + 1 + step;
			}
			if (hi > n)
				hi = n;
			if (hi - b > LDS_GROUP + 1)
				hi = b + LDS_GROUP + 1;
			// first record in [lo, hi) of another context (or hi)
			uint64_t x = lo, y = hi;
			while (x < y) {
				const uint64_t mid = x + ((y - x) >> 1);
				if ((ctxgid[mid] >> GID_NBITS) == c)
					x = mid + 1;
				else
					y = mid;
			}
			gsize = (uint32_t)(x - b);
		}
		__syncthreads();
		const uint32_t g = gsize;
		if (g > LDS_GROUP) {
			if (threadIdx.x == 0)
				over[atomicAdd(nover, 1u)] = b;
			__syncthreads();
			continue;
		}
		uint32_t np = 1;
		while (np < g)
			np <<= 1;
		for (uint32_t k = threadIdx.x; k < np; k += blockDim.x)
			pr[k] = k < g ? ((uint64_t)obj[b + k] << 32) | k : ~0ull;
		__syncthreads();
		for (uint32_t size = 2; size <= np; size <<= 1)
			for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
				for (uint32_t k = threadIdx.x; k < np; k += blockDim.x) {
					const uint32_t p = k ^ stride;
					if (p > k) {
						const bool up = (k & size) == 0;
						const uint64_t a = pr[k], q = pr[p];
						if ((a > q) == up) {
							pr[k] = q;
							pr[p] = a;
						}
					}
				}
				__syncthreads();
			}
		for (uint32_t k = threadIdx.x; k < g; k += blockDim.x) {
			const uint32_t o = (uint32_t)(pr[k] >> 32);
			const bool once = (k == 0 || (uint32_t)(pr[k - 1] >> 32) != o) &&
					  (k + 1 >= g || (uint32_t)(pr[k + 1] >> 32) != o);
			if (once)
				set_bit(bits, b + (uint32_t)pr[k]);
		}
		__syncthreads();
	}
}

// Length of the context group that starts at record b (one thread).
__global__ void k_group_len(const uint64_t *__restrict__ ctxgid, uint64_t n, uint64_t b, uint64_t *len)
{
	const uint64_t c = ctxgid[b] >> GID_NBITS;
	uint64_t lo = b + 1, step = 1, hi = b + 2; // ctxgid[lo - 1] is in the group
	while (hi <= n && (ctxgid[hi - 1] >> GID_NBITS) == c) {
		lo = hi;
		step <<= 1;
		hi = b + 1 + step;
	}
	if (hi > n)
		hi = n;
	while (lo < hi) {
		const uint64_t mid = lo + ((hi - lo) >> 1);
		if ((ctxgid[mid] >> GID_NBITS) == c)
			lo = mid + 1;
		else
			hi = mid;
	}
	*len = lo - b;
}

// Giant group [b, b + g): objects with their positions, sorted in HBM by
// the radix sort; mark the objects that occur once.
__global__ void k_giant_pairs(const uint32_t *__restrict__ obj, uint64_t b, uint64_t g, uint64_t *key, uint32_t *val)
{
	for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < g; k += (uint64_t)gridDim.x * blockDim.x) {
		key[k] = obj[b + k];
		val[k] = (uint32_t)k;
	}
}

__global__ void k_giant_mark(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, uint64_t b, uint64_t g,
			     uint32_t *bits)
{
	for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < g; k += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t o = key[k];
		if ((k == 0 || key[k - 1] != o) && (k + 1 >= g || key[k + 1] != o))
			set_bit(bits, b + val[k]);
	}
}

__global__ void k_pack(const uint64_t *__restrict__ ctxgid, const uint32_t *__restrict__ obj, uint64_t n,
		       kssd_ctxgidobj_t *out)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
		kssd_ctxgidobj_t r;
		r.ctxgid = ctxgid[i];
		r.obj = obj[i];
		out[i] = r;
	}
}

unsigned grid_n(uint64_t n, unsigned block = 256)
{
	uint64_t g = (n + block - 1) / block;
	return (unsigned)std::min<uint64_t>(std::max<uint64_t>(g, 1), 1u << 20);
}

struct Mem {
	std::vector<void *> p;
	~Mem()
	{
		for (void *x : p)
			(void)hipFree(x);
	}
	template <typename T> int get(T **out, size_t count)
	{
		void *x = nullptr;
		if (hipMalloc(&x, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
			return kssdhip_fail(KSSDHIP_ENOMEM, "build_index: device allocation of %zu bytes failed",
					    count * sizeof(T));
		p.push_back(x);
		*out = (T *)x;
		return 0;
	}
};

// Stable LSD radix sort of (key, val)[0, n) by key bits [bit0, bit1),
// 8 bits per pass; the result ends in (k0, v0).  k1/v1: n-record buffers,
// hist: 256 * ceil(n / 4096) words, sums: ceil(that / 4096) words.
int radix_sort(uint64_t *&k0, uint32_t *&v0, uint64_t *&k1, uint32_t *&v1, uint64_t n, int bit0, int bit1,
	       uint32_t *hist, uint32_t *sums, hipStream_t s)
{
	if (n <= 1)
		return 0;
	const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
	const uint64_t hn = (uint64_t)RS_BINS * ntiles;
	const uint32_t nsum = (uint32_t)((hn + SCAN_BLOCK - 1) / SCAN_BLOCK);
	for (int sh = bit0; sh < bit1; sh += 8) {
		hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, k0, n, sh, ntiles, hist);
		hipLaunchKernelGGL(k_scan_sums, dim3(nsum), dim3(SCAN_THREADS), 0, s, hist, hn, sums);
		hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, s, sums, nsum);
		hipLaunchKernelGGL(k_scan_apply, dim3(nsum), dim3(SCAN_THREADS), 0, s, hist, hn, sums);
		hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, k0, v0, n, sh, ntiles, hist, k1,
				   v1);
		std::swap(k0, k1);
		std::swap(v0, v1);
	}
	const hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "build_index: radix sort launch: %s", hipGetErrorString(e));
	return 0;
}

} // namespace

static_assert(SCAN_BLOCK == KSSD_SCAN_BLOCK, "scan block size shared with the callers");

void kssd_scan_u32_excl(uint32_t *a, uint64_t n, uint32_t *sums, hipStream_t s)
{
	if (!n)
		return;
	const uint32_t nsum = (uint32_t)((n + SCAN_BLOCK - 1) / SCAN_BLOCK);
	hipLaunchKernelGGL(k_scan_sums, dim3(nsum), dim3(SCAN_THREADS), 0, s, a, n, sums);
	hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, s, sums, nsum);
	hipLaunchKernelGGL(k_scan_apply, dim3(nsum), dim3(SCAN_THREADS), 0, s, a, n, sums);
}

extern "C" int kssdhip_build_index(const uint64_t *keys, const uint64_t *skidx, int nsamples, unsigned obj_bits,
				   kssd_ctxgidobj_t *records, uint8_t *unique_bits)
{
	if (!skidx || nsamples < 0)
		return kssdhip_fail(KSSDHIP_EINVAL, "build_index: null argument");
	if (nsamples >= (1 << GID_NBITS))
		return kssdhip_fail(KSSDHIP_EINVAL, "build_index: %d samples, at most %d", nsamples,
				    (1 << GID_NBITS) - 1);
	if (obj_bits < 1 || obj_bits > 32)
		return kssdhip_fail(KSSDHIP_EINVAL, "build_index: obj_bits %u out of range 1..32", obj_bits);
	if (skidx[0] != 0)
		return kssdhip_fail(KSSDHIP_EINVAL, "build_index: comblco.index must start at 0");
	for (int g = 0; g < nsamples; ++g)
		if (skidx[g + 1] < skidx[g])
			return kssdhip_fail(KSSDHIP_EINVAL, "build_index: comblco.index decreases at %d", g);
	const uint64_t n = skidx[nsamples];
	if (n && (!keys || !records))
		return kssdhip_fail(KSSDHIP_EINVAL, "build_index: null keys or records");
	if (n >= 0xffffffffull)
		return kssdhip_fail(KSSDHIP_EINVAL, "build_index: %llu records, at most 2^32 - 2", (unsigned long long)n);
	hipStream_t s = nullptr;
	if (hipStreamCreate(&s) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "build_index: stream creation failed");
	struct StreamGuard {
		hipStream_t s;
		~StreamGuard() { (void)hipStreamDestroy(s); }
	} guard{s};
	Mem m;
	uint64_t *d_keys, *d_skidx, *k0, *k1, *large, *over;
	uint32_t *v0, *v1, *hist, *sums, *bits;
	unsigned *flags;
	kssd_ctxgidobj_t *d_out;
	const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
	const uint64_t hn = (uint64_t)RS_BINS * ntiles;
	const uint64_t nsum = (hn + SCAN_BLOCK - 1) / SCAN_BLOCK;
	const uint64_t nwords = (n + 63) / 64 * 2; // whole 64-bit words (k_unique_small stores them)
	const unsigned large_cap = (unsigned)std::min<uint64_t>(n / (SMALL_GROUP + 1) + 1, 0xffffffffull);
	int rc;
	if ((rc = m.get(&d_keys, n)) || (rc = m.get(&d_skidx, (size_t)nsamples + 1)) || (rc = m.get(&k0, n)) ||
	    (rc = m.get(&k1, n)) || (rc = m.get(&v0, n)) || (rc = m.get(&v1, n)) || (rc = m.get(&hist, hn)) ||
	    (rc = m.get(&sums, nsum)) || (rc = m.get(&bits, nwords)) || (rc = m.get(&flags, 4)) ||
	    (rc = m.get(&large, large_cap)) || (rc = m.get(&over, large_cap)) || (rc = m.get(&d_out, n)))
		return rc;
	hipError_t e = copy_h2d(d_keys, keys, n * 8, s);
	if (e == hipSuccess)
		e = hipMemcpyAsync(d_skidx, skidx, ((size_t)nsamples + 1) * 8, hipMemcpyHostToDevice, s);
	if (e == hipSuccess)
		e = hipMemsetAsync(flags, 0, 16, s);
	if (e == hipSuccess)
		e = hipMemsetAsync(bits, 0, nwords * 4, s);
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "build_index: upload: %s", hipGetErrorString(e));
	if (nsamples > 0)
		hipLaunchKernelGGL(k_make_records, dim3(std::min(nsamples, 1 << 16)), dim3(256), 0, s, d_keys, d_skidx,
				   nsamples, obj_bits, k0, v0, flags);
	// stable sort by context: ctxgid bits 20..63
	if ((rc = radix_sort(k0, v0, k1, v1, n, GID_NBITS, 64, hist, sums, s)))
		return rc;
	if (n) {
		hipLaunchKernelGGL(k_unique_small, dim3(grid_n(n)), dim3(256), 0, s, k0, v0, n, bits, large, flags + 1,
				   large_cap);
		hipLaunchKernelGGL(k_pack, dim3(grid_n(n)), dim3(256), 0, s, k0, v0, n, d_out);
	}
	unsigned hflags[4] = {0, 0, 0, 0};
	e = hipMemcpyAsync(hflags, flags, 16, hipMemcpyDeviceToHost, s);
	if (e == hipSuccess)
		e = hipStreamSynchronize(s);
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "build_index: %s", hipGetErrorString(e));
	if (hflags[0])
		return kssdhip_fail(KSSDHIP_EINVAL,
				    "build_index: a sample's keys are not ascending, or a context exceeds 44 bits");
	const unsigned nlarge = hflags[1];
	if (nlarge > large_cap)
		return kssdhip_fail(KSSDHIP_EHIP, "build_index: large-group list overflow");
	if (nlarge) {
		hipLaunchKernelGGL(k_unique_large, dim3(std::min(nlarge, 65535u)), dim3(1024), 0, s, k0, v0, n, large,
				   nlarge, bits, over, flags + 2);
		e = hipMemcpyAsync(hflags, flags, 16, hipMemcpyDeviceToHost, s);
		if (e == hipSuccess)
			e = hipStreamSynchronize(s);
		if (e != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "build_index: %s", hipGetErrorString(e));
		const unsigned nover = hflags[2];
		if (nover > nlarge)
			return kssdhip_fail(KSSDHIP_EHIP, "build_index: giant-group list overflow");
		std::vector<uint64_t> heads(nover);
		if (nover && (e = hipMemcpy(heads.data(), over, nover * 8ull, hipMemcpyDeviceToHost)) != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "build_index: %s", hipGetErrorString(e));
		// groups past the LDS: (object, position) pairs radix-sorted in HBM;
		// k1/v1 (free after the sort) and d_keys (free after the records
		// were made) hold them
		for (const uint64_t b : heads) {
			hipLaunchKernelGGL(k_group_len, dim3(1), dim3(1), 0, s, k0, n, b, large);
			uint64_t g = 0;
			e = hipMemcpyAsync(&g, large, 8, hipMemcpyDeviceToHost, s);
			if (e == hipSuccess)
				e = hipStreamSynchronize(s);
			if (e != hipSuccess)
				return kssdhip_fail(KSSDHIP_EHIP, "build_index: %s", hipGetErrorString(e));
			uint64_t *pk0 = k1, *pk1 = d_keys;
			uint32_t *pv0 = v1, *pv1 = nullptr, *ghist, *gsums;
			const uint64_t gt = (g + RS_TILE - 1) / RS_TILE, ghn = (uint64_t)RS_BINS * gt;
			if ((rc = m.get(&pv1, g)) || (rc = m.get(&ghist, ghn)) ||
			    (rc = m.get(&gsums, (ghn + SCAN_BLOCK - 1) / SCAN_BLOCK)))
				return rc;
			hipLaunchKernelGGL(k_giant_pairs, dim3(grid_n(g)), dim3(256), 0, s, v0, b, g, pk0, pv0);
			if ((rc = radix_sort(pk0, pv0, pk1, pv1, g, 0, (int)obj_bits, ghist, gsums, s)))
				return rc;
			hipLaunchKernelGGL(k_giant_mark, dim3(grid_n(g)), dim3(256), 0, s, pk0, pv0, b, g, bits);
			e = hipStreamSynchronize(s);
			if (e != hipSuccess)
				return kssdhip_fail(KSSDHIP_EHIP, "build_index: %s", hipGetErrorString(e));
		}
	}
	e = copy_d2h(records, d_out, n * sizeof(kssd_ctxgidobj_t), s);
	if (e == hipSuccess && unique_bits) {
		std::vector<uint32_t> hb;
		e = hipStreamSynchronize(s);
		if (e == hipSuccess) {
			hb.resize(nwords);
			e = hipMemcpy(hb.data(), bits, nwords * 4, hipMemcpyDeviceToHost);
			if (e == hipSuccess)
				memcpy(unique_bits, hb.data(), (n + 7) / 8);
		}
	}
	if (e == hipSuccess)
		e = hipStreamSynchronize(s);
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "build_index: download: %s", hipGetErrorString(e));
	return 0;
}
