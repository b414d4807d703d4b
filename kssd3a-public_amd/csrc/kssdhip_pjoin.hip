// Partitioned merge-join: the batched counting path.
//
// The resident index is cut into P partitions of about `chunk` records whose
// boundaries never split the records of one context.  For a block of queries
// a split table S[q][p] (first key of query q whose context is >= the first
// context of partition p) is computed first (k_split).  Then one workgroup
// per (partition, slice of up to QB queries) stages the partition's records
// in LDS with 16-byte loads, builds a bucket table over the partition's
// context range in LDS, prefix-sums the slice's key counts in the partition,
// and deals the keys to its threads in equal contiguous runs.  Each key is
// matched against the staged records: bucket probe plus a short scan, walk
// of the context's per-genome runs, minimum per-base mismatch over the object
// cross product; hits are summed per (query, genome) cell in registers and
// flushed with one 64-bit atomic per cell and counter.  Consecutive
// partitions go to the same XCD so the query-key lines they share stay in
// that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_internal.h"

using namespace kssd;

namespace {

constexpr uint32_t LDS_SLACK = 64; // partitions may exceed `chunk` by a short context run
constexpr uint32_t STAGE_ALIGN = 4; // records are staged in aligned groups of 4 (16-byte loads)
constexpr int NUM_XCD = 8;

__global__ void k_part_cand(const uint64_t *__restrict__ ctxgid, uint64_t n, uint64_t chunk,
			    uint64_t ncand, uint32_t *cand)
{
	for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < ncand;
	     j += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t e = j * chunk;
		if (e >= n) {
			cand[j] = (uint32_t)n;
			continue;
		}
		// first record of the context that record e belongs to
		const uint64_t target = (ctxgid[e] >> GID_NBITS) << GID_NBITS;
		uint64_t lo = 0, hi = e;
		while (lo < hi) {
			const uint64_t mid = lo + ((hi - lo) >> 1);
			if (ctxgid[mid] < target)
				lo = mid + 1;
			else
				hi = mid;
		}
		cand[j] = (uint32_t)lo;
	}
}

__global__ void k_part_ctx(const uint64_t *__restrict__ ctxgid, const uint32_t *__restrict__ pstart,
			   uint32_t P, uint64_t *pctx)
{
	for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= P; p += gridDim.x * blockDim.x)
		pctx[p] = p < P ? (ctxgid[pstart[p]] >> GID_NBITS) : ~0ull;
}

constexpr int PLUT_BITS = 12;    // partition lookup table over the top bits of a 44-bit context
constexpr int PLUT_SHIFT = 44 - PLUT_BITS;
constexpr uint32_t PLUT_N = (1u << PLUT_BITS) + 1;

__global__ void k_plut(const uint64_t *__restrict__ pctx, uint32_t P, uint32_t *plut)
{
	for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < PLUT_N; b += gridDim.x * blockDim.x) {
		const uint64_t c = (uint64_t)b << PLUT_SHIFT;
		uint32_t lo = 0, hi = P; // pctx[P] = ~0 >= c
		while (lo < hi) {
			const uint32_t mid = lo + ((hi - lo) >> 1);
			if (pctx[mid] < c)
				lo = mid + 1;
			else
				hi = mid;
		}
		plut[b] = lo;
	}
}

constexpr uint32_t SPLIT_RUN = 2; // partition boundaries per k_split thread

struct SplitArgs {
	const uint64_t *keys;
	const uint64_t *skidx; // relative to keys, nq + 1
	int nq;
	unsigned obj_bits;
	const uint64_t *pctx;  // P + 1, pctx[P] = ~0
	uint32_t P;
	uint32_t *S; // [nq][P + 1]
};

// S[q][p] = first key of query q whose context is >= pctx[p] (the end of q
// when none).  One thread per (query, run of SPLIT_RUN boundaries): a binary
// search for the run's first boundary, then a galloping search forward from
// the previous boundary for each next one (a few keys apart), so the run's
// entries cost a handful of cached loads each and are written contiguously.
__global__ __launch_bounds__(256) void k_split(SplitArgs a)
{
	const uint32_t width = a.P + 1, runs = (width + SPLIT_RUN - 1) / SPLIT_RUN;
	// XCD-aware order: workgroup b runs on XCD b % 8; give each XCD a
	// contiguous run of workgroups (so of queries), whose binary searches
	// share the upper levels of their search paths in that XCD's L2.
	const uint32_t G = gridDim.x, b = blockIdx.x;
	const uint32_t xcd = b % NUM_XCD, nth = b / NUM_XCD, base = G / NUM_XCD, rem = G % NUM_XCD;
	const uint32_t wg = xcd * base + (xcd < rem ? xcd : rem) + nth;
	const uint64_t t = wg * (uint64_t)blockDim.x + threadIdx.x;
	if (t >= (uint64_t)a.nq * runs)
		return;
	const int q = (int)(t / runs);
	const uint32_t p0 = (uint32_t)(t - (uint64_t)q * runs) * SPLIT_RUN;
	const uint32_t p1 = p0 + SPLIT_RUN < width ? p0 + SPLIT_RUN : width;
	const uint64_t qb = a.skidx[q], qe = a.skidx[q + 1];
	const unsigned ob = a.obj_bits;
	uint32_t *row = a.S + (uint64_t)q * width;
	uint64_t c = a.pctx[p0];
	uint64_t lo = qb, hi = qe;
	while (lo < hi) {
		const uint64_t mid = lo + ((hi - lo) >> 1);
		if ((a.keys[mid] >> ob) < c)
			lo = mid + 1;
		else
			hi = mid;
	}
	uint64_t i = lo;
	row[p0] = (uint32_t)i;
	for (uint32_t p = p0 + 1; p < p1; ++p) {
		c = a.pctx[p];
		if (i < qe && (a.keys[i] >> ob) < c) {
			// keys[l] < c; gallop until keys[h] >= c or h reaches qe
			uint64_t l = i, step = 1, h = i + 1;
			while (h < qe && (a.keys[h] >> ob) < c) {
				l = h;
				step <<= 1;
				h = l + step;
			}
			if (h > qe)
				h = qe;
			uint64_t x = l + 1, y = h;
			while (x < y) {
				const uint64_t mid = x + ((y - x) >> 1);
				if ((a.keys[mid] >> ob) < c)
					x = mid + 1;
				else
					y = mid;
			}
			i = x;
		}
		row[p] = (uint32_t)i;
	}
}

// The split table from the keys' side, for blocks with many keys per table
// entry.  k_split_fill first sets every entry of row q to the end of query q
// (what entries past its last key, and every entry of an empty query, hold).
// Then, with f(i) = number of partition starts <= the context of key i, key
// i of query q is the first key of q at or past partitions [f(i-1), f(i))
// (the first key of q: [0, f(i))), so k_split_keys writes exactly those
// entries.  Each wave takes a contiguous run of 256-key tiles, four
// consecutive keys per lane (independent loads and searches in flight; f
// through the partition lookup table and a short search of pctx, both
// cache-resident).  Every entry holds a key index of its own query
// whatever the key order, as k_pjoin's key ranges require.
constexpr uint32_t SPLITK_WAVES = 4;
constexpr uint32_t SPLITK_KPL = 4;   // keys per lane per tile
constexpr uint32_t SPLITK_TILES = 8; // 256-key tiles per wave

__global__ void k_split_fill(const uint64_t *__restrict__ skidx, int nq, uint32_t width, uint32_t *S)
{
	const uint64_t n = (uint64_t)nq * width;
	for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n;
	     e += (uint64_t)gridDim.x * blockDim.x)
		S[e] = (uint32_t)skidx[e / width + 1];
}

struct SplitKArgs {
	SplitArgs s;
	const uint32_t *plut;
	uint64_t total;
};

__device__ __forceinline__ uint32_t parts_at_or_below(const uint64_t *__restrict__ pctx,
						      const uint32_t *__restrict__ plut, uint64_t c)
{
	if ((c >> PLUT_SHIFT) >= PLUT_N - 1) // past every 44-bit context: all P starts lie below
		return plut[PLUT_N - 1];
	const uint32_t b = (uint32_t)(c >> PLUT_SHIFT);
	uint32_t lo = plut[b], hi = plut[b + 1]; // pctx[p] < c for p < lo, > c for p >= hi
	while (lo < hi) {
		const uint32_t mid = lo + ((hi - lo) >> 1);
		if (pctx[mid] <= c)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

__global__ __launch_bounds__(SPLITK_WAVES * 64) void k_split_keys(SplitKArgs a)
{
	const uint32_t lane = threadIdx.x & 63;
	const uint64_t w = (uint64_t)blockIdx.x * SPLITK_WAVES + (threadIdx.x >> 6);
	const uint64_t tile = 64 * SPLITK_KPL, span = (uint64_t)SPLITK_TILES * tile;
	const uint64_t start = w * span;
	if (start >= a.total)
		return;
	const uint64_t end = start + span < a.total ? start + span : a.total;
	const uint32_t width = a.s.P + 1;
	const uint64_t *__restrict__ skidx = a.s.skidx;
	const unsigned ob = a.s.obj_bits;
	// the query of the run's first key by a 64-way search (each round the
	// lanes probe 64 evenly spaced offsets), and f of the key before it when
	// that key is in the same query (uniform across the wave)
	int lo = 0, hi = a.s.nq; // skidx[lo] <= start < skidx[hi]
	while (hi - lo > 1) {
		const int step = (hi - lo + 63) / 64;
		const int probe = lo + (int)lane * step;
		const bool le = probe < hi && skidx[probe] <= start;
		const uint64_t m = __ballot(le); // a prefix of the lanes
		const int k = 63 - __clzll((long long)m);
		lo = lo + k * step;
		hi = lo + step < hi ? lo + step : hi;
	}
	int q = lo;
	uint64_t qend = skidx[q + 1]; // first key past query q
	uint32_t prevf = start > skidx[q] ? parts_at_or_below(a.s.pctx, a.plut, a.s.keys[start - 1] >> ob) : 0;
	for (uint64_t base = start; base < end; base += tile) {
		// each lane takes SPLITK_KPL consecutive keys: independent loads and
		// searches, gaps between its own keys found locally
		const uint64_t i0 = base + (uint64_t)lane * SPLITK_KPL;
		uint32_t f[SPLITK_KPL];
		uint64_t c[SPLITK_KPL];
#pragma unroll
		for (int j = 0; j < (int)SPLITK_KPL; ++j)
			c[j] = i0 + j < end ? a.s.keys[i0 + j] >> ob : 0;
#pragma unroll
		for (int j = 0; j < (int)SPLITK_KPL; ++j)
			f[j] = i0 + j < end ? parts_at_or_below(a.s.pctx, a.plut, c[j]) : 0;
		const int nv = i0 >= end ? 0 : (int)(end - i0 < SPLITK_KPL ? end - i0 : SPLITK_KPL);
		uint32_t pf = __shfl_up(nv ? f[nv - 1] : 0u, 1);
		if (lane == 0)
			pf = prevf;
		int ql = q;
		uint64_t qe = qend, qb = 0;
		bool first = false;
#pragma unroll
		for (int j = 0; j < (int)SPLITK_KPL; ++j) {
			if (j >= nv)
				break;
			const uint64_t i = i0 + j;
			while (qe <= i) { // rarely: the key starts a later query
				qb = qe;
				++ql;
				qe = skidx[ql + 1];
				first = true;
			}
			const bool starts = first && i == qb;
			uint32_t *row = a.s.S + (uint64_t)ql * width;
			for (uint32_t p = starts || i == 0 ? 0 : pf; p < f[j]; ++p)
				row[p] = (uint32_t)i;
			pf = f[j];
		}
		// carry the last key's query and f into the next tile (a full tile
		// unless this is the run's last)
		prevf = __shfl(pf, 63);
		q = __shfl(ql, 63);
		qend = __shfl(qe, 63);
	}
}

// LDS (address space 3) pointers: accesses through them compile to ds_*
// instructions rather than flat ones.
#define LDS __attribute__((address_space(3)))

struct PJArgs {
	JoinArgs j;
	const uint32_t *pstart;
	const uint32_t *S; // [nq][P + 1]
	uint32_t P;
	uint32_t cap;   // records that fit the LDS stage
	uint32_t group; // consecutive partitions per workgroup
	uint32_t slots; // LDS query slots (0 unless group > 1)
	int bbits;      // log2 of the context buckets of a staged partition
};

// Per-query hit slots of a partition group in LDS: slot k belongs to query
// qb0 + k and holds the first genome flushed for it with the summed
// counters; hits on that query's other genomes go straight to the global
// counters.  With UNIQ a second set of nq slots holds the unique copies.
constexpr uint32_t SLOT_EMPTY = 0xffffffffu;
struct QuerySlots {
	LDS uint32_t *gid;
	LDS unsigned long long *c, *o;
};

// Adds (c, o) to query qi's slot if it holds (or can claim) genome gid.
__device__ __forceinline__ bool slot_add(const QuerySlots &t, uint32_t qi, uint32_t gid, unsigned long long c,
					 unsigned long long o)
{
	uint32_t g = t.gid[qi];
	if (g == SLOT_EMPTY) {
		uint32_t expected = SLOT_EMPTY;
		__hip_atomic_compare_exchange_strong(&t.gid[qi], &expected, gid, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
						     __HIP_MEMORY_SCOPE_WORKGROUP);
		g = expected == SLOT_EMPTY ? gid : expected;
	}
	if (g != gid)
		return false;
	__hip_atomic_fetch_add(&t.c[qi], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
	if (o)
		__hip_atomic_fetch_add(&t.o[qi], o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
	return true;
}

constexpr uint32_t QB = 1024; // queries per workgroup; grid.y covers the block's queries

// Context buckets of a staged partition: bucket b holds the records whose
// context c has (c - c0) >> shift == b; bt[b] = first record of bucket b.
#ifndef MAX_BUCKET_BITS
#define MAX_BUCKET_BITS 13
#endif
// dynamic LDS a k_pjoin workgroup may take (160 KB per CU, less the static part)
constexpr uint32_t PJ_LDS_MAX = 160 * 1024 - 256;
struct PartTable {
	const LDS uint16_t *bt;
	uint64_t c0, clast;
	int shift;
};

// In-place exclusive scan of v[0, n) in LDS; returns the total.  Every
// thread of the block calls it.  wsum: LDS scratch of blockDim/64 + 1 words.
__device__ uint32_t block_exclusive_scan(LDS uint32_t *v, uint32_t n, uint32_t *wsum)
{
	const uint32_t B = blockDim.x, t = threadIdx.x, nw = B / warpSize;
	const uint32_t per = (n + B - 1) / B;
	const uint32_t b0 = t * per < n ? t * per : n, b1 = b0 + per < n ? b0 + per : n;
	uint32_t s = 0;
	for (uint32_t k = b0; k < b1; ++k)
		s += v[k];
	const uint32_t lane = t & (warpSize - 1), w = t / warpSize;
	uint32_t x = s; // inclusive scan over the wave
	for (int d = 1; d < warpSize; d <<= 1) {
		const uint32_t y = __shfl_up(x, d);
		if (lane >= (uint32_t)d)
			x += y;
	}
	if (lane == warpSize - 1)
		wsum[w] = x;
	__syncthreads();
	if (t == 0) {
		uint32_t run = 0;
		for (uint32_t k = 0; k < nw; ++k) {
			const uint32_t c = wsum[k];
			wsum[k] = run;
			run += c;
		}
		wsum[nw] = run;
	}
	__syncthreads();
	uint32_t off = wsum[w] + x - s;
	for (uint32_t k = b0; k < b1; ++k) {
		const uint32_t c = v[k];
		v[k] = off;
		off += c;
	}
	const uint32_t total = wsum[nw];
	__syncthreads();
	return total;
}

// Per-thread accumulation of consecutive hits on one counter cell, flushed
// when the cell changes: into the workgroup's query slots when it has them
// (CT), else with one 64-bit atomic per counter.
template <bool CT>
struct CellAcc {
	uint64_t cell = ~0ull; // global cell index
	uint32_t qi = 0, gid = 0;
	unsigned long long c = 0, o = 0;
	__device__ __forceinline__ void add(uint64_t at, uint32_t q, uint32_t g, int m, const QuerySlots &t,
					    unsigned long long *ctx, unsigned long long *obj)
	{
		if (at != cell) {
			flush(t, ctx, obj);
			cell = at;
			qi = q;
			gid = g;
		}
		c += 1ull | ((unsigned long long)(m > 1) << 32);
		if (m)
			o += 1ull | ((unsigned long long)m << 32);
	}
	__device__ __forceinline__ void flush(const QuerySlots &t, unsigned long long *ctx, unsigned long long *obj)
	{
		if (cell == ~0ull)
			return;
		if (!CT || !slot_add(t, qi, gid, c, o)) {
			atomicAdd(&ctx[cell], c);
			if (o)
				atomicAdd(&obj[cell], o);
		}
		cell = ~0ull;
		c = o = 0;
	}
};

// Count one partition against the workgroup's queries.  pre[0, nqb] =
// exclusive prefix of the queries' key counts in this partition, start[k] =
// first key of query qb0 + k in it.  The K keys, in flattened (query, key)
// order, are dealt to the threads in equal contiguous runs, so every lane
// does the same number of lookups and a lane's keys mostly belong to one
// query; the first key of each context group does the work for the group
// (command_ani.h:506-575, :668-716), and hits are summed per (query, genome)
// cell in registers before one atomic per cell and counter.  UNIQ: also the
// unique-owner copies (command_ani.h:611-617); STAGED: rec/robj and the
// bucket table are in LDS, else rec/robj are the partition in HBM and the
// first record is found by binary search.
template <bool UNIQ, bool CT, bool STAGED, typename RecPtr, typename ObjPtr>
__device__ __forceinline__ void merge_keys(const PJArgs &a, RecPtr rec, ObjPtr robj, uint32_t cnt,
					   const PartTable &tab, const LDS uint32_t *pre, const LDS uint32_t *start,
					   int qb0, int nqb, uint32_t K, const QuerySlots &qs, const QuerySlots &us)
{
	const JoinArgs &j = a.j;
	const unsigned ob = j.obj_bits;
	const uint64_t objmask = ob >= 64 ? ~0ull : ((1ull << ob) - 1ull);
	const bool lower = (j.flags & KSSDHIP_LOWER) != 0;
	const bool ignore = (j.flags & KSSDHIP_IGNORECONFLICT) != 0;
	const uint32_t per = (K + blockDim.x - 1) / blockDim.x;
	const uint32_t f_begin = threadIdx.x * per;
	const uint32_t f_end = f_begin + per < K ? f_begin + per : K;
	if (f_begin >= K)
		return;
	int qi; // largest qi with pre[qi] <= f_begin: the query owning key f_begin
	{
		int lo = 0, hi = nqb;
		while (hi - lo > 1) {
			const int mid = (lo + hi) >> 1;
			if (pre[mid] <= f_begin)
				lo = mid;
			else
				hi = mid;
		}
		qi = lo;
	}
	CellAcc<CT> acc, uacc;
	uint64_t prev_key = 0, next_key = 0;
	uint32_t prev_idx = ~0u, next_idx = ~0u;
	uint32_t q_pre = pre[qi], q_next = pre[qi + 1], q_start = start[qi];
	const uint32_t step = j.vote_step;
	uint32_t q_first = step > 1 ? (uint32_t)j.skidx[qb0 + qi] : 0; // the query's first key
	for (uint32_t f = f_begin; f < f_end; ++f) {
		while (q_next <= f) {
			++qi;
			q_pre = q_next;
			q_next = pre[qi + 1];
			q_start = start[qi];
			if (step > 1)
				q_first = (uint32_t)j.skidx[qb0 + qi];
		}
		const uint32_t kidx = q_start + (f - q_pre);
		const uint32_t qend = q_start + (q_next - q_pre);
		if (step > 1 && (kidx - q_first) % step)
			continue; // not a sampled key
		const uint64_t key = kidx == next_idx ? next_key : j.keys[kidx];
		const uint64_t qctx = key >> ob;
		if (f != q_pre) {
			const uint64_t prev = kidx - 1 == prev_idx ? prev_key : j.keys[kidx - 1];
			if ((prev >> ob) == qctx) {
				prev_key = key;
				prev_idx = kidx;
				continue; // not the first key of its context group
			}
		}
		prev_key = key;
		prev_idx = kidx;
		uint32_t s_end = kidx + 1;
		if (s_end < qend) {
			next_key = j.keys[s_end];
			next_idx = s_end;
			if ((next_key >> ob) == qctx) {
				++s_end;
				while (s_end < qend && (j.keys[s_end] >> ob) == qctx)
					++s_end;
			}
		}
		// first record of the partition whose context is >= qctx
		uint32_t d;
		if (STAGED) {
			if (qctx > tab.clast || qctx < tab.c0) {
				d = cnt;
			} else {
				const uint32_t b = (uint32_t)((qctx - tab.c0) >> tab.shift);
				d = tab.bt[b];
				const uint32_t e = tab.bt[b + 1];
				while (d < e && (rec[d] >> GID_NBITS) < qctx)
					++d;
			}
		} else {
			const uint64_t target = qctx << GID_NBITS;
			uint32_t lo = 0, h2 = cnt;
			while (lo < h2) {
				const uint32_t mid = (lo + h2) >> 1;
				if (rec[mid] < target)
					lo = mid + 1;
				else
					h2 = mid;
			}
			d = lo;
		}
		if (j.vote_max_freq) { // skip contexts carried by more index records
			uint32_t x = d;
			while (x < cnt && x - d <= j.vote_max_freq && (rec[x] >> GID_NBITS) == qctx)
				++x;
			if (x - d > j.vote_max_freq)
				continue;
		}
		const int q = qb0 + qi;
		const int qgid = j.qry_gid_offset + q;
		const uint64_t row = (uint64_t)q * (uint64_t)j.ref_gnum;
		int owners = 0, sole_diff = 0;
		uint32_t sole_gid = 0;
		bool sole_allowed = false;
		while (d < cnt) {
			const uint64_t cg = rec[d];
			if ((cg >> GID_NBITS) != qctx)
				break;
			const uint32_t gid = (uint32_t)(cg & GID_MASK);
			const uint32_t rb = d;
			do {
				++d;
			} while (d < cnt && rec[d] == cg);
			const uint32_t re = d;
			if (lower && (int)gid >= qgid) {
				if ((int)gid > qgid)
					break;
				continue;
			}
			if (UNIQ) {
				++owners;
				sole_gid = gid;
			}
			if (ignore && re - rb > 1)
				continue;
			int m = MIN_INIT;
			for (uint32_t x = kidx; x < s_end && m; ++x) {
				const uint32_t oq = (uint32_t)((x == kidx ? key : j.keys[x]) & objmask);
				for (uint32_t r = rb; r < re; ++r) {
					const uint32_t diff = oq ^ robj[r];
					if (diff == 0) {
						m = 0;
						break;
					}
					const int c = dna_popcount_dev(diff);
					m = c < m ? c : m;
				}
			}
			acc.add(row + gid, (uint32_t)qi, gid, m, qs, j.ctx, j.objc);
			if (UNIQ) {
				sole_diff = m;
				sole_allowed = true;
			}
		}
		if (UNIQ && owners == 1 && sole_allowed)
			uacc.add(row + sole_gid, (uint32_t)qi, sole_gid, sole_diff, us, j.uctx, j.uobj);
	}
	acc.flush(qs, j.ctx, j.objc);
	if (UNIQ)
		uacc.flush(us, j.uctx, j.uobj);
}

// Byte offsets of k_pjoin's LDS arrays for a stage of `cap` records (a
// multiple of 4), 2^bbits buckets, qcap queries and `slots` query slots:
// [cap] u64 ctxgid | [cap] u32 obj | [2^bbits + 2] u16 bucket table |
// [qcap + 1] u32 prefix | [qcap] u32 starts | [slots] u32 genome (8-byte
// aligned) | [slots] u64 ctx counters | [slots] u64 obj counters.
struct PJLayout {
	uint32_t bt, pre, start, sg, sc, so, bytes;
	__host__ __device__ PJLayout(uint32_t cap, int bbits, uint32_t qcap, uint32_t slots)
	{
		bt = cap * 12;
		pre = bt + ((1u << bbits) + 2) * 2;
		start = pre + (qcap + 1) * 4;
		sg = (start + qcap * 4 + 7) & ~7u;
		sc = (sg + slots * 4 + 7) & ~7u;
		so = sc + slots * 8;
		bytes = so + slots * 8;
	}
};

template <bool UNIQ, bool CT>
__global__ void k_pjoin(PJArgs a)
{
	// LDS arrays as PJLayout places them, qcap = min(nq, QB)
	extern __shared__ uint64_t lds_dyn[];
	__shared__ uint32_t wsum[1024 / 64 + 1];
	LDS char *const lds = (LDS char *)lds_dyn;
	const uint32_t qcap = (uint32_t)a.j.nq < QB ? (uint32_t)a.j.nq : QB;
	const PJLayout lay(a.cap, a.bbits, qcap, a.slots);
	LDS uint64_t *const lds_rec = (LDS uint64_t *)lds;
	LDS uint32_t *const lds_obj = (LDS uint32_t *)(lds + a.cap * 8);
	LDS uint16_t *const lds_bt = (LDS uint16_t *)(lds + lay.bt);
	LDS uint32_t *const lds_pre = (LDS uint32_t *)(lds + lay.pre);
	LDS uint32_t *const lds_start = (LDS uint32_t *)(lds + lay.start);
	QuerySlots qs, us; // us: the unique copies' slots, after the nq plain ones
	qs.gid = (LDS uint32_t *)(lds + lay.sg);
	qs.c = (LDS unsigned long long *)(lds + lay.sc);
	qs.o = (LDS unsigned long long *)(lds + lay.so);
	us.gid = qs.gid + qcap;
	us.c = qs.c + qcap;
	us.o = qs.o + qcap;
	// XCD-aware order: workgroup b runs on XCD b % 8; give each XCD a
	// contiguous run of partition groups.
	const uint32_t G = gridDim.x, b = blockIdx.x;
	const uint32_t x = b % NUM_XCD, i = b / NUM_XCD, base = G / NUM_XCD, rem = G % NUM_XCD;
	const uint32_t pg = x * base + (x < rem ? x : rem) + i;
	const uint32_t p_begin = pg * a.group;
	if (p_begin >= a.P)
		return;
	const uint32_t p_end = p_begin + a.group < a.P ? p_begin + a.group : a.P;
	const int qb0 = (int)(blockIdx.y * QB);
	const int nqb = a.j.nq - qb0 < (int)QB ? a.j.nq - qb0 : (int)QB;
	if (CT) {
		for (uint32_t k = threadIdx.x; k < a.slots; k += blockDim.x) {
			qs.gid[k] = SLOT_EMPTY;
			qs.c[k] = 0;
			qs.o[k] = 0;
		}
	}
	const uint64_t width = (uint64_t)a.P + 1;
	for (uint32_t p = p_begin; p < p_end; ++p) {
		__syncthreads(); // the previous partition is done with the shared arrays
		// this partition's key range of each of the workgroup's queries
		for (int k = threadIdx.x; k < nqb; k += blockDim.x) {
			const uint64_t r = (uint64_t)(qb0 + k) * width + p;
			const uint32_t s0 = a.S[r], s1 = a.S[r + 1];
			lds_start[k] = s0;
			lds_pre[k] = s1 > s0 ? s1 - s0 : 0; // (unsorted keys must not underflow)
		}
		const uint32_t rb = a.pstart[p], re = a.pstart[p + 1], cnt = re - rb;
		const uint32_t rb_al = rb & ~(STAGE_ALIGN - 1), re_al = (re + STAGE_ALIGN - 1) & ~(STAGE_ALIGN - 1);
		const bool staged = re_al - rb_al <= a.cap;
		if (staged) {
			// stage [rb_al, re_al) with 16-byte loads; the index arrays are
			// padded past n, so the last group stays in bounds
			const uint32_t groups = (re_al - rb_al) / STAGE_ALIGN;
			const uint4 *g_rec = (const uint4 *)(a.j.ctxgid + rb_al);
			const uint4 *g_obj = (const uint4 *)(a.j.robj + rb_al);
			uint4 *l_rec = (uint4 *)lds_dyn; // generic pointers into the LDS stage
			uint4 *l_obj = (uint4 *)(lds_dyn + a.cap);
			for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
				const uint4 r0 = g_rec[2 * g], r1 = g_rec[2 * g + 1], o = g_obj[g];
				l_rec[2 * g] = r0;
				l_rec[2 * g + 1] = r1;
				l_obj[g] = o;
			}
		}
		__syncthreads();
		const uint32_t K = block_exclusive_scan(lds_pre, (uint32_t)nqb, wsum);
		if (threadIdx.x == 0)
			lds_pre[nqb] = K;
		if (!staged) {
			__syncthreads();
			// a context run longer than the stage: merge straight from HBM
			merge_keys<UNIQ, CT, false>(a, a.j.ctxgid + rb, a.j.robj + rb, cnt, PartTable{}, lds_pre, lds_start,
						    qb0, nqb, K, qs, us);
			continue;
		}
		const uint32_t off = rb - rb_al;
		const LDS uint64_t *srec = lds_rec + off;
		const LDS uint32_t *sobj = lds_obj + off;
		// bucket table over [c0, clast]: at most 2^bbits buckets
		PartTable tab;
		tab.bt = lds_bt;
		tab.c0 = srec[0] >> GID_NBITS;
		tab.clast = srec[cnt - 1] >> GID_NBITS;
		const uint64_t span = tab.clast - tab.c0; // buckets = (span >> shift) + 1
		const int bits = span ? 64 - __clzll((long long)span) : 0;
		tab.shift = bits > a.bbits ? bits - a.bbits : 0;
		for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
			const int bk = (int)(((srec[k] >> GID_NBITS) - tab.c0) >> tab.shift);
			const int bp = k ? (int)(((srec[k - 1] >> GID_NBITS) - tab.c0) >> tab.shift) : -1;
			for (int bb = bp + 1; bb <= bk; ++bb)
				lds_bt[bb] = (uint16_t)k;
		}
		if (threadIdx.x == 0)
			lds_bt[(span >> tab.shift) + 1] = (uint16_t)cnt;
		__syncthreads();
		merge_keys<UNIQ, CT, true>(a, srec, sobj, cnt, tab, lds_pre, lds_start, qb0, nqb, K, qs, us);
	}
	if (!CT)
		return;
	__syncthreads();
	// each claimed slot added to the outputs once
	const JoinArgs &j = a.j;
	for (uint32_t k = threadIdx.x; k < a.slots; k += blockDim.x) {
		const uint32_t g = qs.gid[k];
		if (g == SLOT_EMPTY)
			continue;
		const bool u = k >= qcap;
		const uint64_t cell = (uint64_t)(qb0 + (int)(u ? k - qcap : k)) * (uint64_t)j.ref_gnum + g;
		atomicAdd(&(u ? j.uctx : j.ctx)[cell], qs.c[k]);
		if (qs.o[k])
			atomicAdd(&(u ? j.uobj : j.objc)[cell], qs.o[k]);
	}
}

unsigned grid_cap(uint64_t n, unsigned block)
{
	uint64_t g = (n + block - 1) / block;
	if (g < 1)
		g = 1;
	if (g > (1u << 20))
		g = 1u << 20;
	return (unsigned)g;
}

} // namespace

void kssd_free_partitions(kssdhip_index *ix)
{
	if (ix->d_pstart)
		(void)hipFree(ix->d_pstart);
	if (ix->d_pctx)
		(void)hipFree(ix->d_pctx);
	if (ix->d_plut)
		(void)hipFree(ix->d_plut);
	ix->d_pstart = nullptr;
	ix->d_pctx = nullptr;
	ix->d_plut = nullptr;
	ix->P = 0;
	ix->max_part = 0;
}

int kssd_build_partitions(kssdhip_index *ix, int chunk)
{
	if (chunk < 64 || chunk > 12288)
		return kssdhip_fail(KSSDHIP_EINVAL, "chunk_records %d out of range 64..12288", chunk);
	kssd_free_partitions(ix);
	ix->chunk = chunk;
	if (ix->n == 0)
		return 0;
	const uint64_t ncand = (ix->n + chunk - 1) / chunk;
	uint32_t *d_cand = nullptr;
	std::vector<uint32_t> cand(ncand);
	if (hipMalloc(&d_cand, ncand * 4) != hipSuccess)
		return kssdhip_fail(KSSDHIP_ENOMEM, "partition candidates: device allocation failed");
	hipLaunchKernelGGL(k_part_cand, dim3(grid_cap(ncand, 256)), dim3(256), 0, ix->stream, ix->d_ctxgid,
			   ix->n, (uint64_t)chunk, ncand, d_cand);
	hipError_t e = hipMemcpyAsync(cand.data(), d_cand, ncand * 4, hipMemcpyDeviceToHost, ix->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(ix->stream);
	(void)hipFree(d_cand);
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "partition candidates: %s", hipGetErrorString(e));
	std::vector<uint32_t> pstart;
	pstart.reserve(ncand + 1);
	for (uint64_t j = 0; j < ncand; ++j)
		if (pstart.empty() || cand[j] != pstart.back())
			pstart.push_back(cand[j]);
	if (pstart.empty() || pstart[0] != 0)
		pstart.insert(pstart.begin(), 0u);
	if (pstart.back() != (uint32_t)ix->n)
		pstart.push_back((uint32_t)ix->n);
	const uint32_t P = (uint32_t)pstart.size() - 1;
	uint32_t mx = 0;
	for (uint32_t p = 0; p < P; ++p)
		mx = std::max(mx, pstart[p + 1] - pstart[p]);
	if (hipMalloc(&ix->d_pstart, (P + 1) * 4) != hipSuccess ||
	    hipMalloc(&ix->d_pctx, (P + 1) * 8) != hipSuccess || hipMalloc(&ix->d_plut, PLUT_N * 4) != hipSuccess) {
		kssd_free_partitions(ix);
		return kssdhip_fail(KSSDHIP_ENOMEM, "partition table: device allocation failed");
	}
	e = hipMemcpyAsync(ix->d_pstart, pstart.data(), (P + 1) * 4, hipMemcpyHostToDevice, ix->stream);
	if (e == hipSuccess) {
		hipLaunchKernelGGL(k_part_ctx, dim3(grid_cap(P + 1, 256)), dim3(256), 0, ix->stream,
				   ix->d_ctxgid, ix->d_pstart, P, ix->d_pctx);
		hipLaunchKernelGGL(k_plut, dim3(grid_cap(PLUT_N, 256)), dim3(256), 0, ix->stream, ix->d_pctx, P,
				   ix->d_plut);
		e = hipGetLastError();
	}
	if (e == hipSuccess)
		e = hipStreamSynchronize(ix->stream);
	if (e != hipSuccess) {
		kssd_free_partitions(ix);
		return kssdhip_fail(KSSDHIP_EHIP, "partition table: %s", hipGetErrorString(e));
	}
	ix->P = P;
	ix->max_part = mx;
	return 0;
}

int kssd_launch_pjoin(kssdhip_index *ix, const JoinArgs &ja, hipStream_t s, DevBuf *split, bool *used)
{
	*used = false;
	if (ix->P == 0 || ja.total >= 0xffffffffull)
		return 0;
	const bool want = ix->mode == 2 ||
			  (ix->mode == 0 && (double)ja.total >= ix->auto_min_keys * (double)ix->P);
	if (!want)
		return 0;
	const uint64_t sbytes = ((uint64_t)ix->P + 1) * (uint64_t)ja.nq * 4;
	if (ix->block > 1024)
		return 0;
	int rc = kssd_reserve_for_stream(split, sbytes, s, "split table");
	if (rc)
		return rc;
	uint32_t *S = (uint32_t *)split->ptr;
	SplitArgs sa;
	sa.keys = ja.keys;
	sa.skidx = ja.skidx;
	sa.nq = ja.nq;
	sa.obj_bits = ja.obj_bits;
	sa.pctx = ix->d_pctx;
	sa.P = ix->P;
	sa.S = S;
	const uint64_t entries = ((uint64_t)ix->P + 1) * (uint64_t)ja.nq;
	if (ja.total * 8 >= entries) {
		// keys' side: every row set to its query's end, then a run of key tiles per wave
		hipLaunchKernelGGL(k_split_fill, dim3(grid_cap(entries, 256)), dim3(256), 0, s, ja.skidx, ja.nq, ix->P + 1,
				   S);
		SplitKArgs ka;
		ka.s = sa;
		ka.plut = ix->d_plut;
		ka.total = ja.total;
		const uint64_t span = (uint64_t)SPLITK_TILES * 64 * SPLITK_KPL;
		const uint64_t waves = (ja.total + span - 1) / span;
		hipLaunchKernelGGL(k_split_keys, dim3((unsigned)((waves + SPLITK_WAVES - 1) / SPLITK_WAVES)),
				   dim3(SPLITK_WAVES * 64), 0, s, ka);
	} else {
		const uint64_t runs = ((uint64_t)ix->P + SPLIT_RUN) / SPLIT_RUN;
		const uint64_t nthreads = (uint64_t)ja.nq * runs;
		hipLaunchKernelGGL(k_split, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, s, sa);
	}
	PJArgs a;
	a.j = ja;
	a.pstart = ix->d_pstart;
	a.S = S;
	a.P = ix->P;
	a.cap = ((uint32_t)ix->chunk + LDS_SLACK + STAGE_ALIGN - 1) & ~(STAGE_ALIGN - 1);
	a.group = (uint32_t)ix->group;
	const uint32_t qslices = (uint32_t)((ja.nq + QB - 1) / QB);
	const uint32_t qb = (uint32_t)ja.nq < QB ? (uint32_t)ja.nq : QB;
	const bool uniq = (ja.flags & KSSDHIP_UNIQUE) && !(ja.flags & KSSDHIP_LOWER);
	a.slots = ix->group > 1 ? qb * (uniq ? 2 : 1) : 0;
	// about one staged record per bucket, as far as the LDS allows without
	// fitting fewer workgroups per CU than the smallest table does (a CU has
	// 160 KB of LDS; the kernel's static LDS is under 256 bytes)
	auto per_cu = [](size_t bytes) { return (uint32_t)((160u * 1024u) / (bytes + 256)); };
	a.bbits = 1;
	while (a.bbits < 10 && (2u << a.bbits) <= a.cap)
		++a.bbits;
	const uint32_t wgs = per_cu(PJLayout(a.cap, a.bbits, qb, a.slots).bytes);
	while (a.bbits < MAX_BUCKET_BITS && (2u << a.bbits) <= a.cap &&
	       PJLayout(a.cap, a.bbits + 1, qb, a.slots).bytes <= PJ_LDS_MAX &&
	       per_cu(PJLayout(a.cap, a.bbits + 1, qb, a.slots).bytes) >= wgs)
		++a.bbits;
	const size_t lds = PJLayout(a.cap, a.bbits, qb, a.slots).bytes;
	if (lds > PJ_LDS_MAX)
		return kssdhip_fail(KSSDHIP_EINVAL, "pjoin: LDS %zu bytes over the limit", lds);
	const bool ctab = a.slots != 0;
	const void *fn = uniq ? (ctab ? (const void *)k_pjoin<true, true> : (const void *)k_pjoin<true, false>)
			      : (ctab ? (const void *)k_pjoin<false, true> : (const void *)k_pjoin<false, false>);
	static thread_local size_t lds_set[4] = {0, 0, 0, 0};
	const int fi = (uniq ? 2 : 0) + (ctab ? 1 : 0);
	if (lds > 65536 && lds_set[fi] < lds) {
		hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
		if (e != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "pjoin: LDS %zu bytes: %s", lds, hipGetErrorString(e));
		lds_set[fi] = lds;
	}
	const uint32_t ngroups = (ix->P + a.group - 1) / a.group;
	const dim3 grid(ngroups, qslices), block(ix->block);
	if (uniq && ctab)
		hipLaunchKernelGGL((k_pjoin<true, true>), grid, block, lds, s, a);
	else if (uniq)
		hipLaunchKernelGGL((k_pjoin<true, false>), grid, block, lds, s, a);
	else if (ctab)
		hipLaunchKernelGGL((k_pjoin<false, true>), grid, block, lds, s, a);
	else
		hipLaunchKernelGGL((k_pjoin<false, false>), grid, block, lds, s, a);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "pjoin launch: %s", hipGetErrorString(e));
	*used = true;
	return 0;
}

extern "C" int kssdhip_index_set_tiling(kssdhip_index *ix, const kssdhip_tiling_t *t)
{
	if (!ix || !t)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: null argument");
	if (t->mode < 0 || t->mode > 3)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: mode %d", t->mode);
	if (t->query_threads && (t->query_threads < 64 || t->query_threads > 512 || t->query_threads % 64))
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: query_threads %d", t->query_threads);
	if (t->query_slots && (t->query_slots < 256 || t->query_slots > 8192 || (t->query_slots & (t->query_slots - 1))))
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: query_slots %d (a power of two in 256..8192)",
				    t->query_slots);
	if (t->block_threads < 64 || t->block_threads > 1024 || t->block_threads % 64)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: block_threads %d", t->block_threads);
	if (t->auto_min_keys < 0)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: auto_min_keys < 0");
	if (t->group_partitions < 1 || t->group_partitions > 64)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: group_partitions %d out of range 1..64",
				    t->group_partitions);
	if (t->group_partitions > 1 && t->chunk_records > 6144)
		return kssdhip_fail(KSSDHIP_EINVAL, "set_tiling: group_partitions > 1 needs chunk_records <= 6144");
	if (hipSetDevice(ix->device) != hipSuccess || hipStreamSynchronize(ix->stream) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "set_tiling: device not ready");
	if (t->chunk_records != ix->chunk) {
		int rc = kssd_build_partitions(ix, t->chunk_records);
		if (rc)
			return rc;
	}
	ix->mode = t->mode;
	ix->block = t->block_threads;
	ix->auto_min_keys = t->auto_min_keys;
	ix->group = t->group_partitions;
	ix->qthreads = t->query_threads;
	ix->qslots = t->query_slots;
	return kssd_qjoin_prepare(ix);
}

extern "C" int kssdhip_index_get_tiling(const kssdhip_index *ix, kssdhip_tiling_t *t)
{
	if (!ix || !t)
		return kssdhip_fail(KSSDHIP_EINVAL, "get_tiling: null argument");
	t->mode = ix->mode;
	t->chunk_records = ix->chunk;
	t->block_threads = ix->block;
	t->auto_min_keys = ix->auto_min_keys;
	t->group_partitions = ix->group;
	t->query_threads = ix->qthreads;
	t->query_slots = ix->qslots;
	return 0;
}
