// Query-major join: the default counting path of an index with a record
// table, and the device-side pass gate of the index path.
//
// The reference counts one query per OpenMP iteration: every context group
// of the query finds its first index record through the fenceposts, walks
// the context's per-genome runs and adds into the query's own row of the
// dense counters (command_ani.h:481-575).  Here a persistent workgroup takes
// whole queries from a device work queue (largest first), or, for blocks of
// few queries, key-range items whose sizes halve round by round.  Each wave
// takes 64 consecutive keys; the first key of each context group looks its
// context up in the record table, where one 128-byte home line holds the
// context's records (or a pointer into the sorted index for long ones): the
// wave's 64 lines are loaded eight lanes to a line into LDS rows, each lane
// compares its row's slots with its context's remainder, and the wave walks
// the found records one per lane, adding each genome run's counters into an
// LDS table of the query (one cell per genome when the reference set is
// small enough and no slot count is asked for, else an open-addressed table
// keyed by genome; a genome that finds no slot goes to the workgroup's
// spill row in HBM).  When the query (or item) is done the workgroup writes
// its dense row once, adds it into the outputs, or applies the pass gate of
// the ranking loop (command_ani.h:585-606) to the touched pairs and appends
// the passing ones to a compact list.  No counter is updated per match with
// a global atomic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_epi.h"
#include "kssdhip_internal.h"

using namespace kssd;

#define LDS __attribute__((address_space(3)))

namespace {

constexpr uint32_t EMPTY = 0xffffffffu;
constexpr uint32_t NONE = 0xffffffffu;
constexpr int MAX_PROBES = 32;
constexpr uint32_t DIRECT_LDS_MAX = 64u * 1024u; // direct per-genome rows up to this many LDS bytes
constexpr int DEFAULT_THREADS = 512;
constexpr int QJ_MAX_THREADS = 512; // workgroup size bound of k_qjoin
#ifndef QJ_WAVES_PER_SIMD
#define QJ_WAVES_PER_SIMD 4 // register budget of k_qjoin: at most 128 VGPRs, four waves per SIMD
#endif
constexpr uint32_t DEFAULT_SLOTS = 512;
constexpr uint64_t QJ_MIN_CHUNK = 1024; // least keys per work item when queries are cut into key ranges
// workspace: [0, WS_HEAD) work counter and the order's histogram and
// cursors (65 words), then the spill rows, then the work order (nq words,
// WS_ALIGN-rounded)
constexpr size_t WS_HEAD = 512;
constexpr uint32_t WS_HEAD_WORDS = 65;
constexpr size_t WS_ALIGN = 256;

enum { QJ_DENSE = 0, QJ_ACCUM = 1, QJ_SELECT = 2 };

// The record table: the index's records bucketed by context, so that one
// 128-byte memory line answers a lookup.  A context c of w = ctx_width bits
// is mapped by a bijection p = br_perm(c) of w-bit values; the top bb bits
// of p name c's home line, the low w - bb bits (the remainder) are stored
// in each of c's slots.  As (home, remainder) determines c, a slot match is
// exact: there is no fingerprint to confirm.  A line holds 16 u64 slots
// (rs = max(32, 20 + obj_bits): the remainder lies in the slot's high word,
// so that a lookup compares high words only):
//   record  remainder << rs | gid << obj_bits | obj  (< 2^63)
//   long    remainder << rs | GID_MASK << obj_bits, then a
//           payload slot 3 << 62 | count << 32 | first record: the
//           context's records in the sorted index (contexts of more than
//           BR_LMAX records)
//   cont    1 << 63 | line (slot 15 only): the home's contexts continue in
//           that overflow line
//   empty   all ones
// A context's slots are contiguous and never split between lines; its
// records are in index order ((gid, obj) ascending).  bb >= 21 keeps every
// record and header below 2^63 for object widths up to 32 bits (rs plus the
// remainder's w - bb bits: at most 84 - bb), and gid GID_MASK is reserved
// for headers, so
// an index of 2^20 genomes has no table.  Home lines are sized for about 3
// slots each, so few homes need an overflow line (0.1 % of the lines at
// 171.7M records): a context absent from the index costs one line, a
// present one the same line (plus its run in the sorted index when long).
constexpr int BR_SLOTS = 16;
constexpr uint32_t BR_LMAX = 12;
constexpr int BR_MIN_BITS = 21;
constexpr unsigned long long BR_EMPTY = ~0ull;
constexpr unsigned long long BR_HIGH = 1ull << 63;
constexpr unsigned long long BR_PAYLOAD = 3ull << 62;
constexpr uint64_t BR_MAX_LONG = (1ull << 30) - 1; // records of one long context

__host__ __device__ __forceinline__ int br_ctx_width(unsigned obj_bits)
{
	return obj_bits >= 20 ? 64 - (int)obj_bits : 44;
}

// bit position of the remainder in record and header slots
__host__ __device__ __forceinline__ int br_rem_shift(unsigned obj_bits)
{
	return GID_NBITS + (int)obj_bits > 32 ? GID_NBITS + (int)obj_bits : 32;
}

// A bijection of w-bit values (1 <= w <= 44): an odd multiplier modulo
// 2^w, then an xor-shift that folds the well-mixed high bits into the low
// ones; each step is invertible.
__host__ __device__ __forceinline__ uint64_t br_perm(uint64_t x, int w)
{
	x = (x * 0x9e3779b97f4a7c15ull) & ((1ull << w) - 1ull);
	return x ^ (x >> ((w + 1) / 2));
}

__device__ __forceinline__ bool run_start(const uint64_t *__restrict__ ctxgid, uint64_t i)
{
	return i == 0 || (ctxgid[i] >> GID_NBITS) != (ctxgid[i - 1] >> GID_NBITS);
}

struct BrGeom {
	int w, bb;          // context width, home bits
	unsigned ob;        // object bits
	uint64_t homes;     // 2^bb
};

__device__ __forceinline__ uint64_t br_home(const BrGeom &g, uint64_t c)
{
	return br_perm(c, g.w) >> (g.w - g.bb);
}

// records of the context whose run starts at i (until the context changes)
__device__ __forceinline__ uint64_t br_run_len(const uint64_t *__restrict__ ctxgid, uint64_t n, uint64_t i)
{
	const uint64_t c = ctxgid[i] >> GID_NBITS;
	uint64_t e = i + 1;
	while (e < n && (ctxgid[e] >> GID_NBITS) == c)
		++e;
	return e - i;
}

__device__ __forceinline__ uint32_t br_need(uint64_t len)
{
	return len <= BR_LMAX ? (uint32_t)len : 2u;
}

// pass 1: contexts per home line (contexts wider than w cannot match a query
// and are left out)
__global__ void k_br_count(const uint64_t *__restrict__ ctxgid, uint64_t n, BrGeom g, uint32_t *cnt)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
		if (!run_start(ctxgid, i))
			continue;
		const uint64_t c = ctxgid[i] >> GID_NBITS;
		if (c >> g.w)
			continue;
		atomicAdd(&cnt[br_home(g, c)], 1u);
	}
}

// pass 2: each context's first record into its home's list (off: exclusive
// scan of the counts; cur: zeroed cursors)
__global__ void k_br_scatter(const uint64_t *__restrict__ ctxgid, uint64_t n, BrGeom g, const uint32_t *off,
			     uint32_t *cur, uint32_t *list)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
		if (!run_start(ctxgid, i))
			continue;
		const uint64_t c = ctxgid[i] >> GID_NBITS;
		if (c >> g.w)
			continue;
		const uint64_t h = br_home(g, c);
		list[off[h] + atomicAdd(&cur[h], 1u)] = (uint32_t)i;
	}
}

// pass 3 (count) and 5 (write), one thread per home: its contexts in index
// order (the list segment sorted in place), laid out line by line; a line
// keeps slot 15 for the continuation while more than 16 slots remain.
// Count: overflow lines per home into nov.  Write: the home line and its
// overflow lines from ovf_base + ovoff[h] on (all lines prefilled empty).
template <bool WRITE>
__global__ void k_br_place(const uint64_t *__restrict__ ctxgid, const uint32_t *__restrict__ robj, uint64_t n,
			   BrGeom g, const uint32_t *__restrict__ off, const uint32_t *__restrict__ cnt, uint32_t *list,
			   uint32_t *nov, const uint32_t *__restrict__ ovoff, uint64_t ovf_base,
			   unsigned long long *lines)
{
	const uint64_t objmask = g.ob >= 32 ? 0xffffffffull : ((1ull << g.ob) - 1ull);
	const int rs = br_rem_shift(g.ob);
	const uint64_t remmask = (1ull << (g.w - g.bb)) - 1ull;
	for (uint64_t h = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; h < g.homes;
	     h += (uint64_t)gridDim.x * blockDim.x) {
		const uint32_t k = cnt[h];
		if (!k) {
			if (!WRITE)
				nov[h] = 0;
			continue;
		}
		uint32_t *ls = list + off[h];
		if (!WRITE) { // index order (insertion sort: homes hold a few contexts)
			for (uint32_t x = 1; x < k; ++x) {
				const uint32_t v = ls[x];
				uint32_t y = x;
				while (y > 0 && ls[y - 1] > v) {
					ls[y] = ls[y - 1];
					--y;
				}
				ls[y] = v;
			}
		}
		uint64_t remaining = 0;
		for (uint32_t x = 0; x < k; ++x)
			remaining += br_need(br_run_len(ctxgid, n, ls[x]));
		uint64_t line = h;
		uint32_t used = 0, slot = 0;
		uint32_t cap = remaining <= BR_SLOTS ? BR_SLOTS : BR_SLOTS - 1;
		for (uint32_t x = 0; x < k; ++x) {
			const uint64_t i = ls[x];
			const uint64_t len = br_run_len(ctxgid, n, i);
			const uint32_t need = br_need(len);
			if (slot + need > cap) { // the rest goes to a new overflow line
				const uint64_t next = ovf_base + (WRITE ? ovoff[h] : 0) + used;
				if (WRITE)
					lines[line * BR_SLOTS + (BR_SLOTS - 1)] = BR_HIGH | next;
				++used;
				line = next;
				slot = 0;
				cap = remaining <= BR_SLOTS ? BR_SLOTS : BR_SLOTS - 1;
			}
			remaining -= need;
			if (WRITE) {
				const uint64_t c = ctxgid[i] >> GID_NBITS;
				const uint64_t head = (br_perm(c, g.w) & remmask) << rs;
				unsigned long long *dst = lines + line * BR_SLOTS + slot;
				if (len <= BR_LMAX) {
					for (uint64_t y = 0; y < len; ++y) {
						const uint64_t gid = ctxgid[i + y] & GID_MASK;
						dst[y] = head | (gid << g.ob) | ((uint64_t)robj[i + y] & objmask);
					}
				} else {
					dst[0] = head | (GID_MASK << g.ob);
					dst[1] = BR_PAYLOAD | (len << 32) | i;
				}
			}
			slot += need;
		}
		if (!WRITE)
			nov[h] = used;
	}
}

// longest context run (a long context's count must fit its payload field)
__global__ __launch_bounds__(256) void k_br_longest(const uint64_t *__restrict__ ctxgid, uint64_t n,
						    unsigned long long *longest)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
		if (!run_start(ctxgid, i))
			continue;
		const uint64_t len = br_run_len(ctxgid, n, i);
		if (len > BR_LMAX)
			atomicMax(longest, (unsigned long long)len);
	}
}

// Work order of a block: queries grouped by the bit length of their key
// count, longest first (so the largest queries do not end the launch
// alone).  head[1 .. 32] histogram, head[33 .. 64] cursors (zeroed before).
__device__ __forceinline__ uint32_t size_class(const uint64_t *skidx, uint32_t q)
{
	const uint64_t k = skidx[q + 1] - skidx[q];
	return k ? 63u - (uint32_t)__clzll((long long)k) : 0u; // 0 .. 31 for k < 2^32
}

// per workgroup: an LDS histogram of its queries' classes, then one global
// atomic per class (the classes of a block are few, so per-query global
// atomics would queue on the same words)
__global__ __launch_bounds__(256) void k_order_hist(const uint64_t *__restrict__ skidx, uint32_t nq, uint32_t *head)
{
	__shared__ uint32_t h[32];
	if (threadIdx.x < 32)
		h[threadIdx.x] = 0;
	__syncthreads();
	for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x)
		atomicAdd(&h[size_class(skidx, q) & 31u], 1u);
	__syncthreads();
	if (threadIdx.x < 32 && h[threadIdx.x])
		atomicAdd(&head[1 + threadIdx.x], h[threadIdx.x]);
}

__global__ void k_order_scan(uint32_t *head)
{
	if (blockIdx.x || threadIdx.x)
		return;
	uint32_t at = 0;
	for (int b = 31; b >= 0; --b) {
		head[33 + b] = at;
		at += head[1 + b];
	}
}

// per workgroup: its queries' ranks within their class in LDS, one global
// atomic per class to reserve the workgroup's range of each
__global__ __launch_bounds__(256) void k_order_scatter(const uint64_t *__restrict__ skidx, uint32_t nq, uint32_t *head,
						       uint32_t *order)
{
	__shared__ uint32_t cnt[32], base[32];
	for (uint32_t q0 = blockIdx.x * blockDim.x; q0 < nq; q0 += gridDim.x * blockDim.x) {
		if (threadIdx.x < 32)
			cnt[threadIdx.x] = 0;
		__syncthreads();
		const uint32_t q = q0 + threadIdx.x;
		const uint32_t c = q < nq ? size_class(skidx, q) & 31u : 0u;
		const uint32_t r = q < nq ? atomicAdd(&cnt[c], 1u) : 0u;
		__syncthreads();
		if (threadIdx.x < 32 && cnt[threadIdx.x])
			base[threadIdx.x] = atomicAdd(&head[33 + threadIdx.x], cnt[threadIdx.x]);
		__syncthreads();
		if (q < nq)
			order[base[c] + r] = q;
		__syncthreads();
	}
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n)
{
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
		p[i] = 0;
}

__global__ void k_zero_u64(unsigned long long *p, uint64_t n)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
		p[i] = 0;
}

// n words of each output plane (blockIdx.y: plane 0..3) in one launch, and
// the workspace head's words
__global__ __launch_bounds__(256) void k_zero_planes(unsigned long long *p0, unsigned long long *p1,
						     unsigned long long *p2, unsigned long long *p3, uint64_t n,
						     uint32_t *head, uint32_t head_words)
{
	if (blockIdx.x == 0 && blockIdx.y == 0)
		for (uint32_t k = threadIdx.x; k < head_words; k += blockDim.x)
			head[k] = 0;
	unsigned long long *p = blockIdx.y == 0 ? p0 : blockIdx.y == 1 ? p1 : blockIdx.y == 2 ? p2 : p3;
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
		p[i] = 0;
}

struct QJArgs {
	// index: the record table (2^bb home lines, then overflow lines) and the
	// sorted records (long contexts)
	const unsigned long long *brec;
	int w, bb;
	const uint64_t *ctxgid;
	const uint32_t *robj;
	uint64_t n;
	uint64_t ctx_max;
	uint64_t cmax; // min(ctx_max, 2^w - 1): contexts above it are in no index record
	uint32_t R; // ref_gnum
	unsigned obj_bits;
	// block
	const uint64_t *keys;
	const uint64_t *skidx; // relative, nq + 1
	uint32_t nq;
	int qry_gid_offset;
	uint32_t flags;
	// LDS table: slots (a power of two, sbits = log2) or direct rows
	uint32_t slots;
	int sbits;
	// work queue and spill rows (zero at rest: a workgroup zeroes what it spilled)
	uint32_t *next;
	const uint32_t *order; // queries in processing order
	// key-range work items (chunk > 0): item k is a range of the block's
	// keys (qj_item_range), counted query by query; the outputs are added
	// to (each query's part flushes its table), so they start zeroed
	uint32_t chunk;
	unsigned long long *spill; // gridDim.x rows of R cells x (2 | 4) u64
	int mode;
	// dense outputs (QJ_DENSE, QJ_ACCUM)
	unsigned long long *ctx, *objc, *uctx, *uobj;
	// select (QJ_SELECT)
	const uint32_t *qry_ctx, *ref_ctx;
	uint32_t ctxcut;
	double afcut;
	float anicut;
	int unique_rank;
	kssdhip_epilogue_opts_t vo;
	int vmetric;
	kssdhip_ani_row_t *rows;
	uint32_t cap;
	uint32_t *row_count;
	uint32_t *q_first, *q_count;
};

// The query's counter table in LDS: c = {num_ctx, num_mut2_ctx}, o =
// {diff_obj, diff_obj_section} as one u64 each (low word first, as the dense
// cells), uc/uo the unique-owner copies; gid the key of each slot (hash
// tables only).
struct QTable {
	LDS unsigned long long *c, *o, *uc, *uo;
	LDS uint32_t *gid;
	uint32_t mask;
	int shift;
};

__device__ __forceinline__ uint32_t slot_find(const QTable &t, uint32_t gid, bool insert)
{
	uint32_t s = (gid * 2654435761u) >> t.shift;
	for (int p = 0; p < MAX_PROBES; ++p) {
		const uint32_t g = t.gid[s];
		if (g == gid)
			return s;
		if (g == EMPTY) {
			if (!insert)
				return NONE;
			uint32_t expected = EMPTY;
			__hip_atomic_compare_exchange_strong(&t.gid[s], &expected, gid, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
							     __HIP_MEMORY_SCOPE_WORKGROUP);
			if (expected == EMPTY || expected == gid)
				return s;
		}
		s = (s + 1) & t.mask;
	}
	return NONE;
}

// The counters of genome gid: its cell s of the table (the genome itself
// in a direct row; NONE when the hashed table has no slot left for it).
template <bool DIRECT>
__device__ __forceinline__ uint32_t qj_slot(const QTable &t, uint32_t gid)
{
	return DIRECT ? gid : slot_find(t, gid, true);
}

// Adds one run with minimum m into cell s of genome gid (qj_slot), or into
// the workgroup's spill row when s is NONE.
template <bool UNIQ, bool DIRECT>
__device__ __forceinline__ void qj_add(const QTable &t, uint32_t s, uint32_t gid, int m, bool ucopy,
				       unsigned long long *spill_row, LDS uint32_t *spilled)
{
	const unsigned long long dc = 1ull | ((unsigned long long)(m > 1) << 32);
	const unsigned long long dob = 1ull | ((unsigned long long)m << 32);
	if (s != NONE) {
		LDS unsigned long long *c = ucopy ? t.uc : t.c;
		LDS unsigned long long *o = ucopy ? t.uo : t.o;
		__hip_atomic_fetch_add(&c[s], dc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		if (m)
			__hip_atomic_fetch_add(&o[s], dob, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		return;
	}
	// no slot left for this genome: the workgroup's spill row (read back
	// and zeroed when the query's output is written)
	*spilled = 1u;
	unsigned long long *cell = spill_row + (uint64_t)gid * (UNIQ ? 4u : 2u) + (ucopy ? 2u : 0u);
	atomicAdd(cell, dc);
	if (m)
		atomicAdd(cell + 1, dob);
}

// Lookups stage their lines in LDS: each wave owns 64 rows of 128 bytes,
// row m holding the line of the wave's m-th key.  Eight lanes load one line
// together (a 16-byte piece each, so one wave instruction reads eight whole
// lines), and a lane then reads its own row.  Piece p of row m sits at
// position p ^ (m & 7), so that both the eight lanes writing a row and
// eight consecutive lanes reading their rows meet distinct banks.  Per wave
// the rows are followed by own_wave(uniq) words: the sole-owner bookkeeping
// of unique counters (128 words), the walk's 64 one-byte key marks, and with
// unique counters the objects of the wave's 64 keys (without them, the
// objects take the first 64 words, which the bookkeeping leaves unused).
constexpr uint32_t STAGE_ROW = BR_SLOTS * 8;        // bytes per row
constexpr uint32_t STAGE_WAVE = 64 * STAGE_ROW;     // bytes per wave
__host__ __device__ constexpr uint32_t own_wave(bool uniq) { return uniq ? 128 + 16 + 64 : 128 + 16; }
__host__ __device__ constexpr uint32_t own_objs(bool uniq) { return uniq ? 128 + 16 : 0; }
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// Cross-lane moves within the 64-lane wave (DPP controls): the whole wave
// shifted by one lane (the lane left without a source takes `edge`), row
// shifts within rows of 16 lanes, and the row broadcasts that carry a scan
// from one row to the next.  They need every lane of the wave active.
constexpr int DPP_WAVE_SHL1 = 0x130;  // lane i takes lane i + 1
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane i takes lane i - 1
constexpr int DPP_ROW_SHR = 0x110;    // + n: lane i takes lane i - n of its row
constexpr int DPP_ROW_BCAST15 = 0x142; // lane 15 of each row to the next row
constexpr int DPP_ROW_BCAST31 = 0x143; // lane 31 to rows 2 and 3

__device__ __forceinline__ uint32_t from_prev_lane(uint32_t x, uint32_t edge)
{
	return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)x, DPP_WAVE_SHR1, 0xf, 0xf, false);
}

__device__ __forceinline__ uint32_t from_next_lane(uint32_t x, uint32_t edge)
{
	return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)x, DPP_WAVE_SHL1, 0xf, 0xf, false);
}

__device__ __forceinline__ uint64_t from_prev_lane64(uint64_t x, uint64_t edge)
{
	return ((uint64_t)from_prev_lane((uint32_t)(x >> 32), (uint32_t)(edge >> 32)) << 32) |
	       from_prev_lane((uint32_t)x, (uint32_t)edge);
}

// inclusive sum over lanes 0..i (row steps, then rows 1 and 3 take the row
// below, rows 2 and 3 take rows 0-1)
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x)
{
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 1, 0xf, 0xf, false);
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 2, 0xf, 0xf, false);
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 4, 0xf, 0xf, false);
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 8, 0xf, 0xf, false);
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST15, 0xa, 0xf, false);
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST31, 0xc, 0xf, false);
	return x;
}

// inclusive maximum over lanes 0..i
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t x)
{
	x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 1, 0xf, 0xf, false));
	x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 2, 0xf, 0xf, false));
	x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 4, 0xf, 0xf, false));
	x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_SHR + 8, 0xf, 0xf, false));
	x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST15, 0xa, 0xf, false));
	x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST31, 0xc, 0xf, false));
	return x;
}

// lane 8 * (i / 8) + K's value (each group of eight lanes takes its K-th)
template <uint32_t K>
__device__ __forceinline__ uint32_t group8_lane(uint32_t x)
{
	return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (K << 5));
}

__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long row_slot(const LDS unsigned long long *row, uint32_t lane, uint32_t j)
{
	return row[(((j >> 1) ^ (lane & 7u)) << 1) | (j & 1u)];
}

// Match bits of the 16 slots of a row, slot j at bit 15 - j, and the high
// word of the row's last slot.  Slot j holds a record or header of the
// context with remainder rem when its high word, cleared below the
// remainder (hmask = ~0 << (rs - 32)), equals remhi = rem << (rs - 32).
// Special slots (bit 63 set, empty) never match: rs plus the remainder's
// bits stay below 64, so remhi < 2^31.  Only the high words are read.
__device__ __forceinline__ uint32_t row_scan(const LDS unsigned long long *row, uint32_t lane, uint32_t remhi,
					     uint32_t hmask, uint32_t *lasthi)
{
	uint32_t mm = 0;
#pragma unroll
	for (uint32_t p = 0; p < BR_SLOTS / 2; ++p) {
		const LDS uint32_t *w = (const LDS uint32_t *)(row + ((p ^ (lane & 7u)) << 1));
		const uint32_t hx = w[1], hy = w[3];
		mm = 2u * mm + (uint32_t)((hx & hmask) == remhi);
		mm = 2u * mm + (uint32_t)((hy & hmask) == remhi);
		if (p == BR_SLOTS / 2 - 1)
			*lasthi = hy;
	}
	return mm;
}

// Record at position pos of key `owner`'s context: (gid << 32 | obj), from
// slot pos of the owner's row (inline contexts) or record pos of the sorted
// index (long contexts).
__device__ __forceinline__ uint64_t qj_record(const QJArgs &a, const LDS unsigned long long *wst, uint32_t owner,
					      bool lng, uint32_t pos, uint64_t objmask)
{
	const unsigned ob = a.obj_bits;
	if (lng)
		return ((a.ctxgid[pos] & GID_MASK) << 32) | ((uint64_t)a.robj[pos] & objmask);
	const unsigned long long x = row_slot(wst + owner * BR_SLOTS, owner, pos);
	return (((x >> ob) & GID_MASK) << 32) | (x & objmask);
}

// The walk of the wave's found contexts with one record per lane
// (command_ani.h:516-575; LOWER: :668-716, genomes below the query only).
// Key m of the wave (found) has cnt records, from slot src of its row or,
// for a long context, from record src of the sorted index; its group is
// the query's keys [wb + m, wb + m + gs).  The wave numbers all records in
// lane order (an inclusive sum of the counts); lane L takes record r0 + L
// of each round.  Its key: each key with records marks the position of its
// first record in this round with its lane, and a running maximum over the
// marks carries the key to the records that follow.  The lane holding the
// first record of a genome run counts the run: object-mismatch minimum over
// the cross product of the group's and the run's objects, one add per run
// (reading any further records of the run itself).  With unique counters
// each key's runs are counted in `own` (per lane of the wave: runs, then
// the last run's genome, minimum and whether it was counted), so that the
// key's lane applies the sole-owner rule afterwards.  The group's other
// objects come from the wave's copy of its keys' objects (the first nv
// keys) and, past those, from the keys in memory.  Called by the whole
// wave.
template <bool UNIQ, bool DIRECT>
__device__ __forceinline__ void qj_wave_walk(const QJArgs &a, const QTable &t, const LDS unsigned long long *wst,
					     LDS uint32_t *own, uint32_t lane, bool found, bool lng, uint32_t src,
					     uint32_t cnt, uint64_t wb, uint32_t nv, uint32_t gs, uint64_t key, int qgid,
					     unsigned long long *spill_row, LDS uint32_t *spilled)
{
	const unsigned ob = a.obj_bits;
	const uint64_t objmask = ob >= 32 ? 0xffffffffull : ((1ull << ob) - 1ull);
	const uint32_t mycnt = found ? cnt : 0u;
	const uint32_t incl = wave_incl_sum(mycnt);
	const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
	if (total == 0)
		return;
	const uint32_t first = incl - mycnt; // the key's first record number
	// read by the lanes of the key's records: record number r sits at
	// position base + r (a slot of the row or a record of the index), the
	// key's positions end before `end`
	const uint32_t base = src - first, end = src + mycnt;
	const uint32_t qobj = (uint32_t)(key & objmask);
	const uint32_t shape = (gs << 1) | (lng ? 1u : 0u);
	const bool groups = __any(found && gs > 1); // some key has a group of several keys
	const bool lower = (a.flags & KSSDHIP_LOWER) != 0;
	const bool ignore = (a.flags & KSSDHIP_IGNORECONFLICT) != 0;
	if (!groups && !__any(mycnt > 1)) {
		// every found context has one record and every group one key: each
		// key's lane counts its own record (one genome run of one record)
		if (mycnt) {
			const uint64_t rec = qj_record(a, wst, lane, lng, src, objmask);
			const uint32_t gid = (uint32_t)(rec >> 32);
			int mm = dna_popcount_dev(qobj ^ (uint32_t)rec);
			mm = mm < MIN_INIT ? mm : MIN_INIT;
			if (!(lower && (int)gid >= qgid)) {
				const uint32_t s = qj_slot<DIRECT>(t, gid);
				qj_add<UNIQ, DIRECT>(t, s, gid, mm, false, spill_row, spilled);
				if (UNIQ) // the context's only run: its sole owner
					qj_add<UNIQ, DIRECT>(t, s, gid, mm, true, spill_row, spilled);
			}
		}
		return;
	}
	LDS uint8_t *mark = (LDS uint8_t *)(own + 128);
	const LDS uint32_t *qo = own + own_objs(UNIQ);
	uint32_t carry_gid = NONE, carry_owner = 64;
	for (uint32_t r0 = 0; r0 < total; r0 += 64) { // wave-uniform
		const uint32_t tt = r0 + lane;
		const bool act = tt < total;
		if (mycnt && first - r0 < 64u)
			mark[first - r0] = (uint8_t)(lane + 1u);
		wave_sync();
		uint32_t m = mark[lane];
		mark[lane] = 0; // clear for the next round (the marks start at zero)
		if (lane == 0 && m == 0)
			m = carry_owner + 1u; // record r0 continues the previous round's last key
		const uint32_t owner = (wave_incl_max(m) - 1u) & 63u;
		const uint32_t oshape = __shfl(shape, (int)owner);
		const uint32_t pos = __shfl(base, (int)owner) + tt;
		const uint32_t oend = __shfl(end, (int)owner);
		const uint32_t oobj = __shfl(qobj, (int)owner);
		const uint64_t oi = wb + owner;
		const bool olng = oshape & 1u;
		const uint32_t ogs = oshape >> 1;
		const bool ok = act && pos < oend;
		const uint64_t rec = ok ? qj_record(a, wst, owner, olng, pos, objmask) : 0ull;
		const uint32_t gid = ok ? (uint32_t)(rec >> 32) : NONE;
		// minimum over the group's objects against this record's object
		auto gmin = [&](uint32_t obj) -> int {
			int mm = dna_popcount_dev(oobj ^ obj);
			mm = mm < MIN_INIT ? mm : MIN_INIT;
			if (groups) {
				for (uint32_t x = 1; x < ogs && mm; ++x) {
					const uint32_t y = owner + x;
					const uint32_t o = y < nv ? qo[y] : (uint32_t)(a.keys[oi + x] & objmask);
					const int pc = dna_popcount_dev(o ^ obj);
					mm = pc < mm ? pc : mm;
				}
			}
			return mm;
		};
		int mmin = ok ? gmin((uint32_t)rec) : MIN_INIT;
		const uint32_t pg = from_prev_lane(gid, carry_gid), po = from_prev_lane(owner, carry_owner);
		const uint32_t ng = from_next_lane(gid, NONE), no = from_next_lane(owner, 64u);
		carry_gid = (uint32_t)__builtin_amdgcn_readlane((int)gid, 63);
		carry_owner = (uint32_t)__builtin_amdgcn_readlane((int)owner, 63);
		if (!ok || (po == owner && pg == gid))
			continue; // not the first record of its run
		uint32_t n = 1;
		const bool more = lane < 63 ? (tt + 1 < total && no == owner && ng == gid) : pos + 1 < oend;
		for (uint32_t pp = pos + 1; more && pp < oend; ++pp) {
			const uint64_t y = qj_record(a, wst, owner, olng, pp, objmask);
			if ((uint32_t)(y >> 32) != gid)
				break;
			++n;
			if (mmin) {
				const int pc = gmin((uint32_t)y);
				mmin = pc < mmin ? pc : mmin;
			}
		}
		if (lower && (int)gid >= qgid)
			continue;
		const bool counted = !(ignore && n > 1);
		const uint32_t s = counted ? qj_slot<DIRECT>(t, gid) : NONE;
		if (UNIQ) { // the run's cell (its genome, flagged, when it has none), minimum and whether counted
			__hip_atomic_fetch_add(&own[owner], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
			own[64 + owner] = (s != NONE ? s : gid | (1u << 25)) | ((uint32_t)mmin << 20) | ((uint32_t)counted << 24);
		}
		if (counted)
			qj_add<UNIQ, DIRECT>(t, s, gid, mmin, false, spill_row, spilled);
	}
	if (UNIQ) {
		wave_sync();
		if (mycnt) {
			const uint32_t runs = own[lane], info = own[64 + lane];
			if (runs == 1 && ((info >> 24) & 1u))
				qj_add<UNIQ, DIRECT>(t, (info >> 25) & 1u ? NONE : info & GID_MASK, info & GID_MASK,
						     (int)((info >> 20) & 15u), true, spill_row, spilled);
			own[lane] = 0;
		}
	}
}

// Counts of query q into the table: the first key of each context group
// does the work for the group.  Each wave takes 64 consecutive keys at a
// time (strided by the workgroup size, so loads of the keys coalesce; the
// next keys are loaded while the current ones are looked up), stages their
// home lines in its LDS rows (`stage`: STAGE_WAVE bytes per wave; all eight
// line loads in flight before the first row is written; lanes without a
// lookup load the home line of context 0, a valid address), scans each row
// for its key's context and walks the found contexts together; a home that
// continues in an overflow line has that line loaded into its row and is
// scanned again, until every lookup is resolved.
template <bool UNIQ, bool DIRECT, bool ITEMS>
__device__ void qj_count(const QJArgs &a, const QTable &t, uint32_t q, uint64_t k0, uint64_t k1,
			 unsigned long long *spill_row, LDS uint32_t *spilled, LDS unsigned long long *stage)
{
	// the query's keys [kb, ke); groups headed in [k0, k1) (ITEMS) or all
	const uint64_t kb = a.skidx[q], ke = a.skidx[q + 1];
	if (!ITEMS) {
		k0 = kb;
		k1 = ke;
	}
	const unsigned ob = a.obj_bits;
	const int qgid = a.qry_gid_offset + (int)q;
	const uint64_t B = blockDim.x; // a multiple of 64
	const uint32_t lane = threadIdx.x & 63u, g = lane >> 3, piece = lane & 7u;
	LDS unsigned long long *wst = stage + (threadIdx.x >> 6) * (STAGE_WAVE / 8);
	LDS unsigned long long *row = wst + lane * BR_SLOTS;
	LDS uint32_t *own = (LDS uint32_t *)(stage + (blockDim.x >> 6) * (STAGE_WAVE / 8)) + (threadIdx.x >> 6) * own_wave(UNIQ);
	const int rsh = a.w - a.bb;
	const uint64_t remmask = (1ull << rsh) - 1ull;
	const int hs = br_rem_shift(ob) - 32;
	const uint32_t hmask = ~0u << hs;
	uint64_t wb = k0 + (threadIdx.x & ~63u); // wave-uniform
	uint64_t key = wb + lane < k1 ? a.keys[wb + lane] : 0ull;
	uint64_t before = lane == 0 && wb > kb && wb < k1 ? a.keys[wb - 1] : ~0ull; // the key ahead of the wave's first
	// the key after the last one the wave holds (keys [wb, min(wb + 64, k1)))
	auto beyond_of = [&](uint64_t b) -> uint64_t {
		const uint64_t p = min(b + 64, k1);
		return p < ke ? a.keys[p] : ~0ull;
	};
	uint64_t beyond = wb < k1 ? beyond_of(wb) : ~0ull;
	for (; wb < k1; wb += B) {
		const uint64_t i = wb + lane;
		const bool in = i < k1;
		const uint64_t c = key >> ob;
		const uint64_t prev = from_prev_lane64(key, before);
		// cmax: the largest context the table can hold (index contexts <= ctx_max, below 2^w)
		const bool on = in && c <= a.cmax && (i == kb || (prev >> ob) != c);
		const uint64_t p = br_perm(on ? c : 0ull, a.w);
		const uint32_t remhi = (uint32_t)(p & remmask) << hs;
		const uint32_t line = (uint32_t)(p >> rsh); // tables hold fewer than 2^32 lines
		// load k: the eight lanes of group g read row 8g + k's line.  The
		// loads are nontemporal: lines of the table are rarely read twice,
		// and keeping them out of the L2 leaves it to the page-table
		// entries of the multi-gigabyte table (random line reads of an
		// 8 GiB buffer: 52 G/s nontemporal, 48 G/s otherwise)
		const u64x2 *lines = (const u64x2 *)a.brec + piece;
		u64x2 v[8];
		v[0] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<0>(line) * (BR_SLOTS / 2)]);
		v[1] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<1>(line) * (BR_SLOTS / 2)]);
		v[2] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<2>(line) * (BR_SLOTS / 2)]);
		v[3] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<3>(line) * (BR_SLOTS / 2)]);
		v[4] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<4>(line) * (BR_SLOTS / 2)]);
		v[5] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<5>(line) * (BR_SLOTS / 2)]);
		v[6] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<6>(line) * (BR_SLOTS / 2)]);
		v[7] = __builtin_nontemporal_load(&lines[(uint64_t)group8_lane<7>(line) * (BR_SLOTS / 2)]);
		// the next keys, while the lines are in flight
		const uint64_t nb = wb + B;
		const uint64_t nkey = nb + lane < k1 ? a.keys[nb + lane] : 0ull;
		const uint64_t nbefore = lane == 0 && nb < k1 ? a.keys[nb - 1] : ~0ull;
		const uint64_t nbeyond = nb < k1 ? beyond_of(nb) : ~0ull;
#pragma unroll
		for (uint32_t k = 0; k < 8; ++k)
			*(LDS u64x2 *)(wst + (8 * g + k) * BR_SLOTS + ((piece ^ k) << 1)) = v[k];
		// the group: this key and the keys that follow with the same
		// context.  Bit L of `cont`: lane L's key continues the previous
		// key's context; a head's group runs over the set bits after its
		// lane, and when it reaches the last key the wave holds (lane 63, or
		// the item's end k1) and the key after that (`beyond`) continues it,
		// the rest is read from memory
		const uint64_t cont = __ballot(in && i != kb && (prev >> ob) == c);
		uint32_t gs = 1;
		if (on) {
			const uint32_t run = lane < 63 ? (uint32_t)__builtin_ctzll(~(cont >> (lane + 1))) : 0u;
			gs += run;
			const uint64_t p = min(wb + 64, k1);
			if (i + run + 1 == p && p < ke && (beyond >> ob) == c)
				while (i + gs < ke && (a.keys[i + gs] >> ob) == c)
					++gs;
		}
		// the objects of the wave's keys [wb, wb + nv), for the walk's group minima
		const uint32_t nv = (uint32_t)min((uint64_t)64, k1 - wb);
		if (__any(gs > 1u) && lane < nv)
			own[own_objs(UNIQ) + lane] = (uint32_t)key & (ob >= 32 ? 0xffffffffu : (1u << ob) - 1u);
		bool pending = on;
		for (;;) { // wave-uniform: until every lookup is resolved
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			uint32_t lasthi = 0;
			const uint32_t mm = pending ? row_scan(row, lane, remhi, hmask, &lasthi) : 0u;
			const bool found = mm != 0;
			const uint32_t lo = found ? (uint32_t)__builtin_clz(mm) - 16u : 0u; // the first matching slot
			bool lng = false;
			uint32_t src = lo, cnt = (uint32_t)__builtin_popcount(mm);
			if (found && ((row_slot(row, lane, lo) >> ob) & GID_MASK) == GID_MASK) { // header + payload
				const unsigned long long pay = row_slot(row, lane, lo + 1);
				lng = true;
				src = (uint32_t)pay;
				cnt = (uint32_t)((pay >> 32) & BR_MAX_LONG);
			}
			qj_wave_walk<UNIQ, DIRECT>(a, t, wst, own, lane, found, lng, src, cnt, wb, nv, gs, key, qgid,
						   spill_row, spilled);
			pending = pending && !found && (lasthi >> 30) == 2u; // slot 15 continues the home
			if (!__any(pending))
				break;
			const unsigned long long nline = pending ? row_slot(row, lane, BR_SLOTS - 1) & ~BR_HIGH : 0ull;
			__builtin_amdgcn_wave_barrier(); // rows read by the walk before any is replaced
			if (pending) { // the home continues in this overflow line
				const u64x2 *nsrc = (const u64x2 *)(a.brec + nline * BR_SLOTS);
#pragma unroll
				for (uint32_t pp = 0; pp < BR_SLOTS / 2; ++pp)
					*(LDS u64x2 *)(row + ((pp ^ (lane & 7u)) << 1)) = __builtin_nontemporal_load(&nsrc[pp]);
			}
		}
		__builtin_amdgcn_wave_barrier(); // every row read before the next keys' lines land
		key = nkey;
		before = nbefore;
		beyond = nbeyond;
	}
}

struct Cell {
	unsigned long long c, o, uc, uo;
};

// Final counters of genome g (direct rows or table slot, plus the spill row
// when the query spilled; zero_spill clears the spill cells it read).
template <bool UNIQ, bool DIRECT>
__device__ __forceinline__ Cell cell_of_gid(const QTable &t, uint32_t g, bool spilled, unsigned long long *spill_row,
					    bool zero_spill)
{
	Cell v = {0, 0, 0, 0};
	const uint32_t s = DIRECT ? g : slot_find(t, g, false);
	if (s != NONE) {
		v.c = t.c[s];
		v.o = t.o[s];
		if (UNIQ) {
			v.uc = t.uc[s];
			v.uo = t.uo[s];
		}
	}
	if (!DIRECT && spilled) {
		unsigned long long *sp = spill_row + (uint64_t)g * (UNIQ ? 4u : 2u);
		v.c += sp[0];
		v.o += sp[1];
		if (UNIQ) {
			v.uc += sp[2];
			v.uo += sp[3];
		}
		if (zero_spill) {
			sp[0] = 0;
			sp[1] = 0;
			if (UNIQ) {
				sp[2] = 0;
				sp[3] = 0;
			}
		}
	}
	return v;
}

// The pass gate of the index path's ranking loop (command_ani.h:585-606);
// rank: the ranking value.  With unique_rank the FP64 naive ANI is compared
// 1e-9 below anicut, so that the host's own evaluation keeps a subset.
template <bool UNIQ>
__device__ __forceinline__ bool qj_gate(const QJArgs &a, uint32_t q, uint32_t g, const Cell &v, double *rank)
{
	const uint32_t xny = (uint32_t)v.c, dobj = (uint32_t)v.o, dsec = (uint32_t)(v.o >> 32);
	if (xny < a.ctxcut)
		return false;
	const uint32_t qc = a.qry_ctx[q], rc = a.ref_ctx[g];
	if (qc == 0 || rc == 0)
		return false;
	const double afq = (double)xny / (double)qc, afr = (double)xny / (double)rc;
	if (!((afq > afr ? afq : afr) >= a.afcut))
		return false;
	const float dist = (float)dobj / (float)xny / 7.0f;
	double ani = (double)(float)(1.0f - dist);
	if (UNIQ && a.unique_rank) {
		const uint32_t uxny = (uint32_t)v.uc;
		const bool u = (uint64_t)uxny * 20 >= rc; // ani_use_unique_context, command_ani.h:437-440
		ani = 1.0 - (u ? kssd_naive_dist(uxny, (uint32_t)v.uo, (uint32_t)(v.uo >> 32))
			       : kssd_naive_dist(xny, dobj, dsec));
		*rank = ani;
		return ani >= (double)a.anicut - 1e-9;
	}
	*rank = ani;
	return ani >= (double)a.anicut;
}

template <bool UNIQ, bool DIRECT>
__device__ void qj_output(const QJArgs &a, const QTable &t, uint32_t q, bool spilled, unsigned long long *spill_row,
			  LDS uint32_t *s_n, LDS uint32_t *s_base)
{
	const uint32_t R = a.R;
	const uint64_t row = (uint64_t)q * R;
	if (a.mode == QJ_DENSE) {
		for (uint32_t g = threadIdx.x; g < R; g += blockDim.x) {
			const Cell v = cell_of_gid<UNIQ, DIRECT>(t, g, spilled, spill_row, true);
			a.ctx[row + g] = v.c;
			a.objc[row + g] = v.o;
			if (UNIQ) {
				a.uctx[row + g] = v.uc;
				a.uobj[row + g] = v.uo;
			}
		}
		return;
	}
	// candidates: every genome when the table is direct or the query
	// spilled, else the occupied slots
	const bool by_gid = DIRECT || spilled;
	const uint32_t ncand = by_gid ? R : a.slots;
	if (a.mode == QJ_ACCUM) {
		for (uint32_t k = threadIdx.x; k < ncand; k += blockDim.x) {
			uint32_t g = k;
			Cell v;
			if (by_gid) {
				v = cell_of_gid<UNIQ, DIRECT>(t, g, spilled, spill_row, true);
			} else {
				g = t.gid[k];
				if (g == EMPTY)
					continue;
				v = Cell{t.c[k], t.o[k], UNIQ ? t.uc[k] : 0ull, UNIQ ? t.uo[k] : 0ull};
			}
			if (v.c)
				atomicAdd(&a.ctx[row + g], v.c);
			if (v.o)
				atomicAdd(&a.objc[row + g], v.o);
			if (UNIQ) {
				if (v.uc)
					atomicAdd(&a.uctx[row + g], v.uc);
				if (v.uo)
					atomicAdd(&a.uobj[row + g], v.uo);
			}
		}
		return;
	}
	// QJ_SELECT: count the passing pairs, take a segment of the output, write
	for (int phase = 0; phase < 2; ++phase) {
		uint32_t mine = 0;
		for (uint32_t k = threadIdx.x; k < ncand; k += blockDim.x) {
			uint32_t g = k;
			Cell v;
			if (by_gid) {
				v = cell_of_gid<UNIQ, DIRECT>(t, g, spilled, spill_row, phase == 1);
			} else {
				g = t.gid[k];
				if (g == EMPTY)
					continue;
				v = Cell{t.c[k], t.o[k], UNIQ ? t.uc[k] : 0ull, UNIQ ? t.uo[k] : 0ull};
			}
			if (!v.c)
				continue;
			double rank;
			if (!qj_gate<UNIQ>(a, q, g, v, &rank))
				continue;
			if (phase == 0) {
				++mine;
				continue;
			}
			const uint32_t at = *s_base + __hip_atomic_fetch_add(s_n, 1u, __ATOMIC_RELAXED,
									   __HIP_MEMORY_SCOPE_WORKGROUP);
			if (at >= a.cap)
				continue;
			kssdhip_ani_row_t w;
			w.query = q;
			w.ref = g;
			w.ctx.num_ctx = (uint32_t)v.c;
			w.ctx.num_mut2_ctx = (uint32_t)(v.c >> 32);
			w.obj.diff_obj = (uint32_t)v.o;
			w.obj.diff_obj_section = (uint32_t)(v.o >> 32);
			w.unique_ctx.num_ctx = (uint32_t)v.uc;
			w.unique_ctx.num_mut2_ctx = (uint32_t)(v.uc >> 32);
			w.unique_obj.diff_obj = (uint32_t)v.uo;
			w.unique_obj.diff_obj_section = (uint32_t)(v.uo >> 32);
			w.rank_ani = rank;
			w.value = 0.0; // k_row_values
			a.rows[at] = w;
		}
		if (phase == 0) {
			if (mine)
				__hip_atomic_fetch_add(s_n, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
			__syncthreads();
			if (threadIdx.x == 0) {
				const uint32_t npass = *s_n;
				const uint32_t base = npass ? atomicAdd(a.row_count, npass) : 0u;
				a.q_first[q] = base;
				a.q_count[q] = npass;
				*s_base = base;
				*s_n = 0;
			}
			__syncthreads();
		}
	}
}

// One query's keys [k0, k1) (ITEMS; else all its keys) by the workgroup:
// the table cleared, the keys counted, the query's output written or added.
template <bool UNIQ, bool DIRECT, bool ITEMS>
__device__ __forceinline__ void qj_query(const QJArgs &a, const QTable &t, uint32_t q, uint64_t k0, uint64_t k1,
					 unsigned long long *spill_row, LDS unsigned long long *stage,
					 LDS uint32_t *s_spilled, LDS uint32_t *s_n, LDS uint32_t *s_base)
{
	const uint32_t ncell = DIRECT ? a.R : a.slots;
	for (uint32_t k = threadIdx.x; k < ncell; k += blockDim.x) {
		t.c[k] = 0;
		t.o[k] = 0;
		if (UNIQ) {
			t.uc[k] = 0;
			t.uo[k] = 0;
		}
		if (!DIRECT)
			t.gid[k] = EMPTY;
	}
	if (threadIdx.x == 0) {
		*s_spilled = 0;
		*s_n = 0;
	}
	__syncthreads();
	qj_count<UNIQ, DIRECT, ITEMS>(a, t, q, k0, k1, spill_row, s_spilled, stage);
	__syncthreads();
	const bool spilled = *s_spilled != 0;
	qj_output<UNIQ, DIRECT>(a, t, q, spilled, spill_row, s_n, s_base);
}

// Keys [b, e) of item k, of a block of `total` keys shared by g workgroups
// (the reference's guided schedule over queries, command_ani.h:481, taken
// to keys): items come in rounds of g, round j's items of total / (g *
// 2^(j+1)) keys (rounded up to 64), so each round takes about half of what
// is left, until they would not exceed `chunk` keys; then items of `chunk`
// keys.  A function of k alone: workgroups take item numbers from a counter.
__device__ __forceinline__ void qj_item_range(uint64_t k, uint64_t total, uint32_t g, uint32_t chunk, uint64_t *b,
					      uint64_t *e)
{
	uint64_t at = 0;
	for (int j = 0; j < 62; ++j) {
		const uint64_t size = ((total >> (j + 1)) / g + 63) / 64 * 64;
		if (size <= chunk)
			break;
		if (k < g) {
			at += k * size;
			*b = min(at, total);
			*e = min(at + size, total);
			return;
		}
		at += g * size;
		k -= g;
	}
	at += k * chunk;
	*b = min(at, total);
	*e = min(at + chunk, total);
}

// Persistent workgroups: each takes the next query of the block (largest
// first: the reference's guided schedule over queries, command_ani.h:481)
// or the next key-range item from the work counter until none is left.
template <bool UNIQ, bool DIRECT, bool ITEMS>
__global__ __launch_bounds__(QJ_MAX_THREADS) __attribute__((amdgpu_waves_per_eu(QJ_WAVES_PER_SIMD))) void k_qjoin(QJArgs a)
{
	extern __shared__ unsigned long long qj_lds[];
	__shared__ uint32_t s_q, s_spilled, s_n, s_base;
	__shared__ uint64_t s_k0, s_k1;
	const uint32_t ncell = DIRECT ? a.R : a.slots;
	QTable t;
	t.c = (LDS unsigned long long *)qj_lds;
	t.o = t.c + ncell;
	t.uc = UNIQ ? t.o + ncell : nullptr;
	t.uo = UNIQ ? t.uc + ncell : nullptr;
	t.gid = DIRECT ? nullptr : (LDS uint32_t *)(t.o + (UNIQ ? 3 : 1) * (uint64_t)ncell);
	// the lookup rows follow the table (16-byte aligned: cells are 8 bytes,
	// slot counts multiples of 4)
	LDS unsigned long long *stage =
		(LDS unsigned long long *)qj_lds + (uint64_t)ncell * (UNIQ ? 4 : 2) + (DIRECT ? 0 : ncell / 2);
	t.mask = a.slots - 1;
	t.shift = 32 - a.sbits;
	unsigned long long *spill_row = DIRECT ? nullptr : a.spill + (uint64_t)blockIdx.x * a.R * (UNIQ ? 4u : 2u);
	{ // the walk's marks and sole-owner counts start at zero (each lane clears what it used)
		LDS uint32_t *own = (LDS uint32_t *)(stage + (blockDim.x >> 6) * (STAGE_WAVE / 8));
		for (uint32_t k = threadIdx.x; k < (blockDim.x >> 6) * own_wave(UNIQ); k += blockDim.x)
			own[k] = 0;
	}
	for (;;) {
		if (threadIdx.x == 0) {
			const uint32_t k = atomicAdd(a.next, 1u);
			if (ITEMS) { // item k's keys (absolute indices) and the first query they reach
				const uint64_t base = a.skidx[0], total = a.skidx[a.nq] - base;
				uint64_t b, e;
				qj_item_range(k, total, gridDim.x, a.chunk, &b, &e);
				s_q = a.nq;
				if (b < total) {
					uint32_t lo = 0, hi = a.nq; // largest q with skidx[q] <= key b
					while (hi - lo > 1) {
						const uint32_t mid = (lo + hi) >> 1;
						if (a.skidx[mid] <= base + b)
							lo = mid;
						else
							hi = mid;
					}
					s_q = lo;
					s_k0 = base + b;
					s_k1 = base + e;
				}
			} else {
				s_q = k < a.nq ? a.order[k] : a.nq;
			}
		}
		__syncthreads(); // also: every thread is done with the previous query's table
		uint32_t q = s_q;
		if (q >= a.nq)
			break;
		if (!ITEMS) {
			qj_query<UNIQ, DIRECT, ITEMS>(a, t, q, 0, 0, spill_row, stage, (LDS uint32_t *)&s_spilled,
							      (LDS uint32_t *)&s_n, (LDS uint32_t *)&s_base);
			continue;
		}
		// each query the key range reaches, in turn
		const uint64_t i0 = s_k0, i1 = s_k1;
		for (;;) {
			const uint64_t k0 = max(i0, a.skidx[q]), k1 = min(i1, a.skidx[q + 1]);
			if (k0 < k1)
				qj_query<UNIQ, DIRECT, ITEMS>(a, t, q, k0, k1, spill_row, stage, (LDS uint32_t *)&s_spilled,
							      (LDS uint32_t *)&s_n, (LDS uint32_t *)&s_base);
			if (++q >= a.nq || a.skidx[q] >= i1)
				break;
			__syncthreads(); // every thread is done with the previous query's table
		}
	}
}

// FP64 value of each stored row (the number of rows is read on the device,
// so the launch can sit in a graph)
__global__ __launch_bounds__(256) void k_row_values(kssdhip_ani_row_t *rows, const uint32_t *count, uint32_t cap,
						    const uint32_t *__restrict__ qry_ctx,
						    const uint32_t *__restrict__ ref_ctx, kssdhip_epilogue_opts_t vo,
						    int vmetric)
{
	const uint32_t n = *count < cap ? *count : cap;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		kssdhip_ani_row_t &w = rows[i];
		w.value = epi_cell_value(vo, vmetric, w.ctx.num_ctx, w.ctx.num_mut2_ctx, w.obj.diff_obj,
					 w.obj.diff_obj_section, qry_ctx[w.query], ref_ctx[w.ref]);
	}
}

// Marks the table slots of every index context that occurs in the block
// (marks: one byte per slot, zeroed before); k_count_marks then adds one
// record per marked record slot and a long context's count per marked header.
__global__ __launch_bounds__(256) void k_mark_matched(QJArgs a, uint8_t *marks)
{
	const unsigned ob = a.obj_bits;
	const int rs = br_rem_shift(ob);
	const int rsh = a.w - a.bb;
	const uint64_t total = a.skidx[a.nq];
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t c = a.keys[i] >> ob;
		if (c > a.ctx_max || (c >> a.w))
			continue;
		if (i > 0 && (a.keys[i - 1] >> ob) == c)
			continue; // the previous key (of this or the previous query) marked this context
		const uint64_t p = br_perm(c, a.w);
		const uint64_t rem = p & ((1ull << rsh) - 1ull);
		uint64_t line = p >> rsh;
		for (;;) {
			const unsigned long long *ln = a.brec + line * BR_SLOTS;
			bool hit = false;
			for (int j = 0; j < BR_SLOTS; ++j) {
				const unsigned long long x = ln[j];
				if (x < BR_HIGH && (x >> rs) == rem) {
					marks[line * BR_SLOTS + j] = 1;
					hit = true;
				}
			}
			const unsigned long long last = ln[BR_SLOTS - 1];
			if (hit || (last & BR_PAYLOAD) != BR_HIGH)
				break;
			line = last & ~BR_HIGH;
		}
	}
}

__global__ __launch_bounds__(256) void k_count_marks(const uint8_t *marks, const unsigned long long *__restrict__ brec,
						     unsigned ob, uint64_t slots, unsigned long long *count)
{
	unsigned long long k = 0;
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * blockDim.x) {
		if (!marks[i])
			continue;
		const unsigned long long x = brec[i];
		k += ((x >> ob) & GID_MASK) == GID_MASK ? ((brec[i + 1] >> 32) & BR_MAX_LONG) : 1ull;
	}
	for (int off = 32; off > 0; off >>= 1)
		k += __shfl_xor(k, off, 64);
	if ((threadIdx.x & 63) == 0 && k)
		atomicAdd(count, k);
}

// The pass gate over dense count rows (small blocks, counted by the
// one-thread-per-key lookup): per-query pass counts, their prefix sums as
// the segments' first rows, then the passing rows written into their
// query's segment.
template <bool UNIQ>
__device__ __forceinline__ bool dense_cell(const QJArgs &a, uint64_t i, Cell *v)
{
	v->c = a.ctx[i];
	if (!v->c)
		return false;
	v->o = a.objc[i];
	v->uc = UNIQ ? a.uctx[i] : 0ull;
	v->uo = UNIQ ? a.uobj[i] : 0ull;
	return true;
}

template <bool UNIQ>
__global__ __launch_bounds__(256) void k_gate_dense_count(QJArgs a)
{
	const uint64_t cells = (uint64_t)a.nq * a.R;
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * blockDim.x) {
		Cell v;
		double rank;
		const uint32_t q = (uint32_t)(i / a.R), g = (uint32_t)(i % a.R);
		if (dense_cell<UNIQ>(a, i, &v) && qj_gate<UNIQ>(a, q, g, v, &rank))
			atomicAdd(&a.q_count[q], 1u);
	}
}

// one thread: segments in query order, the total, cursors zeroed
__global__ void k_gate_dense_scan(QJArgs a, uint32_t *cursor)
{
	if (blockIdx.x || threadIdx.x)
		return;
	uint32_t at = 0;
	for (uint32_t q = 0; q < a.nq; ++q) {
		a.q_first[q] = at;
		at += a.q_count[q];
		cursor[q] = 0;
	}
	*a.row_count = at;
}

template <bool UNIQ>
__global__ __launch_bounds__(256) void k_gate_dense_write(QJArgs a, uint32_t *cursor)
{
	const uint64_t cells = (uint64_t)a.nq * a.R;
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * blockDim.x) {
		Cell v;
		double rank;
		const uint32_t q = (uint32_t)(i / a.R), g = (uint32_t)(i % a.R);
		if (!dense_cell<UNIQ>(a, i, &v) || !qj_gate<UNIQ>(a, q, g, v, &rank))
			continue;
		const uint32_t at = a.q_first[q] + atomicAdd(&cursor[q], 1u);
		if (at >= a.cap)
			continue;
		kssdhip_ani_row_t w;
		w.query = q;
		w.ref = g;
		w.ctx.num_ctx = (uint32_t)v.c;
		w.ctx.num_mut2_ctx = (uint32_t)(v.c >> 32);
		w.obj.diff_obj = (uint32_t)v.o;
		w.obj.diff_obj_section = (uint32_t)(v.o >> 32);
		w.unique_ctx.num_ctx = (uint32_t)v.uc;
		w.unique_ctx.num_mut2_ctx = (uint32_t)(v.uc >> 32);
		w.unique_obj.diff_obj = (uint32_t)v.uo;
		w.unique_obj.diff_obj_section = (uint32_t)(v.uo >> 32);
		w.rank_ani = rank;
		w.value = 0.0; // k_row_values
		a.rows[at] = w;
	}
}

struct QJShape {
	bool direct;
	uint32_t slots;
	int sbits;
	int threads;
	size_t lds;
	unsigned grid;
	size_t spill_bytes;
};

int ilog2(uint32_t v)
{
	int b = 0;
	while ((1u << (b + 1)) <= v)
		++b;
	return b;
}

template <bool UNIQ, bool DIRECT, bool ITEMS>
const void *qj_kernel()
{
	return (const void *)k_qjoin<UNIQ, DIRECT, ITEMS>;
}

const void *qj_fn(bool uniq, bool direct, bool items)
{
	if (items)
		return uniq ? (direct ? qj_kernel<true, true, true>() : qj_kernel<true, false, true>())
			    : (direct ? qj_kernel<false, true, true>() : qj_kernel<false, false, true>());
	return uniq ? (direct ? qj_kernel<true, true, false>() : qj_kernel<true, false, false>())
		    : (direct ? qj_kernel<false, true, false>() : qj_kernel<false, false, false>());
}

void qj_layout(const kssdhip_index *ix, bool uniq, QJShape *sh)
{
	const uint32_t R = (uint32_t)ix->ref_gnum;
	const uint32_t cell_bytes = uniq ? 32u : 16u;
	sh->threads = ix->qthreads ? ix->qthreads : DEFAULT_THREADS;
	// a direct per-genome row when it fits, unless slots were asked for (a
	// smaller hashed table lets more workgroups share a CU)
	sh->direct = !ix->qslots && (uint64_t)R * cell_bytes <= DIRECT_LDS_MAX;
	if (sh->direct) {
		sh->slots = 0;
		sh->sbits = 0;
		sh->lds = (size_t)R * cell_bytes;
	} else {
		sh->slots = ix->qslots ? (uint32_t)ix->qslots : DEFAULT_SLOTS;
		sh->sbits = ilog2(sh->slots);
		sh->lds = (size_t)sh->slots * (cell_bytes + 4u);
	}
	sh->lds = (sh->lds + 15) / 16 * 16 + (size_t)(sh->threads / 64) * (STAGE_WAVE + own_wave(uniq) * 4);
}

int qj_shape(const kssdhip_index *ix, uint32_t nq, bool uniq, QJShape *sh)
{
	qj_layout(ix, uniq, sh);
	const int per_cu = ix->qj_per_cu[uniq ? 1 : 0];
	if (per_cu < 1)
		return kssdhip_fail(KSSDHIP_EINVAL, "qjoin: %d threads with %zu bytes of LDS do not fit a CU", sh->threads,
				    sh->lds);
	const uint32_t R = (uint32_t)ix->ref_gnum;
	uint64_t g = (uint64_t)per_cu * (uint64_t)(ix->cus > 0 ? ix->cus : 256);
	if (g > nq)
		g = nq;
	if (g < 1)
		g = 1;
	sh->grid = (unsigned)g;
	sh->spill_bytes = sh->direct ? 0 : (size_t)sh->grid * R * (uniq ? 32u : 16u);
	return 0;
}

// The call's workspace: [0, WS_HEAD) the work counter and the order's
// histogram and cursors, then the spill rows (zero at rest: a workgroup
// clears the cells it spilled), then the work order.  The spill part never
// moves and only grows (ws->zero_bytes), so the order of a later, smaller
// block cannot land on cells that must read zero.  Leaves the counter and
// histogram zeroed and the order built, all by kernels on s.
struct QJWork {
	uint32_t *head;
	uint32_t *order;
	unsigned long long *spill;
};

size_t qj_order_bytes(uint32_t nq)
{
	return ((size_t)nq * 4 + WS_ALIGN - 1) / WS_ALIGN * WS_ALIGN;
}

void qj_zero(void *p, size_t bytes, hipStream_t s)
{
	const uint64_t words = bytes / 8;
	if (words)
		hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<uint64_t>((words + 255) / 256, 8192)), dim3(256), 0, s,
				   (unsigned long long *)p, words);
}

// zeroes the head and builds the work order (largest queries first)
void qj_order(const uint64_t *d_skidx, uint32_t nq, hipStream_t s, const QJWork &w, bool order, bool head = true)
{
	if (head)
		hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(128), 0, s, w.head, WS_HEAD_WORDS);
	if (!order)
		return;
	const unsigned g = (unsigned)std::max<uint32_t>(1, std::min<uint32_t>((nq + 255) / 256, 2048));
	hipLaunchKernelGGL(k_order_hist, dim3(g), dim3(256), 0, s, d_skidx, nq, w.head);
	hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(64), 0, s, w.head);
	hipLaunchKernelGGL(k_order_scatter, dim3(g), dim3(256), 0, s, d_skidx, nq, w.head, w.order);
}

int qj_workspace(DevBuf *ws, const QJShape &sh, const uint64_t *d_skidx, uint32_t nq, hipStream_t s, QJWork *w,
		 bool order = true, bool head = true)
{
	const size_t spill = std::max(ws->zero_bytes, sh.spill_bytes);
	const size_t want = WS_HEAD + spill + qj_order_bytes(nq);
	if (want > ws->bytes) {
		int rc = kssd_reserve_for_stream(ws, want, s, "query-major workspace");
		if (rc)
			return rc;
	}
	if (spill > ws->zero_bytes) {
		// kernels, but the cleared state must outlive the call: not in a capture
		hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
		if (s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
			return kssdhip_fail(KSSDHIP_EINVAL,
					    "query-major workspace: %zu spill bytes to clear while the stream is being "
					    "captured; run the call once outside the capture first",
					    spill);
		qj_zero((char *)ws->ptr + WS_HEAD, spill, s);
		ws->zero_bytes = spill;
	}
	w->head = (uint32_t *)ws->ptr;
	w->spill = (unsigned long long *)((char *)ws->ptr + WS_HEAD);
	w->order = (uint32_t *)((char *)ws->ptr + WS_HEAD + spill);
	qj_order(d_skidx, nq, s, *w, order, head);
	return 0;
}

void fill_index_args(const kssdhip_index *ix, QJArgs &q)
{
	q.brec = ix->d_brec;
	q.w = ix->br_w;
	q.bb = ix->br_bits;
	q.ctxgid = ix->d_ctxgid;
	q.robj = ix->d_obj;
	q.n = ix->n;
	q.ctx_max = ix->ctx_max;
	q.cmax = ix->br_w ? std::min<uint64_t>(ix->ctx_max, (1ull << ix->br_w) - 1ull) : 0ull;
	q.R = (uint32_t)ix->ref_gnum;
	q.obj_bits = ix->obj_bits;
}

int qj_launch(const QJShape &sh, bool uniq, const QJArgs &a, hipStream_t s)
{
	const dim3 grid(sh.grid), block(sh.threads);
	const bool items = a.chunk != 0;
	if (items) {
		if (uniq && sh.direct)
			hipLaunchKernelGGL((k_qjoin<true, true, true>), grid, block, sh.lds, s, a);
		else if (uniq)
			hipLaunchKernelGGL((k_qjoin<true, false, true>), grid, block, sh.lds, s, a);
		else if (sh.direct)
			hipLaunchKernelGGL((k_qjoin<false, true, true>), grid, block, sh.lds, s, a);
		else
			hipLaunchKernelGGL((k_qjoin<false, false, true>), grid, block, sh.lds, s, a);
	} else if (uniq && sh.direct)
		hipLaunchKernelGGL((k_qjoin<true, true, false>), grid, block, sh.lds, s, a);
	else if (uniq)
		hipLaunchKernelGGL((k_qjoin<true, false, false>), grid, block, sh.lds, s, a);
	else if (sh.direct)
		hipLaunchKernelGGL((k_qjoin<false, true, false>), grid, block, sh.lds, s, a);
	else
		hipLaunchKernelGGL((k_qjoin<false, false, false>), grid, block, sh.lds, s, a);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "qjoin launch: %s", hipGetErrorString(e));
	return 0;
}

} // namespace

__global__ void k_fill_empty(unsigned long long *p, uint64_t n)
{
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
		p[i] = BR_EMPTY;
}

// The record table of the index (see BR_SLOTS): contexts counted per home
// line, their first records listed per home, each home laid out by one
// thread (a count pass, a scan for the overflow lines, a write pass).  An
// index of 2^20 genomes (gid GID_MASK in use) or with a context of 2^30
// records or more gets no table; the query-major kernels then refuse it.
int kssd_qjoin_build(kssdhip_index *ix)
{
	kssd_qjoin_free(ix);
	int dev_cus = 0;
	if (hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ix->device) == hipSuccess)
		ix->cus = dev_cus;
	const uint64_t n = ix->n;
	const hipStream_t s = ix->stream;
	auto grid = [](uint64_t items) {
		return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + 255) / 256, 1u << 20)));
	};
	BrGeom g;
	g.ob = ix->obj_bits;
	g.w = br_ctx_width(ix->obj_bits);
	g.bb = BR_MIN_BITS; // about 3 slots per home line (few homes need an overflow line)
	while (g.bb < g.w && (3ull << g.bb) < n)
		++g.bb;
	if ((uint64_t)ix->ref_gnum > GID_MASK || g.bb > g.w)
		return kssd_qjoin_prepare(ix);
	g.homes = 1ull << g.bb;
	// temporaries: contexts per home, list offsets, cursors (then overflow
	// lines per home), the lists, scan sums, and three read-back words
	uint32_t *cnt = nullptr, *off = nullptr, *cur = nullptr, *list = nullptr, *sums = nullptr;
	unsigned long long *misc = nullptr;
	const uint64_t nsum = (g.homes + KSSD_SCAN_BLOCK - 1) / KSSD_SCAN_BLOCK;
	int rc = 0;
	do {
		if (hipMalloc(&cnt, g.homes * 4) != hipSuccess || hipMalloc(&off, g.homes * 4) != hipSuccess ||
		    hipMalloc(&cur, g.homes * 4) != hipSuccess ||
		    hipMalloc(&list, std::max<uint64_t>(n, 1) * 4) != hipSuccess ||
		    hipMalloc(&sums, nsum * 4) != hipSuccess || hipMalloc(&misc, 3 * 8) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_ENOMEM, "record table: device allocation failed");
			break;
		}
		hipLaunchKernelGGL(k_zero_u32, grid(g.homes), dim3(256), 0, s, cnt, (uint32_t)g.homes);
		hipLaunchKernelGGL(k_zero_u32, grid(g.homes), dim3(256), 0, s, cur, (uint32_t)g.homes);
		hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, misc, (uint64_t)3);
		if (n) {
			hipLaunchKernelGGL(k_br_longest, grid(n), dim3(256), 0, s, ix->d_ctxgid, n, misc);
			hipLaunchKernelGGL(k_br_count, grid(n), dim3(256), 0, s, ix->d_ctxgid, n, g, cnt);
		}
		if (hipMemcpyAsync(off, cnt, g.homes * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "record table: copy failed");
			break;
		}
		kssd_scan_u32_excl(off, g.homes, sums, s);
		if (n)
			hipLaunchKernelGGL(k_br_scatter, grid(n), dim3(256), 0, s, ix->d_ctxgid, n, g, off, cur, list);
		// overflow lines per home into cur, their total: last count + its offset
		hipLaunchKernelGGL((k_br_place<false>), grid(g.homes), dim3(256), 0, s, ix->d_ctxgid, ix->d_obj, n, g, off,
				   cnt, list, cur, nullptr, (uint64_t)0, nullptr);
		if (hipMemcpyAsync(&misc[1], cur + (g.homes - 1), 4, hipMemcpyDeviceToDevice, s) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "record table: copy failed");
			break;
		}
		kssd_scan_u32_excl(cur, g.homes, sums, s);
		unsigned long long back[3] = {0, 0, 0};
		if (hipMemcpyAsync(&misc[2], cur + (g.homes - 1), 4, hipMemcpyDeviceToDevice, s) != hipSuccess ||
		    hipMemcpyAsync(back, misc, sizeof back, hipMemcpyDeviceToHost, s) != hipSuccess ||
		    hipStreamSynchronize(s) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "record table: %s", hipGetErrorString(hipGetLastError()));
			break;
		}
		if (back[0] > BR_MAX_LONG)
			break; // a context too long for the payload: no table
		const uint64_t ovf = (back[1] & 0xffffffffull) + (back[2] & 0xffffffffull);
		const uint64_t lines = g.homes + ovf;
		if (hipMalloc(&ix->d_brec, lines * BR_SLOTS * 8) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_ENOMEM, "record table: device allocation of %llu lines failed",
					  (unsigned long long)lines);
			break;
		}
		hipLaunchKernelGGL(k_fill_empty, grid(lines * BR_SLOTS), dim3(256), 0, s, ix->d_brec, lines * BR_SLOTS);
		hipLaunchKernelGGL((k_br_place<true>), grid(g.homes), dim3(256), 0, s, ix->d_ctxgid, ix->d_obj, n, g, off,
				   cnt, list, nullptr, cur, g.homes, ix->d_brec);
		const hipError_t e = hipGetLastError();
		if (e != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "record table: %s", hipGetErrorString(e));
			break;
		}
		ix->br_lines = lines;
		ix->br_ovf = ovf;
		ix->br_w = g.w;
		ix->br_bits = g.bb;
	} while (0);
	for (void *p : {(void *)cnt, (void *)off, (void *)cur, (void *)list, (void *)sums, (void *)misc})
		if (p)
			(void)hipFree(p);
	if (rc) {
		kssd_qjoin_free(ix);
		return rc;
	}
	return kssd_qjoin_prepare(ix);
}

void kssd_qjoin_free(kssdhip_index *ix)
{
	if (ix->d_brec)
		(void)hipFree(ix->d_brec);
	ix->d_brec = nullptr;
	ix->br_lines = 0;
	ix->br_ovf = 0;
}

int kssd_qjoin_prepare(kssdhip_index *ix)
{
	for (int u = 0; u < 2; ++u) {
		QJShape sh;
		qj_layout(ix, u != 0, &sh);
		const void *fn = qj_fn(u != 0, sh.direct, false);
		const void *fn_items = qj_fn(u != 0, sh.direct, true);
		ix->qj_per_cu[u] = 0;
		if (sh.lds > 65536 &&
		    (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds) != hipSuccess ||
		     hipFuncSetAttribute(fn_items, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds) != hipSuccess))
			continue; // does not fit: launches with this tiling fail with EINVAL
		int per_cu = 0;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, sh.threads, sh.lds) == hipSuccess)
			ix->qj_per_cu[u] = per_cu;
	}
	(void)hipGetLastError();
	return 0;
}

bool kssd_qjoin_auto(const kssdhip_index *ix, int nq, uint64_t total)
{
	// the query-major join over the record table, at every block shape: at
	// the handles' default shapes it counted the 60-genome self matrix in
	// 45 us (partitioned merge-join 320 us), 1,000 x 1,000 clade genomes in
	// 0.45 ms (2.4 ms) and one assembly against 10,000 genomes in 26 us
	// (lookup kernel 375 us); indexes without a record table keep the
	// other kernels
	return ix->d_brec && nq >= 1 && total > 0;
}

int kssd_launch_qjoin_dense(kssdhip_index *ix, const JoinArgs &ja, hipStream_t s, DevBuf *ws)
{
	const bool uniq = (ja.flags & KSSDHIP_UNIQUE) && !(ja.flags & KSSDHIP_LOWER);
	if (!ix->d_brec)
		return kssdhip_fail(KSSDHIP_EINVAL, "qjoin: the index has no record table");
	QJShape sh;
	int rc = qj_shape(ix, (uint32_t)ja.nq, uniq, &sh);
	if (rc)
		return rc;
	// fewer queries than twice the resident workgroups: the queries' keys
	// are cut into key ranges that the workgroups take from a cursor
	// (guided sizes, at least QJ_MIN_CHUNK keys), so every workgroup has
	// work until the end; each range adds its table into the rows
	const uint64_t full = (uint64_t)ix->qj_per_cu[uniq ? 1 : 0] * (uint64_t)(ix->cus > 0 ? ix->cus : 256);
	uint32_t chunk = 0;
	if ((uint64_t)ja.nq < 2 * full && ja.total > 0) {
		chunk = (uint32_t)QJ_MIN_CHUNK;
		sh.grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(full, (ja.total + chunk - 1) / chunk));
		sh.spill_bytes = sh.direct ? 0 : (size_t)sh.grid * ix->ref_gnum * (uniq ? 32u : 16u);
	}
	// key-range items into dense outputs: the head is cleared with the planes
	const bool zero_planes = chunk && !(ja.flags & KSSDHIP_ACCUMULATE);
	QJWork w;
	if ((rc = qj_workspace(ws, sh, ja.skidx, (uint32_t)ja.nq, s, &w, chunk == 0, !zero_planes)))
		return rc;
	QJArgs a;
	memset(&a, 0, sizeof a);
	fill_index_args(ix, a);
	a.keys = ja.keys;
	a.skidx = ja.skidx;
	a.nq = (uint32_t)ja.nq;
	a.qry_gid_offset = ja.qry_gid_offset;
	a.flags = ja.flags;
	a.slots = sh.slots;
	a.sbits = sh.sbits;
	a.next = w.head;
	a.order = w.order;
	a.spill = w.spill;
	a.mode = (ja.flags & KSSDHIP_ACCUMULATE) ? QJ_ACCUM : QJ_DENSE;
	a.ctx = ja.ctx;
	a.objc = ja.objc;
	a.uctx = uniq ? ja.uctx : nullptr;
	a.uobj = uniq ? ja.uobj : nullptr;
	if (chunk) {
		if (zero_planes) { // the output planes and the head, in one launch
			const uint64_t words = (uint64_t)ja.nq * (uint64_t)ix->ref_gnum;
			const unsigned gx = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((words + 1023) / 1024, 8192));
			hipLaunchKernelGGL(k_zero_planes, dim3(gx, uniq ? 4u : 2u), dim3(256), 0, s, a.ctx, a.objc,
					   uniq ? a.uctx : nullptr, uniq ? a.uobj : nullptr, words, w.head, WS_HEAD_WORDS);
		}
		a.mode = QJ_ACCUM;
		a.chunk = chunk;
	}
	return qj_launch(sh, uniq, a, s);
}

namespace {

struct SelectCall {
	const uint64_t *keys, *skidx;
	int nq;
	uint64_t total;
	uint32_t flags;
	const uint32_t *qry_ctx, *ref_ctx;
	const kssdhip_select_opts_t *opts;
	kssdhip_ani_row_t *rows;
	uint32_t cap;
	uint32_t *row_count, *q_first, *q_count;
	hipStream_t s;
};

int select_check(const kssdhip_index *ix, const SelectCall &c)
{
	if (!ix || !c.opts || c.nq < 0 || !c.row_count)
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: bad arguments");
	if (c.flags & ~(KSSDHIP_UNIQUE | KSSDHIP_IGNORECONFLICT))
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: flags 0x%x (UNIQUE and IGNORECONFLICT only)", c.flags);
	if (c.opts->ctxcut < 1)
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: ctxcut must be >= 1");
	if (c.opts->unique_rank && !(c.flags & KSSDHIP_UNIQUE))
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: unique_rank needs KSSDHIP_UNIQUE");
	if (!epi_metric_supported(c.opts->value))
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: value metric %d needs the calibration model (use 3..9)",
				    c.opts->value.metric);
	if (c.nq > 0 && (!c.skidx || !c.qry_ctx || !c.ref_ctx || !c.q_first || !c.q_count || (c.cap && !c.rows) ||
			 (c.total && !c.keys)))
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: null device pointer");
	if (!ix->d_brec)
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: the index has no record table");
	return 0;
}

// the join in select mode and the rows' values, with the work order built
int select_launch(const kssdhip_index *ix, const SelectCall &c, const QJShape &sh, const QJWork &w)
{
	const bool uniq = (c.flags & KSSDHIP_UNIQUE) != 0;
	QJArgs a;
	memset(&a, 0, sizeof a);
	fill_index_args(ix, a);
	a.keys = c.keys;
	a.skidx = c.skidx;
	a.nq = (uint32_t)c.nq;
	a.flags = c.flags;
	a.slots = sh.slots;
	a.sbits = sh.sbits;
	a.next = w.head;
	a.order = w.order;
	a.spill = w.spill;
	a.mode = QJ_SELECT;
	a.qry_ctx = c.qry_ctx;
	a.ref_ctx = c.ref_ctx;
	a.ctxcut = c.opts->ctxcut;
	a.afcut = (double)c.opts->afcut;
	a.anicut = c.opts->anicut;
	a.unique_rank = c.opts->unique_rank;
	a.vo = c.opts->value;
	{
		int m = a.vo.metric;
		if (m < 1 || m > 9)
			m = 1;
		a.vmetric = (a.vo.unassembled && m <= 4) ? 4 : m;
	}
	a.rows = c.rows;
	a.cap = c.cap;
	a.row_count = c.row_count;
	a.q_first = c.q_first;
	a.q_count = c.q_count;
	int rc = qj_launch(sh, uniq, a, c.s);
	if (rc)
		return rc;
	if (c.cap) {
		const unsigned g = (unsigned)std::max<uint32_t>(1, std::min<uint32_t>((c.cap + 255) / 256, 1024));
		hipLaunchKernelGGL(k_row_values, dim3(g), dim3(256), 0, c.s, c.rows, c.row_count, c.cap, c.qry_ctx,
				   c.ref_ctx, a.vo, a.vmetric);
		hipError_t e = hipGetLastError();
		if (e != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "select_block values: %s", hipGetErrorString(e));
	}
	return 0;
}

// Small blocks (fewer queries than compute units) leave most of the chip to
// a few query-major workgroups; they are counted into dense rows by the
// lookup kernel instead (the handle's tiling unless it forces the
// query-major join) and gated row by row.  Dense rows up to this size.
constexpr uint64_t DENSE_SELECT_MAX_BYTES = 256ull << 20;

bool select_dense_fits(const kssdhip_index *ix, int nq, bool uniq)
{
	const uint64_t cells = (uint64_t)nq * (uint64_t)ix->ref_gnum;
	return ix->mode != 3 && nq < (ix->cus > 0 ? ix->cus : 256) && cells * (uniq ? 32u : 16u) <= DENSE_SELECT_MAX_BYTES;
}

int select_dense(kssdhip_index *ix, const SelectCall &c, DevBuf *split, DevBuf *dws)
{
	const bool uniq = (c.flags & KSSDHIP_UNIQUE) != 0;
	const uint64_t cells = (uint64_t)c.nq * (uint64_t)ix->ref_gnum;
	const size_t plane = (size_t)cells * 8;
	const size_t want = plane * (uniq ? 4 : 2) + (size_t)c.nq * 4 + 256;
	int rc = kssd_reserve_for_stream(dws, want, c.s, "select workspace");
	if (rc)
		return rc;
	char *base = (char *)dws->ptr;
	kssd_ctx_mut2_t *d_ctx = (kssd_ctx_mut2_t *)base;
	kssd_obj_section_t *d_obj = (kssd_obj_section_t *)(base + plane);
	kssd_ctx_mut2_t *d_uctx = uniq ? (kssd_ctx_mut2_t *)(base + 2 * plane) : nullptr;
	kssd_obj_section_t *d_uobj = uniq ? (kssd_obj_section_t *)(base + 3 * plane) : nullptr;
	uint32_t *cursor = (uint32_t *)(base + plane * (uniq ? 4 : 2));
	if ((rc = kssd_count_dev(ix, c.keys, c.skidx, c.nq, c.total, 0, c.flags, d_ctx, d_obj, d_uctx, d_uobj, c.s,
				 split, nullptr)))
		return rc;
	QJArgs a;
	memset(&a, 0, sizeof a);
	a.nq = (uint32_t)c.nq;
	a.R = (uint32_t)ix->ref_gnum;
	a.ctx = (unsigned long long *)d_ctx;
	a.objc = (unsigned long long *)d_obj;
	a.uctx = (unsigned long long *)d_uctx;
	a.uobj = (unsigned long long *)d_uobj;
	a.qry_ctx = c.qry_ctx;
	a.ref_ctx = c.ref_ctx;
	a.ctxcut = c.opts->ctxcut;
	a.afcut = (double)c.opts->afcut;
	a.anicut = c.opts->anicut;
	a.unique_rank = c.opts->unique_rank;
	a.rows = c.rows;
	a.cap = c.cap;
	a.row_count = c.row_count;
	a.q_first = c.q_first;
	a.q_count = c.q_count;
	const unsigned g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((cells + 255) / 256, 8192));
	hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)std::min<uint32_t>(((uint32_t)c.nq + 255) / 256, 1024u)),
			   dim3(256), 0, c.s, c.q_count, (uint32_t)c.nq);
	if (uniq)
		hipLaunchKernelGGL(k_gate_dense_count<true>, dim3(g), dim3(256), 0, c.s, a);
	else
		hipLaunchKernelGGL(k_gate_dense_count<false>, dim3(g), dim3(256), 0, c.s, a);
	hipLaunchKernelGGL(k_gate_dense_scan, dim3(1), dim3(64), 0, c.s, a, cursor);
	if (uniq)
		hipLaunchKernelGGL(k_gate_dense_write<true>, dim3(g), dim3(256), 0, c.s, a, cursor);
	else
		hipLaunchKernelGGL(k_gate_dense_write<false>, dim3(g), dim3(256), 0, c.s, a, cursor);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "select_block gate: %s", hipGetErrorString(e));
	if (c.cap) {
		kssdhip_epilogue_opts_t vo = c.opts->value;
		int m = vo.metric;
		if (m < 1 || m > 9)
			m = 1;
		const int vmetric = (vo.unassembled && m <= 4) ? 4 : m;
		const unsigned gr = (unsigned)std::max<uint32_t>(1, std::min<uint32_t>((c.cap + 255) / 256, 1024));
		hipLaunchKernelGGL(k_row_values, dim3(gr), dim3(256), 0, c.s, c.rows, c.row_count, c.cap, c.qry_ctx,
				   c.ref_ctx, vo, vmetric);
		e = hipGetLastError();
		if (e != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "select_block values: %s", hipGetErrorString(e));
	}
	return 0;
}

// zeroes the row count; true when the block is empty (nothing else to do)
int select_begin(const SelectCall &c, bool *done)
{
	hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c.s, c.row_count, 1u);
	*done = c.nq == 0;
	if (*done) {
		hipError_t e = hipGetLastError();
		if (e != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "select_block: %s", hipGetErrorString(e));
	}
	return 0;
}

// The no-allocation workspace: one array per device, at fixed offsets:
// [0, WS_HEAD) the head, then the work order of up to
// KSSDHIP_NOALLOC_MAX_QUERIES queries, then the spill rows.  Host code
// never uses the array's name as an address: kssd_noalloc_setup looks the
// device copy up once per device and the entry point passes that pointer.
constexpr size_t NOALLOC_ORDER = (size_t)KSSDHIP_NOALLOC_MAX_QUERIES * 4;
constexpr size_t NOALLOC_SPILL_AT = WS_HEAD + NOALLOC_ORDER;
static_assert(NOALLOC_SPILL_AT % WS_ALIGN == 0, "spill rows stay aligned");
static_assert(NOALLOC_SPILL_AT < KSSDHIP_NOALLOC_SCRATCH_BYTES, "scratch holds the head and the order");
constexpr size_t NOALLOC_SPILL = KSSDHIP_NOALLOC_SCRATCH_BYTES - NOALLOC_SPILL_AT;

std::mutex g_scratch_mu;
std::vector<void *> g_scratch; // device copy of `scratch` per device (nullptr: not looked up)

} // namespace

static __device__ __align__(256) unsigned char scratch[KSSDHIP_NOALLOC_SCRATCH_BYTES];

int kssd_noalloc_setup(int device)
{
	if (device < 0)
		return kssdhip_fail(KSSDHIP_EINVAL, "scratch: device %d", device);
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if ((size_t)device < g_scratch.size() && g_scratch[device])
		return 0;
	void *ptr = nullptr;
	hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
	if (e != hipSuccess || !ptr)
		return kssdhip_fail(KSSDHIP_EHIP, "scratch: address lookup on device %d failed: %s", device,
				    hipGetErrorString(e));
	if ((size_t)device >= g_scratch.size())
		g_scratch.resize((size_t)device + 1, nullptr);
	g_scratch[device] = ptr;
	return 0;
}

// the device copy of `scratch` on `device`, looked up on first use with
// that device current (kssdhip_init does it for its device)
static int noalloc_scratch(int device, void **out)
{
	{
		std::lock_guard<std::mutex> lock(g_scratch_mu);
		if ((size_t)device < g_scratch.size() && g_scratch[device]) {
			*out = g_scratch[device];
			return 0;
		}
	}
	int prev = 0;
	if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "scratch: device %d not available", device);
	int rc = kssd_noalloc_setup(device);
	(void)hipSetDevice(prev);
	if (rc)
		return rc;
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	*out = g_scratch[device];
	return 0;
}

int kssd_select_dev_ws(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx, int nq, uint64_t total,
		       uint32_t flags, const uint32_t *d_qry_ctx, const uint32_t *d_ref_ctx,
		       const kssdhip_select_opts_t *opts, kssdhip_ani_row_t *d_rows, uint32_t cap,
		       uint32_t *d_row_count, uint32_t *d_q_first, uint32_t *d_q_count, hipStream_t s, DevBuf *ws,
		       DevBuf *split, DevBuf *dense)
{
	const SelectCall c{d_keys, d_skidx, nq, total, flags, d_qry_ctx, d_ref_ctx, opts,
			   d_rows, cap, d_row_count, d_q_first, d_q_count, s};
	int rc = select_check(ix, c);
	bool done = false;
	if (rc || (rc = select_begin(c, &done)) || done)
		return rc;
	if (split && dense && select_dense_fits(ix, nq, (flags & KSSDHIP_UNIQUE) != 0))
		return select_dense(ix, c, split, dense);
	QJShape sh;
	if ((rc = qj_shape(ix, (uint32_t)nq, (flags & KSSDHIP_UNIQUE) != 0, &sh)))
		return rc;
	QJWork w;
	if ((rc = qj_workspace(ws, sh, d_skidx, (uint32_t)nq, c.s, &w)))
		return rc;
	return select_launch(ix, c, sh, w);
}

extern "C" int kssdhip_select_block_dev(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx, int nq,
					uint64_t total, uint32_t flags, const uint32_t *d_qry_ctx,
					const uint32_t *d_ref_ctx, const kssdhip_select_opts_t *opts,
					kssdhip_ani_row_t *d_rows, uint32_t cap, uint32_t *d_row_count,
					uint32_t *d_q_first, uint32_t *d_q_count, void *stream)
{
	if (!ix)
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: bad arguments");
	return kssd_select_dev_ws(ix, d_keys, d_skidx, nq, total, flags, d_qry_ctx, d_ref_ctx, opts, d_rows, cap,
				  d_row_count, d_q_first, d_q_count, (hipStream_t)stream, &ix->qws, &ix->split, &ix->sdense);
}

extern "C" int kssdhip_select_block_dev_noalloc(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx,
						int nq, uint64_t total, uint32_t flags, const uint32_t *d_qry_ctx,
						const uint32_t *d_ref_ctx, const kssdhip_select_opts_t *opts,
						kssdhip_ani_row_t *d_rows, uint32_t cap, uint32_t *d_row_count,
						uint32_t *d_q_first, uint32_t *d_q_count, void *stream)
{
	const SelectCall c{d_keys, d_skidx, nq, total, flags, d_qry_ctx, d_ref_ctx, opts,
			   d_rows, cap, d_row_count, d_q_first, d_q_count, (hipStream_t)stream};
	int rc = select_check(ix, c);
	if (rc)
		return rc;
	if ((uint32_t)nq > KSSDHIP_NOALLOC_MAX_QUERIES)
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block_noalloc: %d queries; the scratch array holds the work "
				    "order of at most %u", nq, KSSDHIP_NOALLOC_MAX_QUERIES);
	const bool uniq = (flags & KSSDHIP_UNIQUE) != 0;
	QJShape sh;
	if (nq > 0 && (rc = qj_shape(ix, (uint32_t)nq, uniq, &sh)))
		return rc;
	if (nq > 0 && !sh.direct) {
		const size_t row = (size_t)ix->ref_gnum * (uniq ? 32u : 16u);
		const size_t fit = NOALLOC_SPILL / row;
		if (fit < 1)
			return kssdhip_fail(KSSDHIP_EINVAL, "select_block_noalloc: a spill row of %d references (%zu bytes) "
					    "exceeds the scratch array's %zu spill bytes", ix->ref_gnum, row,
					    NOALLOC_SPILL);
		if (sh.grid > fit)
			sh.grid = (unsigned)fit;
		sh.spill_bytes = (size_t)sh.grid * row;
	}
	void *base = nullptr;
	if ((rc = noalloc_scratch(ix->device, &base)))
		return rc;
	bool done = false;
	if ((rc = select_begin(c, &done)) || done)
		return rc;
	QJWork w;
	w.head = (uint32_t *)base;
	w.order = (uint32_t *)((char *)base + WS_HEAD);
	w.spill = (unsigned long long *)((char *)base + NOALLOC_SPILL_AT);
	if (!sh.direct)
		qj_zero(w.spill, sh.spill_bytes, c.s); // the last call's contents stay in the array
	qj_order(d_skidx, (uint32_t)nq, c.s, w, true);
	return select_launch(ix, c, sh, w);
}

extern "C" int kssdhip_select_block(kssdhip_index *ix, const uint64_t *section_sk, const uint64_t *section_skidx,
				    int nq, uint32_t flags, const uint32_t *qry_ctx, const uint32_t *ref_ctx,
				    const kssdhip_select_opts_t *opts, kssdhip_ani_row_t *rows, size_t cap,
				    size_t *n_rows)
{
	if (!ix || !opts || nq < 0 || !n_rows || (nq > 0 && (!section_skidx || !qry_ctx || !ref_ctx)) ||
	    (cap && !rows))
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: bad arguments");
	*n_rows = 0;
	if (nq == 0)
		return 0;
	const uint64_t base = section_skidx[0];
	std::vector<uint64_t> rel((size_t)nq + 1);
	for (int i = 0; i <= nq; ++i) {
		if (section_skidx[i] < base || (i && section_skidx[i] < section_skidx[i - 1]))
			return kssdhip_fail(KSSDHIP_EINVAL, "select_block: section_skidx not ascending");
		rel[i] = section_skidx[i] - base;
	}
	const uint64_t total = rel[nq];
	if (total && !section_sk)
		return kssdhip_fail(KSSDHIP_EINVAL, "select_block: null keys");
	const uint32_t dcap = (uint32_t)std::min<size_t>(cap, 0xffffffffu);
	if (hipSetDevice(ix->device) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "select_block: device not available");
	int rc;
	const size_t R = (size_t)ix->ref_gnum;
	if ((rc = ix->keys.reserve((total ? total : 1) * 8)) || (rc = ix->skidx.reserve(rel.size() * 8)) ||
	    (rc = ix->qctxb.reserve((size_t)nq * 4)) || (rc = ix->rctxb.reserve(R * 4)) ||
	    (rc = ix->rows.reserve((dcap ? dcap : 1) * sizeof(kssdhip_ani_row_t))) ||
	    (rc = ix->qmeta.reserve(((size_t)nq * 2 + 1) * 4)))
		return rc;
	hipStream_t s = ix->stream;
	if (total && copy_h2d(ix->keys.ptr, section_sk, total * 8, s) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "select_block: key upload failed");
	if (hipMemcpyAsync(ix->skidx.ptr, rel.data(), rel.size() * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
	    hipMemcpyAsync(ix->qctxb.ptr, qry_ctx, (size_t)nq * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
	    hipMemcpyAsync(ix->rctxb.ptr, ref_ctx, R * 4, hipMemcpyHostToDevice, s) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "select_block: upload failed");
	uint32_t *meta = (uint32_t *)ix->qmeta.ptr; // [0] row count, [1, nq] first, [nq + 1, 2 nq] count
	rc = kssdhip_select_block_dev(ix, (const uint64_t *)ix->keys.ptr, (const uint64_t *)ix->skidx.ptr, nq, total,
				      flags, (const uint32_t *)ix->qctxb.ptr, (const uint32_t *)ix->rctxb.ptr, opts,
				      (kssdhip_ani_row_t *)ix->rows.ptr, dcap, meta, meta + 1, meta + 1 + nq, s);
	if (rc)
		return rc;
	uint32_t count = 0;
	if (hipMemcpyAsync(&count, meta, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "select_block: %s", hipGetErrorString(hipGetLastError()));
	*n_rows = count;
	if (count > cap)
		return kssdhip_fail(KSSDHIP_ENOSPC, "select_block: %u passing pairs exceed the %zu-row buffer", count, cap);
	if (count) {
		if (copy_d2h(rows, ix->rows.ptr, (size_t)count * sizeof(kssdhip_ani_row_t), s) != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "select_block: row download failed");
		std::sort(rows, rows + count, [](const kssdhip_ani_row_t &x, const kssdhip_ani_row_t &y) {
			return x.query != y.query ? x.query < y.query : x.ref < y.ref;
		});
	}
	return 0;
}

extern "C" int kssdhip_index_matched_records(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx,
					     int nq, uint64_t total, uint64_t *matched)
{
	if (!ix || nq < 0 || !matched || (nq > 0 && !d_skidx) || (total && !d_keys))
		return kssdhip_fail(KSSDHIP_EINVAL, "matched_records: bad arguments");
	*matched = 0;
	if (!nq || !total || !ix->n)
		return 0;
	if (!ix->d_brec)
		return kssdhip_fail(KSSDHIP_EINVAL, "matched_records: the index has no record table");
	if (hipSetDevice(ix->device) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "matched_records: device not available");
	DevBuf marks;
	const uint64_t slots = ix->br_lines * BR_SLOTS;
	const uint64_t words = (slots + 7) / 8 + 1; // a mark per table slot, then the count (8-byte aligned)
	int rc = marks.reserve(words * 8);
	if (rc)
		return rc;
	hipStream_t s = ix->stream;
	unsigned long long *d_count = (unsigned long long *)marks.ptr + (words - 1);
	QJArgs a;
	memset(&a, 0, sizeof a);
	fill_index_args(ix, a);
	a.keys = d_keys;
	a.skidx = d_skidx;
	a.nq = (uint32_t)nq;
	hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<uint64_t>((words + 255) / 256, 8192)), dim3(256), 0, s,
			   (unsigned long long *)marks.ptr, words);
	hipLaunchKernelGGL(k_mark_matched, dim3((unsigned)std::min<uint64_t>((total + 255) / 256, 1u << 16)), dim3(256),
			   0, s, a, (uint8_t *)marks.ptr);
	hipLaunchKernelGGL(k_count_marks, dim3((unsigned)std::min<uint64_t>((slots + 255) / 256, 8192)), dim3(256), 0, s,
			   (const uint8_t *)marks.ptr, (const unsigned long long *)ix->d_brec, ix->obj_bits, slots, d_count);
	unsigned long long count = 0;
	hipError_t e = hipGetLastError();
	if (e == hipSuccess)
		e = hipMemcpyAsync(&count, d_count, 8, hipMemcpyDeviceToHost, s);
	if (e == hipSuccess)
		e = hipStreamSynchronize(s);
	marks.release();
	if (e != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "matched_records: %s", hipGetErrorString(e));
	*matched = count;
	return 0;
}

extern "C" int kssdhip_index_table_info(const kssdhip_index *ix, uint64_t *lines, uint64_t *overflow_lines,
					int *home_bits, int *ctx_bits)
{
	if (!ix || !lines || !overflow_lines || !home_bits || !ctx_bits)
		return kssdhip_fail(KSSDHIP_EINVAL, "index_table_info: bad arguments");
	*lines = ix->d_brec ? ix->br_lines : 0;
	*overflow_lines = ix->d_brec ? ix->br_ovf : 0;
	*home_bits = ix->d_brec ? ix->br_bits : 0;
	*ctx_bits = ix->d_brec ? ix->br_w : 0;
	return 0;
}
