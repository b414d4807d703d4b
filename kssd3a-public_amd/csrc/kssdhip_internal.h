// Shared between the translation units of libkssdhip (not part of the ABI).
#ifndef KSSDHIP_INTERNAL_H
#define KSSDHIP_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace kssd {
constexpr int GID_NBITS = 20;                         // sketch_rearrange.h:7
constexpr uint64_t GID_MASK = (1ull << GID_NBITS) - 1ull;
constexpr int MIN_INIT = 12;                          // NUM_CODENS + 1, command_ani.h:535

struct JoinArgs {
	const uint64_t *ctxgid;
	const uint32_t *robj;
	const uint32_t *bucket;
	uint64_t n;
	int shift;
	uint64_t ctx_max;
	int ref_gnum;
	const uint64_t *keys;
	const uint64_t *skidx;
	int nq;
	uint64_t total;
	int qry_gid_offset;
	unsigned obj_bits;
	unsigned flags;
	unsigned long long *ctx;
	unsigned long long *objc;
	unsigned long long *uctx;
	unsigned long long *uobj;
};
// Workspace of the one-launch plain pair kernel: per-block partial sums and
// the count of finished blocks (left at zero by every launch).
constexpr unsigned PAIR_MAX_BLOCKS = 1024;
struct PairWorkspace {
	unsigned done;
	unsigned pad[3];
	uint4 partial[PAIR_MAX_BLOCKS];
};
} // namespace kssd

// grow-only device buffer
struct DevBuf {
	void *ptr = nullptr;
	size_t bytes = 0;
	int reserve(size_t want);
	void release();
};

// read-only view of an index handle for other translation units
struct IndexView {
	const uint64_t *ctxgid;
	const uint32_t *obj;
	const uint32_t *bucket;
	uint64_t n;
	int ref_gnum;
	unsigned obj_bits;
	int shift;
	int k;
	uint64_t ctx_max;
	int device;
};
struct kssdhip_index;
int kssdhip_index_info(const kssdhip_index *ix, IndexView *v);
int kssdhip_fail(int code, const char *fmt, ...);

#endif
