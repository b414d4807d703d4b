// Serving loop over a resident index: continuous batching with admission by
// a device-memory budget.
//
// Requests (one query sketch each) are queued by kssdhip_serve_submit.  A
// dispatcher thread forms a batch from the head of the queue whenever one of
// the server's slots is free: a slot owns a HIP stream, a device arena and a
// pinned host arena of budget/slots bytes, and a batch is admitted while its
// keys, offsets and dense counter rows fit that arena.  The batch is copied
// in, counted (the same kernels as kssdhip_count_block_dev, with the slot's
// own split-table workspace so slots never share scratch), and its counter
// rows are copied back, all asynchronously on the slot's stream; requests
// that arrive meanwhile form the next batch on another slot, so batches
// overlap on the device.  A completion thread waits for the slots in launch
// order and hands each request its rows.  A server opened with `select`
// runs the device pass gate instead (the query-major join in select mode,
// with the slot's own workspace), copies back the row count and per-query
// segments, then only the passing rows, and ranks each request's rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_epi.h"
#include "kssdhip_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

size_t align_up(size_t x)
{
	return (x + 255) & ~(size_t)255;
}

struct Request {
	int64_t ticket = 0;
	std::vector<uint64_t> keys;
	Clock::time_point t_submit, t_launch, t_done;
	int state = 0; // 0 queued, 1 running, 2 done, 3 failed
	int err = 0;
	std::string msg;
	int batch_queries = 0;
	std::vector<unsigned long long> ctx, obj, uctx, uobj; // ref_gnum cells each
	uint32_t qctx = 0;                                    // select: distinct contexts of the keys
	std::vector<kssdhip_ani_row_t> rows;                  // select: ranked passing rows
};

// arena layout of a batch: counter rows [keys | skidx | ctx | obj (| uctx |
// uobj)]; select [keys | skidx | qctx | meta: count, first[nq], count[nq] |
// rows (nq * R)]
struct Batch {
	std::vector<Request *> reqs;
	size_t off_skidx = 0, off_ctx = 0, off_obj = 0, off_uctx = 0, off_uobj = 0, end = 0;
	size_t off_qctx = 0, off_meta = 0, off_rows = 0;
	uint64_t total = 0;
};

struct Slot {
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	void *d_arena = nullptr;
	void *h_arena = nullptr;
	DevBuf split;
	DevBuf qws;    // select: the query-major join's workspace
	DevBuf sdense; // select: dense rows of small batches
	Batch batch;
	bool busy = false;
	int launch_err = 0;
	std::string launch_msg;
};

} // namespace

struct kssdhip_server {
	kssdhip_index *ix = nullptr;
	kssdhip_serve_opts_t opts{};
	int device = 0;
	int R = 0;
	unsigned obj_bits = 0;
	bool unique = false;
	bool results = false;
	kssdhip_select_opts_t sel{};
	uint32_t *d_ref_ctx = nullptr;
	int ntop = 0;
	size_t slot_cap = 0;
	std::mutex mu;
	std::condition_variable cv_disp, cv_done, cv_complete;
	std::deque<Request *> queue;
	std::unordered_map<int64_t, Request *> reqs;
	std::vector<Slot> slots;
	std::deque<int> inflight;
	std::thread dispatcher, completer;
	bool stop = false;
	bool dispatcher_done = false;
	int64_t next_ticket = 0;
	kssdhip_serve_info_t info{};
};

namespace {

// bytes of a batch's arena layout; fills the offsets
size_t layout(Batch &b, int R, bool unique, bool results)
{
	size_t nkeys = 0;
	for (const Request *r : b.reqs)
		nkeys += r->keys.size();
	const size_t nq = b.reqs.size(), cells = nq * (size_t)R;
	b.total = nkeys;
	b.off_skidx = align_up(nkeys * 8);
	if (results) {
		b.off_qctx = b.off_skidx + align_up((nq + 1) * 8);
		b.off_meta = b.off_qctx + align_up(nq * 4);
		b.off_rows = b.off_meta + align_up((1 + 2 * nq) * 4);
		b.end = b.off_rows + align_up(cells * sizeof(kssdhip_ani_row_t));
		return b.end;
	}
	b.off_ctx = b.off_skidx + align_up((nq + 1) * 8);
	b.off_obj = b.off_ctx + align_up(cells * 8);
	b.end = b.off_obj + align_up(cells * 8);
	if (unique) {
		b.off_uctx = b.end;
		b.off_uobj = b.off_uctx + align_up(cells * 8);
		b.end = b.off_uobj + align_up(cells * 8);
	}
	return b.end;
}

size_t request_bytes(size_t nkeys, int R, bool unique, bool results)
{
	if (results)
		return align_up(nkeys * 8) + align_up(16) + align_up(4) + align_up(12) +
		       align_up((size_t)R * sizeof(kssdhip_ani_row_t));
	return align_up(nkeys * 8) + align_up(16) + (unique ? 4 : 2) * align_up((size_t)R * 8);
}

// distinct contexts among sorted keys
uint32_t count_contexts(const std::vector<uint64_t> &k, unsigned obj_bits)
{
	uint32_t n = 0;
	for (size_t i = 0; i < k.size(); ++i)
		n += (i == 0 || (k[i] >> obj_bits) != (k[i - 1] >> obj_bits)) ? 1u : 0u;
	return n;
}

void fail_batch(Batch &b, int err, const std::string &msg)
{
	for (Request *r : b.reqs) {
		r->state = 3;
		r->err = err;
		r->msg = msg;
		r->t_done = Clock::now();
	}
}

// Copy the batch in, count, copy the rows out; all on the slot's stream.
int launch(kssdhip_server *s, Slot &sl)
{
	Batch &b = sl.batch;
	char *h = (char *)sl.h_arena;
	char *d = (char *)sl.d_arena;
	uint64_t *hk = (uint64_t *)h, *hs = (uint64_t *)(h + b.off_skidx);
	uint64_t at = 0;
	for (size_t i = 0; i < b.reqs.size(); ++i) {
		const std::vector<uint64_t> &k = b.reqs[i]->keys;
		hs[i] = at;
		if (!k.empty())
			std::memcpy(hk + at, k.data(), k.size() * 8);
		at += k.size();
	}
	hs[b.reqs.size()] = at;
	const int nq = (int)b.reqs.size();
	if (s->results) {
		uint32_t *hq = (uint32_t *)(h + b.off_qctx);
		for (int i = 0; i < nq; ++i)
			hq[i] = b.reqs[i]->qctx;
		if (hipMemcpyAsync(d, h, b.off_meta, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "serve: copy in failed");
		uint32_t *meta = (uint32_t *)(d + b.off_meta);
		const uint64_t cap = std::min<uint64_t>((uint64_t)nq * (uint64_t)s->R, 0xffffffffu);
		int rc = kssd_select_dev_ws(s->ix, (const uint64_t *)d, (const uint64_t *)(d + b.off_skidx), nq, b.total,
					    s->opts.flags, (const uint32_t *)(d + b.off_qctx), s->d_ref_ctx, &s->sel,
					    (kssdhip_ani_row_t *)(d + b.off_rows), (uint32_t)cap, meta, meta + 1, meta + 1 + nq,
					    sl.stream, &sl.qws, &sl.split, &sl.sdense);
		if (rc)
			return rc;
		if (hipMemcpyAsync(h + b.off_meta, d + b.off_meta, b.off_rows - b.off_meta, hipMemcpyDeviceToHost,
				   sl.stream) != hipSuccess ||
		    hipEventRecord(sl.done, sl.stream) != hipSuccess)
			return kssdhip_fail(KSSDHIP_EHIP, "serve: copy out failed");
		return 0;
	}
	if (hipMemcpyAsync(d, h, b.off_ctx, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "serve: copy in failed");
	int rc = kssd_count_dev(s->ix, (const uint64_t *)d, (const uint64_t *)(d + b.off_skidx), nq, b.total, 0,
				s->opts.flags, (kssd_ctx_mut2_t *)(d + b.off_ctx), (kssd_obj_section_t *)(d + b.off_obj),
				s->unique ? (kssd_ctx_mut2_t *)(d + b.off_uctx) : nullptr,
				s->unique ? (kssd_obj_section_t *)(d + b.off_uobj) : nullptr, sl.stream, &sl.split, nullptr);
	if (rc)
		return rc;
	if (hipMemcpyAsync(h + b.off_ctx, d + b.off_ctx, b.end - b.off_ctx, hipMemcpyDeviceToHost, sl.stream) !=
		    hipSuccess ||
	    hipEventRecord(sl.done, sl.stream) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "serve: copy out failed");
	return 0;
}

void dispatch_loop(kssdhip_server *s)
{
	(void)hipSetDevice(s->device);
	std::unique_lock<std::mutex> lk(s->mu);
	for (;;) {
		int sid = -1;
		s->cv_disp.wait(lk, [&] {
			if (s->stop)
				return true;
			if (s->queue.empty())
				return false;
			for (size_t i = 0; i < s->slots.size(); ++i)
				if (!s->slots[i].busy) {
					sid = (int)i;
					return true;
				}
			return false;
		});
		if (s->stop) {
			s->dispatcher_done = true;
			s->cv_complete.notify_all();
			return;
		}
		Slot &sl = s->slots[sid];
		Batch &b = sl.batch;
		b = Batch();
		// admission: FIFO while the batch fits the slot's arena
		while (!s->queue.empty() && (int)b.reqs.size() < s->opts.max_batch) {
			Request *r = s->queue.front();
			b.reqs.push_back(r);
			if (layout(b, s->R, s->unique, s->results) > s->slot_cap) {
				b.reqs.pop_back();
				break;
			}
			s->queue.pop_front();
		}
		layout(b, s->R, s->unique, s->results);
		const Clock::time_point now = Clock::now();
		for (Request *r : b.reqs) {
			r->state = 1;
			r->t_launch = now;
			r->batch_queries = (int)b.reqs.size();
		}
		sl.busy = true;
		s->info.batches += 1;
		s->info.queries += (int64_t)b.reqs.size();
		if ((int)b.reqs.size() > s->info.max_batch_seen)
			s->info.max_batch_seen = (int)b.reqs.size();
		lk.unlock();
		sl.launch_err = launch(s, sl);
		if (sl.launch_err)
			sl.launch_msg = kssdhip_last_error();
		lk.lock();
		s->inflight.push_back(sid);
		s->cv_complete.notify_one();
	}
}

// select: the passing rows of a finished batch (after its event), copied
// back and split into the requests' ranked lists
int collect_rows(kssdhip_server *s, Slot &sl, std::string *msg)
{
	Batch &b = sl.batch;
	const size_t nq = b.reqs.size();
	const char *h = (const char *)sl.h_arena;
	const uint32_t *meta = (const uint32_t *)(h + b.off_meta);
	const uint32_t n = meta[0];
	if ((uint64_t)n > (uint64_t)nq * (uint64_t)s->R) {
		*msg = "serve: row count beyond the batch's pairs";
		return KSSDHIP_EHIP;
	}
	if (n && (hipMemcpyAsync((char *)sl.h_arena + b.off_rows, (const char *)sl.d_arena + b.off_rows,
				 (size_t)n * sizeof(kssdhip_ani_row_t), hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
		  hipStreamSynchronize(sl.stream) != hipSuccess)) {
		*msg = "serve: row copy failed";
		return KSSDHIP_EHIP;
	}
	const kssdhip_ani_row_t *rows = (const kssdhip_ani_row_t *)(h + b.off_rows);
	for (size_t i = 0; i < nq; ++i) {
		const uint32_t first = meta[1 + i], cnt = meta[1 + nq + i];
		if ((uint64_t)first + cnt > n) {
			*msg = "serve: row segment beyond the row count";
			return KSSDHIP_EHIP;
		}
		std::vector<kssdhip_ani_row_t> &v = b.reqs[i]->rows;
		v.assign(rows + first, rows + first + cnt);
		for (kssdhip_ani_row_t &r : v)
			r.query = 0;
		std::sort(v.begin(), v.end(), [](const kssdhip_ani_row_t &x, const kssdhip_ani_row_t &y) {
			return x.value != y.value ? x.value > y.value : x.ref < y.ref;
		});
		if (s->ntop > 0 && v.size() > (size_t)s->ntop)
			v.resize((size_t)s->ntop);
	}
	return 0;
}

void complete_loop(kssdhip_server *s)
{
	(void)hipSetDevice(s->device);
	std::unique_lock<std::mutex> lk(s->mu);
	for (;;) {
		s->cv_complete.wait(lk, [&] { return !s->inflight.empty() || (s->stop && s->dispatcher_done); });
		if (s->inflight.empty())
			return; // stopped, and every launched batch has been handed back
		const int sid = s->inflight.front();
		Slot &sl = s->slots[sid];
		lk.unlock();
		int err = sl.launch_err;
		std::string msg = sl.launch_msg;
		if (err)
			(void)hipStreamSynchronize(sl.stream); // let any copy already queued finish
		else if (hipEventSynchronize(sl.done) != hipSuccess) {
			err = KSSDHIP_EHIP;
			msg = "serve: device work failed";
		}
		Batch &b = sl.batch;
		if (!err && s->results)
			err = collect_rows(s, sl, &msg);
		const Clock::time_point now = Clock::now();
		if (!err && s->results) {
			for (Request *r : b.reqs)
				r->t_done = now;
		} else if (!err) {
			const char *h = (const char *)sl.h_arena;
			const size_t row = (size_t)s->R;
			for (size_t i = 0; i < b.reqs.size(); ++i) {
				Request *r = b.reqs[i];
				const unsigned long long *c = (const unsigned long long *)(h + b.off_ctx) + i * row;
				const unsigned long long *o = (const unsigned long long *)(h + b.off_obj) + i * row;
				r->ctx.assign(c, c + row);
				r->obj.assign(o, o + row);
				if (s->unique) {
					const unsigned long long *uc = (const unsigned long long *)(h + b.off_uctx) + i * row;
					const unsigned long long *uo = (const unsigned long long *)(h + b.off_uobj) + i * row;
					r->uctx.assign(uc, uc + row);
					r->uobj.assign(uo, uo + row);
				}
				r->t_done = now;
			}
		}
		lk.lock();
		if (err)
			fail_batch(b, err, msg);
		else
			for (Request *r : b.reqs)
				r->state = 2;
		b.reqs.clear();
		sl.busy = false;
		s->inflight.pop_front();
		s->cv_done.notify_all();
		s->cv_disp.notify_one();
	}
}

void free_slots(kssdhip_server *s)
{
	for (Slot &sl : s->slots) {
		if (sl.stream)
			(void)hipStreamSynchronize(sl.stream);
		sl.split.release();
		sl.qws.release();
		sl.sdense.release();
		if (sl.d_arena)
			(void)hipFree(sl.d_arena);
		if (sl.h_arena)
			(void)hipHostFree(sl.h_arena);
		if (sl.done)
			(void)hipEventDestroy(sl.done);
		if (sl.stream)
			(void)hipStreamDestroy(sl.stream);
	}
	s->slots.clear();
}

} // namespace

extern "C" int kssdhip_serve_open(kssdhip_index *ix, const kssdhip_serve_opts_t *opts, kssdhip_server **out)
{
	if (!ix || !opts || !out)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: null argument");
	*out = nullptr;
	if (opts->slots < 1 || opts->slots > 16 || opts->max_batch < 1 || opts->max_queued < 1)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: slots 1..16, max_batch >= 1, max_queued >= 1");
	if (opts->flags & ~(KSSDHIP_UNIQUE | KSSDHIP_IGNORECONFLICT))
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: flags may hold UNIQUE and IGNORECONFLICT only");
	IndexView v;
	int rc = kssdhip_index_info(ix, &v);
	if (rc)
		return rc;
	if (opts->select) {
		const kssdhip_select_opts_t &so = opts->select_opts;
		if (so.ctxcut < 1 || !opts->ref_ctx || opts->ntop < 0)
			return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: select needs ctxcut >= 1, ref_ctx and ntop >= 0");
		if (so.unique_rank && !(opts->flags & KSSDHIP_UNIQUE))
			return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: unique_rank needs KSSDHIP_UNIQUE");
		if (!kssd::epi_metric_supported(so.value))
			return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: value metric %d (use 3..9)", so.value.metric);
	}
	kssdhip_server *s = new kssdhip_server();
	s->ix = ix;
	s->opts = *opts;
	s->opts.ref_ctx = nullptr; // copied below
	s->device = v.device;
	s->R = v.ref_gnum;
	s->obj_bits = v.obj_bits;
	s->unique = (opts->flags & KSSDHIP_UNIQUE) != 0;
	s->results = opts->select != 0;
	s->sel = opts->select_opts;
	s->ntop = opts->ntop;
	s->slot_cap = (opts->budget_bytes / (size_t)opts->slots) & ~(size_t)255;
	if (s->slot_cap < request_bytes(1, s->R, s->unique, s->results)) {
		delete s;
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: budget %zu bytes cannot hold one query per slot",
				    opts->budget_bytes);
	}
	if (hipSetDevice(s->device) != hipSuccess) {
		delete s;
		return kssdhip_fail(KSSDHIP_EHIP, "serve_open: device %d", s->device);
	}
	if (s->results) {
		const size_t rb = (size_t)(s->R ? s->R : 1) * 4;
		if (hipMalloc(&s->d_ref_ctx, rb) != hipSuccess ||
		    (s->R && hipMemcpy(s->d_ref_ctx, opts->ref_ctx, (size_t)s->R * 4, hipMemcpyHostToDevice) != hipSuccess)) {
			if (s->d_ref_ctx)
				(void)hipFree(s->d_ref_ctx);
			delete s;
			return kssdhip_fail(KSSDHIP_ENOMEM, "serve_open: reference context totals");
		}
	}
	s->slots.resize((size_t)opts->slots);
	for (Slot &sl : s->slots) {
		if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
		    hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
		    hipMalloc(&sl.d_arena, s->slot_cap) != hipSuccess ||
		    hipHostMalloc(&sl.h_arena, s->slot_cap, hipHostMallocDefault) != hipSuccess) {
			free_slots(s);
			if (s->d_ref_ctx)
				(void)hipFree(s->d_ref_ctx);
			const size_t cap = s->slot_cap;
			delete s;
			return kssdhip_fail(KSSDHIP_ENOMEM, "serve_open: slot resources (%zu bytes each)", cap);
		}
	}
	s->dispatcher = std::thread(dispatch_loop, s);
	s->completer = std::thread(complete_loop, s);
	*out = s;
	return 0;
}

extern "C" int64_t kssdhip_serve_submit(kssdhip_server *s, const uint64_t *keys, size_t n)
{
	if (!s || (n && !keys))
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_submit: null argument");
	if (request_bytes(n, s->R, s->unique, s->results) > s->slot_cap)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_submit: %zu keys exceed the per-slot budget of %zu bytes", n,
				    s->slot_cap);
	Request *r = new Request();
	r->keys.assign(keys, keys + n);
	if (s->results)
		r->qctx = count_contexts(r->keys, s->obj_bits);
	std::lock_guard<std::mutex> lk(s->mu);
	if (s->stop) {
		delete r;
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_submit: server is closing");
	}
	if ((int)s->queue.size() >= s->opts.max_queued) {
		delete r;
		s->info.rejected += 1;
		return kssdhip_fail(KSSDHIP_EAGAIN, "serve_submit: %d requests already queued", s->opts.max_queued);
	}
	r->ticket = s->next_ticket++;
	r->t_submit = Clock::now();
	s->reqs[r->ticket] = r;
	s->queue.push_back(r);
	s->cv_disp.notify_one();
	return r->ticket;
}

extern "C" int kssdhip_serve_wait(kssdhip_server *s, int64_t ticket, kssd_ctx_mut2_t *ctx, kssd_obj_section_t *obj,
				  kssd_ctx_mut2_t *uctx, kssd_obj_section_t *uobj, int timeout_ms,
				  kssdhip_serve_stats_t *st)
{
	if (!s)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_wait: null server");
	if (s->results)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_wait: the server returns rows (kssdhip_serve_wait_rows)");
	std::unique_lock<std::mutex> lk(s->mu);
	auto it = s->reqs.find(ticket);
	if (it == s->reqs.end())
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_wait: unknown ticket %lld", (long long)ticket);
	Request *r = it->second;
	auto ready = [&] { return r->state >= 2; };
	if (timeout_ms < 0)
		s->cv_done.wait(lk, ready);
	else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
		return kssdhip_fail(KSSDHIP_EAGAIN, "serve_wait: ticket %lld not done", (long long)ticket);
	s->reqs.erase(it);
	lk.unlock();
	int rc = 0;
	if (r->state == 3) {
		rc = kssdhip_fail(r->err ? r->err : KSSDHIP_EHIP, "%s", r->msg.c_str());
	} else {
		const size_t bytes = (size_t)s->R * 8;
		if (ctx && bytes)
			std::memcpy(ctx, r->ctx.data(), bytes);
		if (obj && bytes)
			std::memcpy(obj, r->obj.data(), bytes);
		if (s->unique && uctx && bytes)
			std::memcpy(uctx, r->uctx.data(), bytes);
		if (s->unique && uobj && bytes)
			std::memcpy(uobj, r->uobj.data(), bytes);
	}
	if (st) {
		using ms = std::chrono::duration<double, std::milli>;
		st->queue_ms = r->state >= 1 ? ms(r->t_launch - r->t_submit).count() : 0.0;
		st->device_ms = ms(r->t_done - r->t_launch).count();
		st->total_ms = ms(r->t_done - r->t_submit).count();
		st->batch_queries = r->batch_queries;
	}
	delete r;
	return rc;
}

static void fill_stats(const Request *r, kssdhip_serve_stats_t *st)
{
	if (!st)
		return;
	using ms = std::chrono::duration<double, std::milli>;
	st->queue_ms = r->state >= 1 ? ms(r->t_launch - r->t_submit).count() : 0.0;
	st->device_ms = ms(r->t_done - r->t_launch).count();
	st->total_ms = ms(r->t_done - r->t_submit).count();
	st->batch_queries = r->batch_queries;
}

extern "C" int kssdhip_serve_wait_rows(kssdhip_server *s, int64_t ticket, kssdhip_ani_row_t *rows, size_t cap,
				       size_t *n_rows, int timeout_ms, kssdhip_serve_stats_t *st)
{
	if (!s || !n_rows || (cap && !rows))
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_wait_rows: null argument");
	if (!s->results)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_wait_rows: the server returns counter rows (kssdhip_serve_wait)");
	*n_rows = 0;
	std::unique_lock<std::mutex> lk(s->mu);
	auto it = s->reqs.find(ticket);
	if (it == s->reqs.end())
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_wait_rows: unknown ticket %lld", (long long)ticket);
	Request *r = it->second;
	auto ready = [&] { return r->state >= 2; };
	if (timeout_ms < 0)
		s->cv_done.wait(lk, ready);
	else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
		return kssdhip_fail(KSSDHIP_EAGAIN, "serve_wait_rows: ticket %lld not done", (long long)ticket);
	if (r->state == 2 && r->rows.size() > cap) {
		*n_rows = r->rows.size();
		return kssdhip_fail(KSSDHIP_ENOSPC, "serve_wait_rows: %zu rows exceed the %zu-row buffer", r->rows.size(),
				    cap);
	}
	s->reqs.erase(it);
	lk.unlock();
	int rc = 0;
	if (r->state == 3) {
		rc = kssdhip_fail(r->err ? r->err : KSSDHIP_EHIP, "%s", r->msg.c_str());
	} else {
		*n_rows = r->rows.size();
		if (!r->rows.empty())
			std::memcpy(rows, r->rows.data(), r->rows.size() * sizeof(kssdhip_ani_row_t));
	}
	fill_stats(r, st);
	delete r;
	return rc;
}

extern "C" int kssdhip_serve_info(kssdhip_server *s, kssdhip_serve_info_t *info)
{
	if (!s || !info)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_info: null argument");
	std::lock_guard<std::mutex> lk(s->mu);
	*info = s->info;
	info->slot_bytes = s->slot_cap;
	return 0;
}

extern "C" int kssdhip_serve_close(kssdhip_server *s)
{
	if (!s)
		return 0;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		s->stop = true;
		// queued requests will not run
		for (Request *r : s->queue) {
			r->state = 3;
			r->err = KSSDHIP_EINVAL;
			r->msg = "serve: server closed before the request ran";
		}
		s->queue.clear();
	}
	s->cv_disp.notify_all();
	s->cv_complete.notify_all();
	s->dispatcher.join();
	s->completer.join();
	for (auto &kv : s->reqs)
		delete kv.second;
	s->reqs.clear();
	(void)hipSetDevice(s->device);
	free_slots(s);
	if (s->d_ref_ctx)
		(void)hipFree(s->d_ref_ctx);
	delete s;
	return 0;
}
