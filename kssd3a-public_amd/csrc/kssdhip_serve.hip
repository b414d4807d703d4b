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
// order and hands each request its rows.
#include <hip/hip_runtime.h>

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
};

struct Batch {
	std::vector<Request *> reqs;
	size_t off_skidx = 0, off_ctx = 0, off_obj = 0, off_uctx = 0, off_uobj = 0, end = 0;
	uint64_t total = 0;
};

struct Slot {
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	void *d_arena = nullptr;
	void *h_arena = nullptr;
	DevBuf split;
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
	bool unique = false;
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
size_t layout(Batch &b, int R, bool unique)
{
	size_t nkeys = 0;
	for (const Request *r : b.reqs)
		nkeys += r->keys.size();
	const size_t nq = b.reqs.size(), cells = nq * (size_t)R;
	b.total = nkeys;
	b.off_skidx = align_up(nkeys * 8);
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

size_t request_bytes(size_t nkeys, int R, bool unique)
{
	return align_up(nkeys * 8) + align_up(16) + (unique ? 4 : 2) * align_up((size_t)R * 8);
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
			if (layout(b, s->R, s->unique) > s->slot_cap) {
				b.reqs.pop_back();
				break;
			}
			s->queue.pop_front();
		}
		layout(b, s->R, s->unique);
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
		const Clock::time_point now = Clock::now();
		if (!err) {
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
	kssdhip_server *s = new kssdhip_server();
	s->ix = ix;
	s->opts = *opts;
	s->device = v.device;
	s->R = v.ref_gnum;
	s->unique = (opts->flags & KSSDHIP_UNIQUE) != 0;
	s->slot_cap = (opts->budget_bytes / (size_t)opts->slots) & ~(size_t)255;
	if (s->slot_cap < request_bytes(1, s->R, s->unique)) {
		delete s;
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_open: budget %zu bytes cannot hold one query per slot",
				    opts->budget_bytes);
	}
	if (hipSetDevice(s->device) != hipSuccess) {
		delete s;
		return kssdhip_fail(KSSDHIP_EHIP, "serve_open: device %d", s->device);
	}
	s->slots.resize((size_t)opts->slots);
	for (Slot &sl : s->slots) {
		if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
		    hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
		    hipMalloc(&sl.d_arena, s->slot_cap) != hipSuccess ||
		    hipHostMalloc(&sl.h_arena, s->slot_cap, hipHostMallocDefault) != hipSuccess) {
			free_slots(s);
			delete s;
			return kssdhip_fail(KSSDHIP_ENOMEM, "serve_open: slot resources (%zu bytes each)", (size_t)0);
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
	if (request_bytes(n, s->R, s->unique) > s->slot_cap)
		return kssdhip_fail(KSSDHIP_EINVAL, "serve_submit: %zu keys exceed the per-slot budget of %zu bytes", n,
				    s->slot_cap);
	Request *r = new Request();
	r->keys.assign(keys, keys + n);
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
	delete s;
	return 0;
}
