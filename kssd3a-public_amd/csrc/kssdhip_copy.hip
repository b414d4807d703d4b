// Large copies between pageable host memory and the device.
//
// hipMemcpyAsync from or to pageable memory moves about 12 GB/s here (one
// host thread stages it). These helpers cut the copy into 2 MiB chunks and
// let COPY_THREADS host threads each stage their chunks through two pinned
// buffers of their own, so host-side memcpy of one chunk overlaps the DMA of
// the previous one and the threads together keep the PCIe link busy. The
// pinned buffers (COPY_THREADS * 2 * 2 MiB per device) are allocated on the
// first large copy and kept until the process ends; a mutex serialises
// copies that share them. Copies under COPY_MIN bytes go straight to
// hipMemcpyAsync.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_internal.h"

namespace kssd {
namespace {

constexpr int COPY_THREADS = 4;
constexpr size_t COPY_CHUNK = (size_t)2 << 20;
constexpr size_t COPY_MIN = (size_t)8 << 20;
constexpr int COPY_MAX_DEVICES = 16;

struct StagePool {
	std::mutex mu;
	bool ready = false;
	char *buf[COPY_THREADS * 2] = {};
	hipEvent_t ev[COPY_THREADS * 2] = {};
};
StagePool g_pool[COPY_MAX_DEVICES];

hipError_t pool_for(int dev, StagePool **out)
{
	StagePool &p = g_pool[dev];
	if (!p.ready) {
		for (int b = 0; b < COPY_THREADS * 2; ++b) {
			hipError_t e = hipSuccess;
			if (!p.buf[b])
				e = hipHostMalloc((void **)&p.buf[b], COPY_CHUNK, hipHostMallocDefault);
			if (e == hipSuccess && !p.ev[b])
				e = hipEventCreateWithFlags(&p.ev[b], hipEventDisableTiming);
			if (e != hipSuccess)
				return e; // the buffers made so far stay for the next attempt to reuse
		}
		p.ready = true;
	}
	*out = &p;
	return hipSuccess;
}

// One worker's share: chunks w, w + COPY_THREADS, ... through its buffers
// 2w and 2w + 1.
void h2d_worker(int dev, StagePool *p, int w, char *dst, const char *src, size_t bytes, hipStream_t s,
		std::atomic<int> *err)
{
	if (hipSetDevice(dev) != hipSuccess) {
		err->store((int)hipErrorInvalidDevice);
		return;
	}
	const size_t nch = (bytes + COPY_CHUNK - 1) / COPY_CHUNK;
	int k = 0;
	for (size_t c = (size_t)w; c < nch && !err->load(std::memory_order_relaxed); c += COPY_THREADS, ++k) {
		const int b = 2 * w + (k & 1);
		const size_t off = c * COPY_CHUNK, len = bytes - off < COPY_CHUNK ? bytes - off : COPY_CHUNK;
		hipError_t e = hipEventSynchronize(p->ev[b]); // its previous DMA has finished reading it
		if (e == hipSuccess) {
			memcpy(p->buf[b], src + off, len);
			e = hipMemcpyAsync(dst + off, p->buf[b], len, hipMemcpyHostToDevice, s);
		}
		if (e == hipSuccess)
			e = hipEventRecord(p->ev[b], s);
		if (e != hipSuccess)
			err->store((int)e);
	}
}

// Chunk j of the worker's share is queued into buffer 2w + (j & 1) before
// chunk j - 1 is copied out of the other one.
void d2h_worker(int dev, StagePool *p, int w, char *dst, const char *src, size_t bytes, hipStream_t s,
		std::atomic<int> *err)
{
	if (hipSetDevice(dev) != hipSuccess) {
		err->store((int)hipErrorInvalidDevice);
		return;
	}
	const size_t nch = (bytes + COPY_CHUNK - 1) / COPY_CHUNK;
	const size_t mine = nch > (size_t)w ? (nch - (size_t)w + COPY_THREADS - 1) / COPY_THREADS : 0;
	for (size_t j = 0; j <= mine && !err->load(std::memory_order_relaxed); ++j) {
		hipError_t e = hipSuccess;
		if (j < mine) {
			const int b = 2 * w + (int)(j & 1);
			const size_t off = ((size_t)w + j * COPY_THREADS) * COPY_CHUNK;
			const size_t len = bytes - off < COPY_CHUNK ? bytes - off : COPY_CHUNK;
			e = hipEventSynchronize(p->ev[b]);
			if (e == hipSuccess)
				e = hipMemcpyAsync(p->buf[b], src + off, len, hipMemcpyDeviceToHost, s);
			if (e == hipSuccess)
				e = hipEventRecord(p->ev[b], s);
		}
		if (e == hipSuccess && j > 0) {
			const int b = 2 * w + (int)((j - 1) & 1);
			const size_t off = ((size_t)w + (j - 1) * COPY_THREADS) * COPY_CHUNK;
			const size_t len = bytes - off < COPY_CHUNK ? bytes - off : COPY_CHUNK;
			e = hipEventSynchronize(p->ev[b]);
			if (e == hipSuccess)
				memcpy(dst + off, p->buf[b], len);
		}
		if (e != hipSuccess)
			err->store((int)e);
	}
}

hipError_t staged_copy(void *dst, const void *src, size_t bytes, hipStream_t s, bool h2d)
{
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess)
		return e;
	if (dev < 0 || dev >= COPY_MAX_DEVICES)
		return hipErrorInvalidDevice;
	StagePool *p = nullptr;
	std::lock_guard<std::mutex> lock(g_pool[dev].mu);
	if ((e = pool_for(dev, &p)) != hipSuccess)
		return e;
	std::atomic<int> err{0};
	std::vector<std::thread> th;
	bool started = true;
	try {
		th.reserve(COPY_THREADS);
		for (int w = 0; w < COPY_THREADS; ++w)
			th.emplace_back(h2d ? h2d_worker : d2h_worker, dev, p, w, (char *)dst, (const char *)src, bytes, s,
					&err);
	} catch (...) { // no thread to be had: the workers started finish, then one plain copy of it all
		started = false;
	}
	for (std::thread &t : th)
		t.join();
	if (!started) {
		e = hipMemcpyAsync(dst, src, bytes, h2d ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, s);
		return e == hipSuccess && !h2d ? hipStreamSynchronize(s) : e;
	}
	return (hipError_t)err.load();
}

} // namespace

hipError_t copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s)
{
	if (bytes < COPY_MIN)
		return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
	return staged_copy(dst, src, bytes, s, true);
}

hipError_t copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s)
{
	if (bytes < COPY_MIN) {
		hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
		return e == hipSuccess ? hipStreamSynchronize(s) : e;
	}
	return staged_copy(dst, src, bytes, s, false);
}

} // namespace kssd
