// Autotuner of the counting path's kernel shape and its JSON cache.
//
// For a resident index and a device-resident query block the candidates --
// the lookup kernel, the partitioned merge-join over LDS chunk sizes,
// workgroup sizes and partition groups (including chunk sizes fitted so the
// partition groups fill whole rounds of the compute units), and the
// query-major join over workgroup sizes -- are timed with HIP events on the
// handle's stream, and the fastest is applied.  Results are cached in a
// JSON file keyed by the device and the shape of the problem (index
// records, query keys, queries, references, each rounded down to a power of
// two, and the flags), so later processes pick the shape up without timing.
// The reference's analogue is its fixed block sizes (command_ani.c:403,
// :1517-1529; command_matrix.c:689-699).
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/file.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "kssdhip.h"
#include "kssdhip_internal.h"

namespace {

struct Entry {
	kssdhip_tiling_t t{};
	double ms = 0;
};

uint64_t bucket(uint64_t x)
{
	if (!x)
		return 0;
	uint64_t b = 1;
	while (b <= x / 2)
		b *= 2;
	return b;
}

std::string device_tag(int device)
{
	hipDeviceProp_t p;
	if (hipGetDeviceProperties(&p, device) != hipSuccess)
		return "hip";
	std::string s = std::string(p.gcnArchName) + "/" + p.name;
	for (char &c : s)
		if (c == '"' || c == '\\' || c == '|' || (unsigned char)c < 32)
			c = '_';
	return s;
}

std::string index_prefix(const std::string &dev, uint64_t n, int R)
{
	char b[128];
	snprintf(b, sizeof b, "|N%llu|R%llu|", (unsigned long long)bucket(n), (unsigned long long)bucket((uint64_t)R));
	return dev + b;
}

std::string shape_key(const std::string &dev, uint64_t n, int R, uint64_t total, int nq, uint32_t flags)
{
	char b[128];
	snprintf(b, sizeof b, "K%llu|Q%llu|F%u", (unsigned long long)bucket(total), (unsigned long long)bucket((uint64_t)nq),
		 flags);
	return index_prefix(dev, n, R) + b;
}

std::string cache_path(const char *path)
{
	if (path && path[0])
		return path;
	const char *e = getenv("KSSDHIP_TUNE_CACHE");
	if (e && e[0])
		return e;
	const char *home = getenv("HOME");
	return std::string(home && home[0] ? home : "/tmp") + "/.cache/kssdhip/tuning.json";
}

// The cache file: {"version": 1, "entries": {"<key>": {"mode": .., ...}, ...}}
// as this file writes it; entries it cannot read are skipped.
std::map<std::string, Entry> read_cache(const std::string &path)
{
	std::map<std::string, Entry> out;
	FILE *f = fopen(path.c_str(), "r");
	if (!f)
		return out;
	std::string text;
	char buf[65536];
	size_t k;
	while ((k = fread(buf, 1, sizeof buf, f)) > 0)
		text.append(buf, k);
	fclose(f);
	size_t at = text.find("\"entries\"");
	if (at == std::string::npos)
		return out;
	at = text.find('{', at);
	while (at != std::string::npos) {
		const size_t q0 = text.find('"', at + 1);
		if (q0 == std::string::npos)
			break;
		const size_t q1 = text.find('"', q0 + 1);
		const size_t ob = q1 == std::string::npos ? q1 : text.find('{', q1);
		const size_t cb = ob == std::string::npos ? ob : text.find('}', ob);
		if (cb == std::string::npos)
			break;
		const std::string key = text.substr(q0 + 1, q1 - q0 - 1);
		const std::string body = text.substr(ob + 1, cb - ob - 1);
		Entry e;
		auto num = [&](const char *name, double *v) {
			const size_t p = body.find(std::string("\"") + name + "\"");
			if (p == std::string::npos)
				return false;
			const size_t c = body.find(':', p);
			if (c == std::string::npos)
				return false;
			char *end = nullptr;
			*v = strtod(body.c_str() + c + 1, &end);
			return end != body.c_str() + c + 1;
		};
		double v[7];
		const char *names[7] = {"mode", "chunk_records", "block_threads", "group_partitions", "query_threads",
					"query_slots", "ms"};
		bool ok = true;
		for (int i = 0; i < 7; ++i)
			ok = ok && num(names[i], &v[i]);
		if (ok) {
			e.t.mode = (int)v[0];
			e.t.chunk_records = (int)v[1];
			e.t.block_threads = (int)v[2];
			e.t.group_partitions = (int)v[3];
			e.t.query_threads = (int)v[4];
			e.t.query_slots = (int)v[5];
			e.t.auto_min_keys = 512.0;
			e.ms = v[6];
			out[key] = e;
		}
		at = cb;
	}
	return out;
}

int write_cache(const std::string &path, const std::map<std::string, Entry> &m)
{
	const size_t slash = path.rfind('/');
	if (slash != std::string::npos && slash > 0) { // make the directory chain
		std::string d = path.substr(0, slash);
		for (size_t p = 1; p <= d.size(); ++p)
			if (p == d.size() || d[p] == '/')
				(void)mkdir(d.substr(0, p).c_str(), 0755);
	}
	const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
	FILE *f = fopen(tmp.c_str(), "w");
	if (!f)
		return -1;
	fprintf(f, "{\n \"version\": 1,\n \"entries\": {");
	bool first = true;
	for (const auto &kv : m) {
		const kssdhip_tiling_t &t = kv.second.t;
		fprintf(f,
			"%s\n  \"%s\": {\"mode\": %d, \"chunk_records\": %d, \"block_threads\": %d, \"group_partitions\": %d, "
			"\"query_threads\": %d, \"query_slots\": %d, \"ms\": %.6f}",
			first ? "" : ",", kv.first.c_str(), t.mode, t.chunk_records, t.block_threads, t.group_partitions,
			t.query_threads, t.query_slots, kv.second.ms);
		first = false;
	}
	fprintf(f, "\n }\n}\n");
	const bool ok = fclose(f) == 0;
	if (!ok || rename(tmp.c_str(), path.c_str()) != 0) {
		(void)unlink(tmp.c_str());
		return -1;
	}
	return 0;
}

// an advisory lock beside the cache, held while it is read and rewritten
struct CacheLock {
	int fd = -1;
	explicit CacheLock(const std::string &path)
	{
		fd = open((path + ".lock").c_str(), O_RDWR | O_CREAT, 0644);
		if (fd >= 0)
			(void)flock(fd, LOCK_EX);
	}
	~CacheLock()
	{
		if (fd >= 0) {
			(void)flock(fd, LOCK_UN);
			close(fd);
		}
	}
};

kssdhip_tiling_t tiling(int mode, int chunk, int block, int group, int qthreads = 0, int qslots = 0)
{
	kssdhip_tiling_t t{};
	t.mode = mode;
	t.chunk_records = chunk;
	t.block_threads = block;
	t.auto_min_keys = 512.0;
	t.group_partitions = group;
	t.query_threads = qthreads;
	t.query_slots = qslots;
	return t;
}

std::vector<kssdhip_tiling_t> candidates(uint64_t n, int cus, int R)
{
	std::vector<kssdhip_tiling_t> v;
	v.push_back(tiling(1, 4096, 256, 1));
	for (int c : {1024, 2048, 4096, 8192, 12288})
		for (int b : {128, 256, 512})
			v.push_back(tiling(2, c, b, 1));
	std::vector<std::pair<int, int>> cg;
	for (int c : {2048, 3072, 4096, 6144})
		for (int g : {2, 4, 8})
			cg.push_back({c, g});
	// chunks whose partition groups come to whole rounds of one workgroup per CU
	if (n && cus > 0)
		for (int g : {2, 4, 8})
			for (int rounds = 1; rounds <= 4; ++rounds) {
				uint64_t c = (n + (uint64_t)cus * g * rounds - 1) / ((uint64_t)cus * g * rounds);
				c = (c + 63) / 64 * 64;
				if (c >= 2048 && c <= 6144 && std::find(cg.begin(), cg.end(), std::make_pair((int)c, g)) == cg.end())
					cg.push_back({(int)c, g});
			}
	for (auto &p : cg)
		for (int b : {512, 1024})
			v.push_back(tiling(2, p.first, b, p.second));
	for (int qt : {128, 256, 512})
		v.push_back(tiling(3, 4096, 256, 1, qt, 0));
	// hashed LDS tables: their size sets the workgroups per CU (where a
	// direct row fits, small hashed tables leave room for more workgroups)
	const bool direct_fits = (uint64_t)R * 16 <= 64 * 1024;
	for (int qt : {256, 512})
		for (int qs : {256, 512, 1024, 2048})
			if (direct_fits ? qs <= 512 : qs != 512)
				v.push_back(tiling(3, 4096, 256, 1, qt, qs));
	return v;
}

} // namespace

// The default shape of a new handle: partition groups of four, fitted to whole
// rounds of the compute units, for indexes large enough to fill them; the
// kernel is chosen per call (mode 0).
void kssd_default_tiling(const kssdhip_index *ix, kssdhip_tiling_t *t)
{
	const int cus = ix->cus > 0 ? ix->cus : 256;
	*t = tiling(0, 4096, 256, 1);
	for (int rounds = 1; rounds <= 8; ++rounds) {
		uint64_t c = (ix->n + (uint64_t)cus * 4 * rounds - 1) / ((uint64_t)cus * 4 * rounds);
		c = (c + 63) / 64 * 64;
		if (c < 2048)
			break;
		if (c <= 6144) {
			*t = tiling(0, (int)c, 1024, 4);
			break;
		}
	}
}

// The cached shape parameters of the largest block tuned for an index of
// this shape on this device (the kernel still chosen per call); 1 when one
// was found.
int kssd_cached_index_tiling(const kssdhip_index *ix, kssdhip_tiling_t *t)
{
	const std::string path = cache_path(nullptr);
	if (getenv("KSSDHIP_NO_TUNE_CACHE"))
		return 0;
	std::map<std::string, Entry> m;
	{
		CacheLock lock(path);
		m = read_cache(path);
	}
	const std::string pre = index_prefix(device_tag(ix->device), ix->n, ix->ref_gnum);
	uint64_t best_k = 0;
	bool found = false;
	for (const auto &kv : m) {
		if (kv.first.compare(0, pre.size(), pre) != 0)
			continue;
		const uint64_t k = strtoull(kv.first.c_str() + pre.size() + 1, nullptr, 10);
		if (!found || k > best_k) {
			*t = kv.second.t;
			best_k = k;
			found = true;
		}
	}
	if (found)
		t->mode = 0;
	return found ? 1 : 0;
}

extern "C" int kssdhip_index_autotune(kssdhip_index *ix, const uint64_t *d_keys, const uint64_t *d_skidx, int nq,
				      uint64_t total, uint32_t flags, const char *path, int force,
				      kssdhip_tune_result_t *res)
{
	if (!ix || !res || nq < 1 || !d_skidx || (total && !d_keys))
		return kssdhip_fail(KSSDHIP_EINVAL, "autotune: bad arguments");
	if (flags & ~(KSSDHIP_LOWER | KSSDHIP_UNIQUE | KSSDHIP_IGNORECONFLICT))
		return kssdhip_fail(KSSDHIP_EINVAL, "autotune: flags 0x%x", flags);
	memset(res, 0, sizeof *res);
	const std::string cpath = cache_path(path);
	const std::string key = shape_key(device_tag(ix->device), ix->n, ix->ref_gnum, total, nq, flags);
	if (!force) {
		std::map<std::string, Entry> m;
		{
			CacheLock lock(cpath);
			m = read_cache(cpath);
		}
		auto it = m.find(key);
		if (it != m.end()) {
			int rc = kssdhip_index_set_tiling(ix, &it->second.t);
			if (rc)
				return rc;
			res->tiling = it->second.t;
			res->ms = it->second.ms;
			res->from_cache = 1;
			return 0;
		}
	}
	if (hipSetDevice(ix->device) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "autotune: device not available");
	const bool uniq = (flags & KSSDHIP_UNIQUE) && !(flags & KSSDHIP_LOWER);
	const size_t plane = (size_t)nq * (size_t)ix->ref_gnum * 8;
	DevBuf out;
	int rc = out.reserve(plane * (uniq ? 4 : 2));
	if (rc)
		return rc;
	char *o = (char *)out.ptr;
	kssd_ctx_mut2_t *c0 = (kssd_ctx_mut2_t *)o;
	kssd_obj_section_t *o0 = (kssd_obj_section_t *)(o + plane);
	kssd_ctx_mut2_t *uc = uniq ? (kssd_ctx_mut2_t *)(o + 2 * plane) : nullptr;
	kssd_obj_section_t *uo = uniq ? (kssd_obj_section_t *)(o + 3 * plane) : nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		out.release();
		return kssdhip_fail(KSSDHIP_EHIP, "autotune: events");
	}
	kssdhip_tiling_t before{};
	kssdhip_index_get_tiling(ix, &before);
	const hipStream_t s = ix->stream;
	double best_ms = 0;
	kssdhip_tiling_t best{};
	int timed = 0;
	for (const kssdhip_tiling_t &t : candidates(ix->n, ix->cus, ix->ref_gnum)) {
		if ((rc = kssdhip_index_set_tiling(ix, &t)))
			break;
		// a candidate whose launch fails (e.g. no LDS fit) is skipped
		if (kssd_count_dev(ix, d_keys, d_skidx, nq, total, 0, flags, c0, o0, uc, uo, s, &ix->split, &ix->qws) ||
		    hipStreamSynchronize(s) != hipSuccess) {
			(void)hipGetLastError();
			continue;
		}
		const int reps = 3;
		(void)hipEventRecord(e0, s);
		for (int r = 0; r < reps; ++r)
			(void)kssd_count_dev(ix, d_keys, d_skidx, nq, total, 0, flags, c0, o0, uc, uo, s, &ix->split,
					     &ix->qws);
		(void)hipEventRecord(e1, s);
		float ms = 0;
		if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
			rc = kssdhip_fail(KSSDHIP_EHIP, "autotune: %s", hipGetErrorString(hipGetLastError()));
			break;
		}
		++timed;
		if (timed == 1 || ms / reps < best_ms) {
			best_ms = ms / reps;
			best = t;
		}
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	out.release();
	if (rc || !timed) {
		(void)kssdhip_index_set_tiling(ix, &before);
		return rc ? rc : kssdhip_fail(KSSDHIP_EHIP, "autotune: no candidate ran");
	}
	if ((rc = kssdhip_index_set_tiling(ix, &best)))
		return rc;
	res->tiling = best;
	res->ms = best_ms;
	res->candidates = timed;
	{
		CacheLock lock(cpath);
		std::map<std::string, Entry> m = read_cache(cpath);
		Entry e;
		e.t = best;
		e.ms = best_ms;
		m[key] = e;
		res->cached = write_cache(cpath, m) == 0;
	}
	return 0;
}

extern "C" int kssdhip_index_autotune_host(kssdhip_index *ix, const uint64_t *section_sk,
					   const uint64_t *section_skidx, int nq, uint32_t flags, const char *path,
					   int force, kssdhip_tune_result_t *res)
{
	if (!ix || !res || nq < 1 || !section_skidx)
		return kssdhip_fail(KSSDHIP_EINVAL, "autotune: bad arguments");
	const uint64_t base = section_skidx[0];
	std::vector<uint64_t> rel((size_t)nq + 1);
	for (int i = 0; i <= nq; ++i) {
		if (section_skidx[i] < base || (i && section_skidx[i] < section_skidx[i - 1]))
			return kssdhip_fail(KSSDHIP_EINVAL, "autotune: section_skidx not ascending");
		rel[i] = section_skidx[i] - base;
	}
	const uint64_t total = rel[nq];
	if (total && !section_sk)
		return kssdhip_fail(KSSDHIP_EINVAL, "autotune: null keys");
	if (hipSetDevice(ix->device) != hipSuccess)
		return kssdhip_fail(KSSDHIP_EHIP, "autotune: device not available");
	DevBuf keys, skidx;
	int rc = keys.reserve((total ? total : 1) * 8);
	if (!rc)
		rc = skidx.reserve(rel.size() * 8);
	if (!rc && ((total && hipMemcpy(keys.ptr, section_sk, total * 8, hipMemcpyHostToDevice) != hipSuccess) ||
		    hipMemcpy(skidx.ptr, rel.data(), rel.size() * 8, hipMemcpyHostToDevice) != hipSuccess))
		rc = kssdhip_fail(KSSDHIP_EHIP, "autotune: upload failed");
	if (!rc)
		rc = kssdhip_index_autotune(ix, (const uint64_t *)keys.ptr, (const uint64_t *)skidx.ptr, nq, total, flags,
					    path, force, res);
	keys.release();
	skidx.release();
	return rc;
}

extern "C" int kssdhip_index_tiling_from_cache(kssdhip_index *ix, int nq, uint64_t total, uint32_t flags,
					       const char *path, kssdhip_tune_result_t *res)
{
	if (!ix || !res || nq < 0)
		return kssdhip_fail(KSSDHIP_EINVAL, "tiling_from_cache: bad arguments");
	memset(res, 0, sizeof *res);
	const std::string cpath = cache_path(path);
	std::map<std::string, Entry> m;
	{
		CacheLock lock(cpath);
		m = read_cache(cpath);
	}
	auto it = m.find(shape_key(device_tag(ix->device), ix->n, ix->ref_gnum, total, nq, flags));
	if (it == m.end())
		return 0;
	int rc = kssdhip_index_set_tiling(ix, &it->second.t);
	if (rc)
		return rc;
	res->tiling = it->second.t;
	res->ms = it->second.ms;
	res->from_cache = 1;
	return 1;
}
