"""Python bindings (ctypes) of the package's native libraries.

  libkssdhip.so   HIP kernels for gfx950 behind the C ABI of include/kssdhip.h
  libkssdhost.so  host components: sketcher, genome simulator, index builder
  libamdhip64.so  the HIP runtime, for device buffers, streams and events

The bindings fail loudly when a library is missing; there is no fallback.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
REC_DTYPE = np.dtype([("ctxgid", "<u8"), ("obj", "<u4")], align=False)
V, S, U, I, D = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_int, ctypes.c_double
U32, U64 = ctypes.c_uint32, ctypes.c_uint64

LOWER, UNIQUE, IGNORECONFLICT, ACCUMULATE = 1, 2, 4, 8
PAIR_WORKSPACE_BYTES = 16400


def _p(a):
    return a.ctypes.data_as(V) if a is not None else None


class Features(ctypes.Structure):
    """kssd_ani_features_t == ani_features_t (model_ani.h:215-222)."""
    _fields_ = [("XnY_ctx", U32), ("X_ctx", U32), ("N_diff_obj_section", U32),
                ("N_mut2_ctx", U32), ("N_diff_obj", U32)]

    def tup(self):
        return (self.XnY_ctx, self.N_diff_obj, self.N_diff_obj_section, self.N_mut2_ctx)


class EpilogueOpts(ctypes.Structure):
    _fields_ = [("metric", I), ("values_ani", I), ("unassembled", I), ("float_af", I),
                ("lower", I), ("qry_gid_offset", I), ("drfold", I), ("obj_bits", U),
                ("ctx_bits", U), ("afcut", D), ("exception", D)]


class Tiling(ctypes.Structure):
    """kssdhip_tiling_t: mode 0 auto, 1 lookup, 2 partitioned merge-join,
    3 query-major join; group_partitions > 1 lets one workgroup of the
    partitioned join walk that many consecutive partitions, summing hits per
    cell in LDS first; query_threads / query_slots shape the query-major
    kernel (0 = defaults)."""
    _fields_ = [("mode", I), ("chunk_records", I), ("block_threads", I), ("auto_min_keys", D),
                ("group_partitions", I), ("query_threads", I), ("query_slots", I)]


class TuneResult(ctypes.Structure):
    """kssdhip_tune_result_t."""
    _fields_ = [("tiling", Tiling), ("ms", D), ("from_cache", I), ("candidates", I), ("cached", I)]


class SelectOpts(ctypes.Structure):
    """kssdhip_select_opts_t: the index path's pass gate and the FP64 value
    computed for each passing pair."""
    _fields_ = [("ctxcut", U32), ("afcut", ctypes.c_float), ("anicut", ctypes.c_float), ("unique_rank", I),
                ("value", EpilogueOpts)]


ROW_DTYPE = np.dtype([("query", "<u4"), ("ref", "<u4"), ("num_ctx", "<u4"), ("num_mut2_ctx", "<u4"),
                      ("diff_obj", "<u4"), ("diff_obj_section", "<u4"), ("u_num_ctx", "<u4"),
                      ("u_num_mut2_ctx", "<u4"), ("u_diff_obj", "<u4"), ("u_diff_obj_section", "<u4"),
                      ("rank_ani", "<f8"), ("value", "<f8")])
assert ROW_DTYPE.itemsize == 56
ENOSPC = -6
NOALLOC_SCRATCH_BYTES = 64 << 20
NOALLOC_MAX_QUERIES = 1 << 20


class KssdStat(ctypes.Structure):
    """kssd_stat_t (lcofiles.stat header, 32 bytes)."""
    _fields_ = [("hash_id", U32), ("koc", ctypes.c_bool), ("conflict", ctypes.c_bool), ("coden_len", I),
                ("klen", I), ("hclen", I), ("holen", I), ("drfold", I), ("infile_num", I)]


class U64Vec(ctypes.Structure):
    _fields_ = [("a", ctypes.POINTER(U64)), ("n", S), ("cap", S)]


class SimSpec(ctypes.Structure):
    _fields_ = [("anc_seed", U64), ("anc_len", U64), ("mut_seed", U64), ("sub_rate", D),
                ("indel_frac", D), ("anc2_seed", U64), ("anc2_rate", D)]


class ReportOpts(ctypes.Structure):
    """kssd_report_opts_t (the output-shaping fields of the reference's ani_opt_t)."""
    _fields_ = [("fmt", I), ("s", I), ("v", ctypes.c_bool), ("unassembled", ctypes.c_bool),
                ("unified_metric", ctypes.c_bool), ("ignoreconflict", ctypes.c_bool),
                ("raw_output", ctypes.c_bool), ("max_pdist_set", ctypes.c_bool), ("max_pdist", D),
                ("max_diff_obj_section", I), ("e", I), ("ntop", I), ("ctxcut", I),
                ("afcut", ctypes.c_float), ("anicut", ctypes.c_float), ("unique_ctx", ctypes.c_bool),
                ("d", ctypes.c_bool), ("drfold", I), ("obj_bits", U), ("ctx_bits", U),
                ("estimate_coverage", ctypes.c_bool)]

    @classmethod
    def defaults(cls, **kw):
        """The reference's ani defaults (command_ani_wrapper.c:174-221)."""
        o = cls(fmt=0, s=1, v=False, unassembled=False, unified_metric=False, ignoreconflict=False,
                raw_output=False, max_pdist_set=False, max_pdist=0.0, max_diff_obj_section=-1, e=1,
                ntop=0, ctxcut=3, afcut=0.5, anicut=0.95, unique_ctx=False, d=False, drfold=8,
                obj_bits=20, ctx_bits=44, estimate_coverage=False)
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class QueryCounts(ctypes.Structure):
    _fields_ = [("ctx", V), ("obj", V), ("uctx", V), ("uobj", V), ("qry_ctx", U32), ("ref_ctx", V),
                ("nref", I), ("qry_meta", V), ("ref_meta", V), ("cov", V), ("ids", V), ("ncand", I)]


class Buf(ctypes.Structure):
    _fields_ = [("s", ctypes.c_char_p), ("l", S), ("m", S)]


# ------------------------------------------------------------------ HIP runtime
class HipError(RuntimeError):
    pass


class Hip:
    """Minimal HIP runtime access: device memory, streams and events."""
    H2D, D2H, D2D = 1, 2, 3

    def __init__(self, device=0):
        self.lib = ctypes.CDLL("libamdhip64.so")
        L = self.lib
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipMalloc.argtypes = [ctypes.POINTER(V), S]
        L.hipFree.argtypes = [V]
        L.hipMemcpy.argtypes = [V, V, S, I]
        L.hipMemcpyAsync.argtypes = [V, V, S, I, V]
        L.hipStreamCreate.argtypes = [ctypes.POINTER(V)]
        L.hipStreamSynchronize.argtypes = [V]
        L.hipStreamDestroy.argtypes = [V]
        L.hipEventCreate.argtypes = [ctypes.POINTER(V)]
        L.hipEventRecord.argtypes = [V, V]
        L.hipEventSynchronize.argtypes = [V]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), V, V]
        L.hipEventDestroy.argtypes = [V]
        L.hipSetDevice.argtypes = [I]
        L.hipMemsetAsync.argtypes = [V, I, S, V]
        L.hipStreamBeginCapture.argtypes = [V, I]
        L.hipStreamEndCapture.argtypes = [V, ctypes.POINTER(V)]
        L.hipGraphInstantiate.argtypes = [ctypes.POINTER(V), V, V, V, S]
        L.hipGraphLaunch.argtypes = [V, V]
        L.hipGraphExecDestroy.argtypes = [V]
        L.hipGraphDestroy.argtypes = [V]
        L.hipHostMalloc.argtypes = [ctypes.POINTER(V), S, U32]
        L.hipHostFree.argtypes = [V]
        L.hipStreamWaitEvent.argtypes = [V, V, U32]
        self.check(L.hipSetDevice(device))
        self.device = device

    # ---- graphs: capture the launches a callable makes on `stream`
    def capture(self, stream, fn):
        """Capture fn()'s work on stream into an instantiated graph; returns
        (graph, exec).  fn must only enqueue device work (no allocation)."""
        self.check(self.lib.hipStreamBeginCapture(stream, 0))  # hipStreamCaptureModeGlobal
        try:
            fn()
        except BaseException:
            g = V()
            self.lib.hipStreamEndCapture(stream, ctypes.byref(g))
            if g:
                self.lib.hipGraphDestroy(g)
            raise
        g = V()
        self.check(self.lib.hipStreamEndCapture(stream, ctypes.byref(g)))
        ex = V()
        self.check(self.lib.hipGraphInstantiate(ctypes.byref(ex), g, None, None, 0))
        return g, ex

    def graph_launch(self, ex, stream):
        self.check(self.lib.hipGraphLaunch(ex, stream))

    def graph_destroy(self, g, ex):
        self.check(self.lib.hipGraphExecDestroy(ex))
        self.check(self.lib.hipGraphDestroy(g))

    def check(self, rc):
        if rc != 0:
            raise HipError(f"HIP error {rc}: {self.lib.hipGetErrorString(rc).decode()}")

    def malloc(self, nbytes):
        p = V()
        self.check(self.lib.hipMalloc(ctypes.byref(p), max(int(nbytes), 1)))
        return p

    def free(self, p):
        if p:
            self.check(self.lib.hipFree(p))

    def host_malloc(self, nbytes):
        """Pinned host memory (hipHostMalloc), for asynchronous copies."""
        p = V()
        self.check(self.lib.hipHostMalloc(ctypes.byref(p), max(int(nbytes), 1), 0))
        return p

    def host_free(self, p):
        if p:
            self.check(self.lib.hipHostFree(p))

    def copy_async(self, dst, src, nbytes, kind, stream):
        if nbytes:
            self.check(self.lib.hipMemcpyAsync(dst, src, int(nbytes), kind, stream))

    def wait_event(self, stream, ev):
        self.check(self.lib.hipStreamWaitEvent(stream, ev, 0))

    def upload(self, arr, dptr=None):
        arr = np.ascontiguousarray(arr)
        if dptr is None:
            dptr = self.malloc(arr.nbytes)
        if arr.nbytes:
            self.check(self.lib.hipMemcpy(dptr, _p(arr), arr.nbytes, self.H2D))
        return dptr

    def download(self, dptr, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            self.check(self.lib.hipMemcpy(_p(out), dptr, out.nbytes, self.D2H))
        return out

    def stream(self):
        s = V()
        self.check(self.lib.hipStreamCreate(ctypes.byref(s)))
        return s

    def sync(self, stream):
        self.check(self.lib.hipStreamSynchronize(stream))

    def event(self):
        e = V()
        self.check(self.lib.hipEventCreate(ctypes.byref(e)))
        return e

    def record(self, ev, stream):
        self.check(self.lib.hipEventRecord(ev, stream))

    def elapsed_ms(self, a, b):
        self.check(self.lib.hipEventSynchronize(b))
        ms = ctypes.c_float()
        self.check(self.lib.hipEventElapsedTime(ctypes.byref(ms), a, b))
        return ms.value


# ------------------------------------------------------------------ serving
HIT_DTYPE = np.dtype([("query", "<u4"), ("ref", "<u4"), ("value", "<f8")])  # kssdhip_hit_t


class ServeOpts(ctypes.Structure):
    _fields_ = [("budget_bytes", S), ("slots", I), ("max_batch", I), ("max_queued", I), ("flags", U32),
                ("select", I), ("select_opts", SelectOpts), ("ref_ctx", V), ("ntop", I)]


class ServeStats(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("batch_queries", I)]


class ServeInfo(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_int64), ("queries", ctypes.c_int64), ("rejected", ctypes.c_int64),
                ("max_batch_seen", I), ("slot_bytes", S)]


EAGAIN = -5


# ------------------------------------------------------------------ libkssdhip
def load_kssdhip():
    so = PKG / "libkssdhip.so"
    if not so.exists():
        raise RuntimeError(f"{so} is not built: run `make -C kssd3a-public_amd`")
    lib = ctypes.CDLL(str(so))
    lib.kssdhip_last_error.restype = ctypes.c_char_p
    lib.kssdhip_init.argtypes = [I]
    lib.kssdhip_index_open.argtypes = [V, S, I, U, ctypes.POINTER(V)]
    lib.kssdhip_index_close.argtypes = [V]
    lib.kssdhip_build_index.argtypes = [V, V, I, U, V, V]
    lib.kssdhip_index_device_bytes.argtypes = [V]
    lib.kssdhip_index_device_bytes.restype = S
    lib.kssdhip_count_block.argtypes = [V, V, V, I, I, U32, V, V, V, V]
    lib.kssdhip_count_block_dev.argtypes = [V, V, V, I, U64, I, U32, V, V, V, V, V]
    lib.kssdhip_find_first.argtypes = [V, V, S, V]
    lib.kssdhip_count_votes.argtypes = [V, V, V, I, I, U32, U32, V]
    lib.kssdhip_pair_features.argtypes = [V, S, V, S, U, I, ctypes.POINTER(Features)]
    lib.kssdhip_pair_workspace_init_dev.argtypes = [V, V]
    lib.kssdhip_pair_features_dev.argtypes = [V, S, V, S, U, I, V, V, V]
    lib.kssdhip_index_set_tiling.argtypes = [V, ctypes.POINTER(Tiling)]
    lib.kssdhip_index_get_tiling.argtypes = [V, ctypes.POINTER(Tiling)]
    lib.kssdhip_epilogue_dev.argtypes = [V, V, I, I, V, V, ctypes.POINTER(EpilogueOpts), V, V]
    lib.kssdhip_select_hits_dev.argtypes = [V, I, I, ctypes.c_double, U32, V, V, V]
    lib.kssdhip_index_matched_records.argtypes = [V, V, V, I, U64, ctypes.POINTER(U64)]
    lib.kssdhip_probe_random_reads.argtypes = [S, U64, ctypes.POINTER(D)]
    lib.kssdhip_index_table_info.argtypes = [V, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(I),
                                             ctypes.POINTER(I)]
    lib.kssdhip_index_autotune.argtypes = [V, V, V, I, U64, U32, ctypes.c_char_p, I, ctypes.POINTER(TuneResult)]
    lib.kssdhip_index_autotune_host.argtypes = [V, V, V, I, U32, ctypes.c_char_p, I, ctypes.POINTER(TuneResult)]
    lib.kssdhip_index_tiling_from_cache.argtypes = [V, I, U64, U32, ctypes.c_char_p, ctypes.POINTER(TuneResult)]
    lib.kssdhip_select_block.argtypes = [V, V, V, I, U32, V, V, ctypes.POINTER(SelectOpts), V, S,
                                         ctypes.POINTER(S)]
    lib.kssdhip_select_block_dev.argtypes = [V, V, V, I, U64, U32, V, V, ctypes.POINTER(SelectOpts), V, U32,
                                             V, V, V, V]
    lib.kssdhip_select_block_dev_noalloc.argtypes = lib.kssdhip_select_block_dev.argtypes
    lib.kssdhip_serve_open.argtypes = [V, ctypes.POINTER(ServeOpts), ctypes.POINTER(V)]
    lib.kssdhip_serve_submit.argtypes = [V, V, S]
    lib.kssdhip_serve_submit.restype = ctypes.c_int64
    lib.kssdhip_serve_wait.argtypes = [V, ctypes.c_int64, V, V, V, V, I, ctypes.POINTER(ServeStats)]
    lib.kssdhip_serve_wait_rows.argtypes = [V, ctypes.c_int64, V, S, ctypes.POINTER(S), I, ctypes.POINTER(ServeStats)]
    lib.kssdhip_serve_info.argtypes = [V, ctypes.POINTER(ServeInfo)]
    lib.kssdhip_serve_close.argtypes = [V]
    return lib


class KssdHip:
    """numpy-facing wrapper of the C ABI (include/kssdhip.h)."""

    def __init__(self, device=0):
        self.lib = load_kssdhip()
        self.check(self.lib.kssdhip_init(device))

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(f"kssdhip error {rc}: {self.lib.kssdhip_last_error().decode()}")

    def open_index(self, records, ref_gnum, obj_bits):
        h = V()
        records = np.ascontiguousarray(records)
        self.check(self.lib.kssdhip_index_open(_p(records), len(records), ref_gnum, obj_bits,
                                               ctypes.byref(h)))
        return h

    def count_votes(self, h, ref_gnum, keys, skidx_abs, qry_gid_offset=0, max_ctx_freq=0, sample_step=1):
        nq = len(skidx_abs) - 1
        votes = np.zeros((nq, ref_gnum), dtype=np.uint32)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        skidx_abs = np.ascontiguousarray(skidx_abs, dtype=np.uint64)
        self.check(self.lib.kssdhip_count_votes(h, _p(keys), _p(skidx_abs), nq, qry_gid_offset, max_ctx_freq,
                                                sample_step, _p(votes)))
        return votes

    def build_index(self, keys, skidx, obj_bits=20, unique=True):
        """Sorted index records (structured array ctxgid/obj) and, with
        `unique`, the unique-marker bitset bytes, built on the device."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        skidx = np.ascontiguousarray(skidx, dtype=np.uint64)
        n = int(skidx[-1]) if len(skidx) else 0
        rec = np.empty(n, dtype=REC_DTYPE)
        bits = np.zeros(max((n + 7) // 8, 1), dtype=np.uint8) if unique else None
        self.check(self.lib.kssdhip_build_index(_p(keys) if len(keys) else None, _p(skidx), len(skidx) - 1,
                                                obj_bits, _p(rec) if n else _p(np.zeros(1, np.uint8)),
                                                _p(bits) if unique else None))
        return rec, (bits[:(n + 7) // 8].tobytes() if unique else None)

    def set_tiling(self, h, mode=None, chunk_records=None, block_threads=None, auto_min_keys=None,
                   group_partitions=None, query_threads=None, query_slots=None):
        t = self.get_tiling(h)
        if query_threads is not None:
            t.query_threads = query_threads
        if query_slots is not None:
            t.query_slots = query_slots
        if group_partitions is not None:
            t.group_partitions = group_partitions
        if mode is not None:
            t.mode = mode
        if chunk_records is not None:
            t.chunk_records = chunk_records
            if chunk_records > 6144 and group_partitions is None:
                t.group_partitions = 1  # partition groups need chunk_records <= 6144
        if block_threads is not None:
            t.block_threads = block_threads
        if auto_min_keys is not None:
            t.auto_min_keys = auto_min_keys
        self.check(self.lib.kssdhip_index_set_tiling(h, ctypes.byref(t)))

    def autotune(self, h, d_keys, d_skidx, nq, total, flags=0, cache_path=None, force=False):
        """Time the kernel shapes on a device-resident block (or take the
        cached one), apply the fastest; returns a TuneResult."""
        r = TuneResult()
        self.check(self.lib.kssdhip_index_autotune(h, d_keys, d_skidx, nq, total, flags,
                                                   str(cache_path).encode() if cache_path else None, int(force),
                                                   ctypes.byref(r)))
        return r

    def autotune_host(self, h, keys, skidx_abs, flags=0, cache_path=None, force=False):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        skidx_abs = np.ascontiguousarray(skidx_abs, dtype=np.uint64)
        r = TuneResult()
        self.check(self.lib.kssdhip_index_autotune_host(h, _p(keys), _p(skidx_abs), len(skidx_abs) - 1, flags,
                                                        str(cache_path).encode() if cache_path else None,
                                                        int(force), ctypes.byref(r)))
        return r

    def tiling_from_cache(self, h, nq, total, flags=0, cache_path=None):
        """(applied?, TuneResult)."""
        r = TuneResult()
        rc = self.lib.kssdhip_index_tiling_from_cache(h, nq, total, flags,
                                                      str(cache_path).encode() if cache_path else None,
                                                      ctypes.byref(r))
        if rc < 0:
            self.check(rc)
        return rc == 1, r

    def get_tiling(self, h):
        t = Tiling()
        self.check(self.lib.kssdhip_index_get_tiling(h, ctypes.byref(t)))
        return t

    def close_index(self, h):
        self.check(self.lib.kssdhip_index_close(h))

    def count_block(self, h, ref_gnum, keys, skidx_abs, qry_gid_offset=0, flags=0):
        nq = len(skidx_abs) - 1
        ctx = np.zeros((nq, ref_gnum, 2), dtype=np.uint32)
        obj = np.zeros((nq, ref_gnum, 2), dtype=np.uint32)
        uctx = np.zeros_like(ctx)
        uobj = np.zeros_like(obj)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        skidx_abs = np.ascontiguousarray(skidx_abs, dtype=np.uint64)
        self.check(self.lib.kssdhip_count_block(h, _p(keys), _p(skidx_abs), nq, qry_gid_offset,
                                                flags, _p(ctx), _p(obj), _p(uctx), _p(uobj)))
        return ctx, obj, uctx, uobj

    def count_block_dev(self, h, d_keys, d_skidx, nq, total, qry_gid_offset, flags,
                        d_ctx, d_obj, d_uctx=None, d_uobj=None, stream=None):
        self.check(self.lib.kssdhip_count_block_dev(h, d_keys, d_skidx, nq, total, qry_gid_offset,
                                                    flags, d_ctx, d_obj, d_uctx, d_uobj, stream))

    def find_first(self, h, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        idx = np.zeros(len(keys), dtype=np.uint64)
        self.check(self.lib.kssdhip_find_first(h, _p(keys), len(keys), _p(idx)))
        return idx

    def pair(self, a, b, obj_bits, ignore_b_conflict=False):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        f = Features()
        self.check(self.lib.kssdhip_pair_features(_p(a), len(a), _p(b), len(b), obj_bits,
                                                  int(ignore_b_conflict), ctypes.byref(f)))
        return f.tup()

    def pair_dev(self, d_a, n, d_b, m, obj_bits, ignore, d_out, d_ws, stream=None):
        self.check(self.lib.kssdhip_pair_features_dev(d_a, n, d_b, m, obj_bits, int(ignore),
                                                      d_out, d_ws, stream))

    # ---- serving loop (kssdhip_serve_*)
    def serve_open(self, h, budget_bytes, slots=2, max_batch=64, max_queued=4096, flags=0, select_opts=None,
                   ref_ctx=None, ntop=0):
        """A server returning counter rows, or with select_opts (and the
        references' context totals) ranked passing rows."""
        srv = V()
        rc = None
        if select_opts is not None:
            rc = np.ascontiguousarray(ref_ctx, dtype=np.uint32)
        opts = ServeOpts(budget_bytes, slots, max_batch, max_queued, flags, int(select_opts is not None),
                         select_opts if select_opts is not None else SelectOpts(), _p(rc), ntop)
        self.check(self.lib.kssdhip_serve_open(h, ctypes.byref(opts), ctypes.byref(srv)))
        return srv

    def serve_wait_rows(self, srv, ticket, timeout_ms=-1, cap=256):
        """(ROW_DTYPE array of the request's ranked rows, ServeStats); the
        buffer grows on KSSDHIP_ENOSPC."""
        st = ServeStats()
        n = S()
        while True:
            rows = np.zeros(max(cap, 1), dtype=ROW_DTYPE)
            rc = self.lib.kssdhip_serve_wait_rows(srv, ticket, _p(rows), cap, ctypes.byref(n), timeout_ms,
                                                  ctypes.byref(st))
            if rc == ENOSPC:
                cap = int(n.value)
                continue
            self.check(rc)
            return rows[: int(n.value)].copy(), st

    def serve_submit(self, srv, keys):
        """Ticket (>= 0) or the negative error code (EAGAIN: queue full)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        return int(self.lib.kssdhip_serve_submit(srv, _p(keys), len(keys)))

    def serve_wait(self, srv, ticket, ref_gnum, unique=False, timeout_ms=-1):
        ctx = np.zeros((ref_gnum, 2), np.uint32)
        obj = np.zeros((ref_gnum, 2), np.uint32)
        uctx = np.zeros_like(ctx) if unique else None
        uobj = np.zeros_like(obj) if unique else None
        st = ServeStats()
        self.check(self.lib.kssdhip_serve_wait(srv, ticket, _p(ctx), _p(obj), _p(uctx), _p(uobj), timeout_ms,
                                               ctypes.byref(st)))
        return ctx, obj, uctx, uobj, st

    def serve_info(self, srv):
        info = ServeInfo()
        self.check(self.lib.kssdhip_serve_info(srv, ctypes.byref(info)))
        return info

    def serve_close(self, srv):
        self.check(self.lib.kssdhip_serve_close(srv))

    def epilogue_dev(self, d_ctx, d_obj, nq, nr, d_qctx, d_rctx, opts, d_out, stream=None):
        self.check(self.lib.kssdhip_epilogue_dev(d_ctx, d_obj, nq, nr, d_qctx, d_rctx,
                                                 ctypes.byref(opts), d_out, stream))

    def select_block(self, h, keys, skidx_abs, qry_ctx, ref_ctx, opts, flags=0, cap=None):
        """Passing pairs of the index path's gate as a ROW_DTYPE array
        ordered by (query, ref); the buffer grows on KSSDHIP_ENOSPC."""
        nq = len(skidx_abs) - 1
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        skidx_abs = np.ascontiguousarray(skidx_abs, dtype=np.uint64)
        qry_ctx = np.ascontiguousarray(qry_ctx, dtype=np.uint32)
        ref_ctx = np.ascontiguousarray(ref_ctx, dtype=np.uint32)
        cap = max(1, 64 * nq) if cap is None else cap
        while True:
            rows = np.zeros(max(cap, 1), dtype=ROW_DTYPE)
            n = S()
            rc = self.lib.kssdhip_select_block(h, _p(keys) if len(keys) else None, _p(skidx_abs), nq, flags,
                                               _p(qry_ctx), _p(ref_ctx), ctypes.byref(opts), _p(rows), cap,
                                               ctypes.byref(n))
            if rc == ENOSPC:
                cap = int(n.value)
                continue
            self.check(rc)
            return rows[: int(n.value)].copy()

    def select_block_dev(self, h, d_keys, d_skidx, nq, total, flags, d_qctx, d_rctx, opts, d_rows, cap,
                         d_count, d_first, d_qcount, stream=None):
        self.check(self.lib.kssdhip_select_block_dev(h, d_keys, d_skidx, nq, total, flags, d_qctx, d_rctx,
                                                     ctypes.byref(opts), d_rows, cap, d_count, d_first, d_qcount,
                                                     stream))

    def select_block_dev_noalloc(self, h, d_keys, d_skidx, nq, total, flags, d_qctx, d_rctx, opts, d_rows, cap,
                                 d_count, d_first, d_qcount, stream=None):
        """select_block_dev with its workspace in the library's static
        per-device scratch array (no allocation; calls on one device must
        not overlap)."""
        self.check(self.lib.kssdhip_select_block_dev_noalloc(h, d_keys, d_skidx, nq, total, flags, d_qctx, d_rctx,
                                                             ctypes.byref(opts), d_rows, cap, d_count, d_first,
                                                             d_qcount, stream))

    def probe_random_reads(self, nbytes=2 << 30, reads=400_000_000):
        """Random 128-byte line reads per second from an nbytes buffer (the join's load form)."""
        r = D()
        self.check(self.lib.kssdhip_probe_random_reads(nbytes, reads, ctypes.byref(r)))
        return float(r.value)

    def table_info(self, h):
        """(lines, overflow lines, home bits, context bits) of the handle's record table."""
        lines, ovf, hb, cb = U64(), U64(), I(), I()
        self.check(self.lib.kssdhip_index_table_info(h, ctypes.byref(lines), ctypes.byref(ovf), ctypes.byref(hb),
                                                     ctypes.byref(cb)))
        return int(lines.value), int(ovf.value), int(hb.value), int(cb.value)

    def matched_records(self, h, d_keys, d_skidx, nq, total):
        m = U64()
        self.check(self.lib.kssdhip_index_matched_records(h, d_keys, d_skidx, nq, total, ctypes.byref(m)))
        return int(m.value)

    def select_hits_dev(self, d_values, nq, nr, min_value, cap, d_hits, d_count, stream=None):
        self.check(self.lib.kssdhip_select_hits_dev(d_values, nq, nr, float(min_value), cap, d_hits, d_count,
                                                    stream))


# ------------------------------------------------------------------ libkssdhost
def load_kssdhost():
    so = PKG / "libkssdhost.so"
    if not so.exists():
        raise RuntimeError(f"{so} is not built: run `make -C kssd3a-public_amd`")
    lib = ctypes.CDLL(str(so))
    lib.kssd_u64vec_free.argtypes = [ctypes.POINTER(U64Vec)]
    lib.kssd_u64vec_free.restype = None
    lib.kssd_sketch_fasta_file.argtypes = [ctypes.c_char_p, I, I, ctypes.POINTER(U64Vec)]
    lib.kssd_sketchdir_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(KssdStat), V, V, V]
    lib.kssd_write_index_files.argtypes = [ctypes.c_char_p, V, S, V]
    lib.kssd_sketching_id.argtypes = [I, I, I, I]
    lib.kssd_sketching_id.restype = U32
    lib.kssd_simsketch.argtypes = [V, I, I, I, I, ctypes.POINTER(ctypes.POINTER(U64)), V]
    lib.kssd_sim_genome.argtypes = [ctypes.POINTER(SimSpec), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                                    ctypes.POINTER(S)]
    lib.kssd_build_index.argtypes = [V, V, I, U, I, V]
    lib.kssd_sample_ctx_counts.argtypes = [V, V, I, U, I, I, V]
    lib.kssd_sample_ctx_counts.restype = None
    lib.kssd_sketch_codes.argtypes = [V, S, I, ctypes.POINTER(U64Vec)]
    lib.kssd_sketch_codes.restype = None
    lib.kssd_finalize_keys.argtypes = [ctypes.POINTER(U64Vec), I]
    lib.kssd_finalize_keys.restype = None
    for f in ("kssd_emit_index_detail", "kssd_emit_hash_detail"):
        getattr(lib, f).argtypes = [ctypes.POINTER(Buf), ctypes.POINTER(ReportOpts),
                                    ctypes.POINTER(QueryCounts), ctypes.c_char_p, V, V]
        getattr(lib, f).restype = None
    lib.kssd_emit_stream_detail.argtypes = [ctypes.POINTER(Buf), ctypes.POINTER(ReportOpts),
                                            ctypes.POINTER(QueryCounts), ctypes.c_char_p, V, V, I]
    lib.kssd_emit_stream_detail.restype = None
    lib.kssd_emit_index_matrix_row.argtypes = [ctypes.POINTER(Buf), ctypes.POINTER(ReportOpts),
                                               ctypes.POINTER(QueryCounts), ctypes.c_char_p]
    lib.kssd_emit_index_matrix_row.restype = None
    lib.kssd_buf_free.argtypes = [ctypes.POINTER(Buf)]
    lib.kssd_buf_free.restype = None
    lib.kssd_matrix_value.argtypes = [ctypes.POINTER(ReportOpts), U32, U32, U32, U32, U32, U32, V, V]
    lib.kssd_matrix_value.restype = D
    lib.kssd_matrix_metric.argtypes = [ctypes.c_char_p]
    lib.kssd_matrix_context_value.argtypes = [I, D, U32, U32, U32, U32, U32, U32, I, U, U]
    lib.kssd_matrix_context_value.restype = D
    lib.kssd_pairwise_value.argtypes = [I, D, S, S, U32, U32, U32, U32, U32, U32, I, U, U]
    lib.kssd_pairwise_value.restype = D
    lib.kssd_matrix_put_fixed6.argtypes = [ctypes.POINTER(Buf), D]
    lib.kssd_matrix_scale_fixed6.argtypes = [D]
    lib.kssd_matrix_scale_fixed6.restype = U32
    lib.kssd_matrix_put_scaled6.argtypes = [ctypes.POINTER(Buf), U32]
    lib.kssd_refaf_predict_ani.argtypes = [D, D, U, U, U, U, I]
    lib.kssd_refaf_predict_ani.restype = D
    return lib


def default_threads():
    n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


class CountOpts(ctypes.Structure):
    """kssd_count_opts_t: the sketcher's k-mer count options (-n,
    --npercentile, --ncap, --readsQC, -A)."""
    _fields_ = [("min_count", I), ("percentile", D), ("cap", I), ("reads_qc", I), ("record_qc", I)]


class CountFilter(ctypes.Structure):
    _fields_ = [("raw_cutoff", U32), ("cutoff", U32), ("upper", U32), ("mode", U32), ("used_reads_qc", I),
                ("qc", U32 * 4)]


class Host:
    """numpy-facing wrapper of libkssdhost.so."""
    _libc = ctypes.CDLL(None)

    def __init__(self):
        self.lib = load_kssdhost()
        self._libc.free.argtypes = [V]

    def sketch_file(self, path, drfold=8, keep_conflict=False):
        v = U64Vec()
        if self.lib.kssd_sketch_fasta_file(str(path).encode(), drfold, int(keep_conflict),
                                           ctypes.byref(v)) != 0:
            raise OSError(f"cannot sketch {path}")
        out = np.ctypeslib.as_array(v.a, shape=(v.n,)).copy() if v.n else np.zeros(0, np.uint64)
        self.lib.kssd_u64vec_free(ctypes.byref(v))
        return out

    def sketch_pooled(self, paths, drfold=8, keep_conflict=True, record_qc=True, threads=None):
        """The read-set sketch of `sketch --asone [--conflict] [-A]` over
        FASTA/FASTQ files, in process: (sorted keys, per-key counts)."""
        arr = (ctypes.c_char_p * len(paths))(*[str(p).encode() for p in paths])
        v = U64Vec()
        counts = ctypes.POINTER(U32)()
        flt = CountFilter()
        bad = I(-1)
        co = CountOpts(1, 0.0, 0, 0, int(record_qc))
        self.lib.kssd_sketch_pooled.argtypes = [V, I, I, I, I, ctypes.POINTER(CountOpts), ctypes.POINTER(U64Vec),
                                                ctypes.POINTER(ctypes.POINTER(U32)), ctypes.POINTER(CountFilter),
                                                ctypes.POINTER(I)]
        if self.lib.kssd_sketch_pooled(ctypes.cast(arr, V), len(paths), drfold, int(threads or default_threads()),
                                       int(keep_conflict), ctypes.byref(co), ctypes.byref(v), ctypes.byref(counts),
                                       ctypes.byref(flt), ctypes.byref(bad)) != 0:
            raise OSError(f"cannot sketch {paths[bad.value] if bad.value >= 0 else 'the inputs'}")
        keys = np.ctypeslib.as_array(v.a, shape=(v.n,)).copy() if v.n else np.zeros(0, np.uint64)
        cnt = np.ctypeslib.as_array(counts, shape=(v.n,)).copy() if (v.n and counts) else np.zeros(0, np.uint32)
        self.lib.kssd_u64vec_free(ctypes.byref(v))
        if counts:
            self._libc.free(ctypes.cast(counts, V))
        return keys, cnt

    def sim_fastq_pair(self, codes, coverage, seed, path1, path2, read_len=150, insert=300, sd=30, err=0.01):
        """Paired 2 x read_len FASTQ files of a genome (2-bit codes) at
        `coverage` (kssd_sim_fastq_pair); returns the number of pairs."""
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        self.lib.kssd_sim_fastq_pair.argtypes = [V, S, D, U64, I, I, I, D, ctypes.c_char_p, ctypes.c_char_p]
        self.lib.kssd_sim_fastq_pair.restype = ctypes.c_long
        n = self.lib.kssd_sim_fastq_pair(_p(codes), len(codes), float(coverage), int(seed), read_len, insert, sd,
                                         float(err), str(path1).encode(), str(path2).encode())
        if n < 0:
            raise OSError(f"cannot write {path1} / {path2}")
        return int(n)

    def simsketch(self, specs, drfold=8, keep_conflict=False, threads=None):
        arr = (SimSpec * len(specs))(*specs)
        keys = ctypes.POINTER(U64)()
        index = np.zeros(len(specs) + 1, np.uint64)
        if self.lib.kssd_simsketch(arr, len(specs), drfold, int(keep_conflict),
                                   int(threads or default_threads()), ctypes.byref(keys),
                                   _p(index)) != 0:
            raise RuntimeError("simsketch failed")
        n = int(index[-1])
        out = np.ctypeslib.as_array(keys, shape=(n,)).copy() if n else np.zeros(0, np.uint64)
        self._libc.free(ctypes.cast(keys, V))
        return out, index

    def sim_genome(self, spec):
        """The simulator's genome for `spec` as 2-bit base codes (uint8 0..3)."""
        codes = ctypes.POINTER(ctypes.c_uint8)()
        n = S()
        if self.lib.kssd_sim_genome(ctypes.byref(spec), ctypes.byref(codes), ctypes.byref(n)) != 0:
            raise RuntimeError("sim_genome failed")
        out = np.ctypeslib.as_array(codes, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint8)
        self._libc.free(ctypes.cast(codes, V))
        return out

    def write_sketch(self, path, keys, index, names, drfold=8, conflict=False, records=None, unique_bits=None):
        """A sketch directory (lcofiles.stat, comblco, comblco.index) of
        coden-pattern keys; with `records`, also its sorted index files."""
        Path(path).mkdir(parents=True, exist_ok=True)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        index = np.ascontiguousarray(index, dtype=np.uint64)
        st = KssdStat(int(self.lib.kssd_sketching_id(11, 0, 0, drfold)), False, conflict, 11, 32, 11, 0, drfold,
                      len(index) - 1)
        nm = (ctypes.c_char * 256 * max(len(names), 1))()
        for i, n in enumerate(names):
            nm[i].value = n.encode()[:255]
        if self.lib.kssd_sketchdir_write(str(path).encode(), ctypes.byref(st), ctypes.cast(nm, V), _p(keys),
                                         _p(index)) != 0:
            raise OSError(f"cannot write the sketch {path}")
        if records is not None:
            ub = np.frombuffer(unique_bits, np.uint8) if unique_bits is not None else None
            if self.lib.kssd_write_index_files(str(path).encode(), _p(records), len(records),
                                               _p(ub) if ub is not None else None) != 0:
                raise OSError(f"cannot write the index of {path}")

    def build_index(self, keys, index, obj_bits=20, threads=None):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        index = np.ascontiguousarray(index, dtype=np.uint64)
        out = np.empty(len(keys), dtype=REC_DTYPE)
        if self.lib.kssd_build_index(_p(keys), _p(index), len(index) - 1, obj_bits,
                                     int(threads or default_threads()), _p(out)) != 0:
            raise ValueError("index limits exceeded")
        return out

    def emit(self, path, opts, ctx, obj, qry_ctx, ref_ctx, qry_name, ref_names, uctx=None, uobj=None,
             ref_anno=None):
        """Detail rows of one query on one of the reference's output paths
        ('index', 'stream', 'hash') or its index-path matrix row ('matrix')."""
        ctx = np.ascontiguousarray(ctx, dtype=np.uint32)
        obj = np.ascontiguousarray(obj, dtype=np.uint32)
        ref_ctx = np.ascontiguousarray(ref_ctx, dtype=np.uint32)
        uctx = None if uctx is None else np.ascontiguousarray(uctx, dtype=np.uint32)
        uobj = None if uobj is None else np.ascontiguousarray(uobj, dtype=np.uint32)
        qc = QueryCounts(_p(ctx), _p(obj), _p(uctx), _p(uobj), int(qry_ctx), _p(ref_ctx),
                         len(ref_ctx), None, None)
        names = ctypes.create_string_buffer(b"".join(n.encode().ljust(256, b"\0") for n in ref_names),
                                            256 * len(ref_names))
        anno = None
        if ref_anno is not None:
            anno = ctypes.create_string_buffer(b"".join(n.encode().ljust(256, b"\0") for n in ref_anno),
                                               256 * len(ref_anno))
        b = Buf()
        if path == "matrix":
            self.lib.kssd_emit_index_matrix_row(ctypes.byref(b), ctypes.byref(opts), ctypes.byref(qc),
                                                qry_name.encode())
        elif path in ("stream", "multistream"):
            self.lib.kssd_emit_stream_detail(ctypes.byref(b), ctypes.byref(opts), ctypes.byref(qc),
                                             qry_name.encode(), names, anno, int(path == "stream"))
        else:
            fn = getattr(self.lib, f"kssd_emit_{path}_detail")
            fn(ctypes.byref(b), ctypes.byref(opts), ctypes.byref(qc), qry_name.encode(), names, anno)
        text = ctypes.string_at(b.s, b.l).decode() if b.l else ""
        self.lib.kssd_buf_free(ctypes.byref(b))
        return text

    def ctx_counts(self, keys, index, obj_bits=20, conflict=False, ignoreconflict=False):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        index = np.ascontiguousarray(index, dtype=np.uint64)
        out = np.zeros(len(index) - 1, np.uint32)
        self.lib.kssd_sample_ctx_counts(_p(keys), _p(index), len(index) - 1, obj_bits,
                                        int(conflict), int(ignoreconflict), _p(out))
        return out
