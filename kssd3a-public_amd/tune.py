"""Autotuner for the counting path's kernel shape (kssdhip_tiling_t).

For a resident index and a device-resident query block it times each
candidate tiling with HIP events on the caller's stream -- the lookup
kernel and the partitioned merge-join over a grid of LDS chunk sizes,
workgroup sizes and partition groups (partitions walked by one workgroup,
which sums its hits per cell in LDS) -- and keeps the fastest.  Results are cached in a JSON
file keyed by the device name and the shape of the problem (index records,
query keys, queries, references, flags), so a serving process or a later
run picks the shape up without re-measuring.
"""
import ctypes
import json
import os
from pathlib import Path

import kssdhip as kh

DEFAULT_CACHE = Path(os.environ.get("KSSDHIP_TUNE_CACHE",
                                    Path.home() / ".cache" / "kssdhip" / "tuning.json"))
CHUNKS = (1024, 2048, 4096, 8192, 12288)
BLOCKS = (128, 256, 512)
GROUP_CHUNKS = (2048, 3072, 4096, 6144)  # partition groups need chunk_records <= 6144
GROUP_BLOCKS = (512, 1024)
GROUPS = (2, 4, 8)


def _bucket(x):
    """Shape keys are rounded to a power of two so nearby sizes share an entry."""
    x = int(x)
    return 0 if x <= 0 else 1 << (x.bit_length() - 1)


def device_name(rt):
    try:
        rt.lib.hipDeviceGetName.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        buf = ctypes.create_string_buffer(256)
        rt.check(rt.lib.hipDeviceGetName(buf, 256, rt.device))
        return buf.value.decode() or "hip"
    except Exception:
        return "hip"


def shape_key(dev, n_index, total_keys, nq, nref, flags):
    return f"{dev}|N{_bucket(n_index)}|K{_bucket(total_keys)}|Q{_bucket(nq)}|R{_bucket(nref)}|F{flags}"


def load_cache(path=DEFAULT_CACHE):
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=DEFAULT_CACHE):
    try:
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        tmp = Path(str(path) + ".tmp")
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass


def cu_count(rt):
    """Compute units of the runtime's device (hipDeviceAttributeMultiprocessorCount = 63)."""
    try:
        v = ctypes.c_int(0)
        rt.lib.hipDeviceGetAttribute.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
        rt.check(rt.lib.hipDeviceGetAttribute(ctypes.byref(v), 63, rt.device))
        return v.value
    except Exception:
        return 0


def fitted_chunks(n_index, cus):
    """(chunk, group) pairs whose partition groups come to one to four whole
    rounds of workgroups over the compute units (one workgroup per CU at a
    time), where the fixed chunk sizes leave the last round part-empty."""
    out = []
    if n_index <= 0 or cus <= 0:
        return out
    for g in GROUPS:
        for rounds in (1, 2, 3, 4):
            c = -(-n_index // (cus * g * rounds))
            c = (c + 63) // 64 * 64
            if GROUP_CHUNKS[0] <= c <= GROUP_CHUNKS[-1] and c not in GROUP_CHUNKS and (c, g) not in out:
                out.append((c, g))
    return out


def candidates(n_index=0, cus=0):
    yield {"mode": 1, "chunk_records": 4096, "block_threads": 256, "group_partitions": 1}
    for c in CHUNKS:
        for b in BLOCKS:
            yield {"mode": 2, "chunk_records": c, "block_threads": b, "group_partitions": 1}
    for c in GROUP_CHUNKS:
        for b in GROUP_BLOCKS:
            for g in GROUPS:
                yield {"mode": 2, "chunk_records": c, "block_threads": b, "group_partitions": g}
    for c, g in fitted_chunks(n_index, cus):
        for b in GROUP_BLOCKS:
            yield {"mode": 2, "chunk_records": c, "block_threads": b, "group_partitions": g}


def candidate_name(c):
    name = f"m{c['mode']}_c{c['chunk_records']}_b{c['block_threads']}"
    g = c.get("group_partitions", 1)
    return name if g == 1 else f"{name}_g{g}"


def tune(hip, rt, h, d_keys, d_skidx, nq, total, nref, n_index, d_ctx, d_obj, stream,
         flags=0, reps=3, cache_path=DEFAULT_CACHE, force=False, log=None):
    """Pick and apply the fastest tiling for this block shape.  Returns
    (tiling dict, {candidate: ms}).  The outputs d_ctx/d_obj are scratch."""
    dev = device_name(rt)
    key = shape_key(dev, n_index, total, nq, nref, flags)
    cache = load_cache(cache_path)
    if not force and key in cache:
        best = dict(cache[key]["best"])
        best.setdefault("group_partitions", 1)
        hip.set_tiling(h, mode=best["mode"], chunk_records=best["chunk_records"],
                       block_threads=best["block_threads"], group_partitions=best["group_partitions"])
        return best, cache[key].get("times", {})
    e0, e1 = rt.event(), rt.event()
    times = {}
    best, best_ms = None, float("inf")
    for cand in candidates(n_index, cu_count(rt)):
        hip.set_tiling(h, **cand)
        hip.count_block_dev(h, d_keys, d_skidx, nq, total, 0, flags, d_ctx, d_obj, None, None, stream)
        rt.sync(stream)
        rt.record(e0, stream)
        for _ in range(reps):
            hip.count_block_dev(h, d_keys, d_skidx, nq, total, 0, flags, d_ctx, d_obj, None, None,
                                stream)
        rt.record(e1, stream)
        ms = rt.elapsed_ms(e0, e1) / reps
        name = candidate_name(cand)
        times[name] = ms
        if log:
            log(f"tune {name}: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = cand, ms
    hip.set_tiling(h, **best)
    cache[key] = {"best": best, "ms": best_ms, "times": times}
    save_cache(cache, cache_path)
    return best, times
