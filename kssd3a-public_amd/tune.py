"""Autotuner front end: the counting path's kernel shape for a block, from
the library's tuner (kssdhip_index_autotune) and its JSON cache.

The library times every candidate shape -- the lookup kernel, the
partitioned merge-join over LDS chunk sizes, workgroup sizes and partition
groups, and the query-major join -- on a device-resident block, applies the
fastest and caches it keyed by the device and the block's shape; a later
call (in this process, another one, or the kssd3a-hip CLI) takes the cached
shape without timing.
"""
import os
from pathlib import Path

DEFAULT_CACHE = Path(os.environ.get("KSSDHIP_TUNE_CACHE",
                                    Path.home() / ".cache" / "kssdhip" / "tuning.json"))


def tiling_dict(t):
    return {"mode": t.mode, "chunk_records": t.chunk_records, "block_threads": t.block_threads,
            "group_partitions": t.group_partitions, "query_threads": t.query_threads,
            "query_slots": t.query_slots}


def tune(hip, rt, h, d_keys, d_skidx, nq, total, nref, n_index, d_ctx, d_obj, stream,
         flags=0, reps=3, cache_path=DEFAULT_CACHE, force=False, log=None):
    """Pick and apply the fastest tiling for this block shape (the library's
    tuner).  Returns (tiling dict, summary dict).  nref, n_index, d_ctx,
    d_obj, stream and reps are accepted for the callers' convenience; the
    library uses the handle's own stream and scratch outputs."""
    r = hip.autotune(h, d_keys, d_skidx, nq, total, flags, cache_path=cache_path, force=force)
    best = tiling_dict(r.tiling)
    summary = {"best_ms": r.ms, "from_cache": bool(r.from_cache), "candidates": r.candidates,
               "cache": str(cache_path)}
    if log:
        log(f"tune: {best} {'from the cache' if r.from_cache else f'of {r.candidates} shapes'}, {r.ms:.3f} ms")
    return best, summary
