"""The library's autotuner and its JSON cache, through the C ABI: a block's
candidates timed and the fastest applied and cached; the cache entry applied
without timing by a later call and by kssdhip_index_tiling_from_cache; a new
handle on an index of the same shape starting from the cached shape
parameters; the CLI tuning with KSSDHIP_TUNE=1 and reading the cache
otherwise, its output unchanged."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import kssdtest as kt

pytestmark = pytest.mark.gpu
SCALE = kt.GOLDEN / "scale"


def _c2():
    ref = kt.read_sketch_dir(SCALE / "c2" / "ref")
    qry = kt.read_sketch_dir(SCALE / "c2" / "qry")
    return ref, qry, kt.build_index_np(ref.keys, ref.index, ref.obj_bits)


def test_autotune_caches_and_applies(hip, oracle, tmp_path):
    rt = kt.Hip(0)
    ref, qry, rec = _c2()
    R, Q = ref.n, qry.n
    cache = tmp_path / "tuning.json"
    d_k, d_i = rt.upload(qry.keys), rt.upload(qry.index)
    h = hip.open_index(rec, R, ref.obj_bits)
    try:
        r = hip.autotune(h, d_k, d_i, Q, int(qry.index[-1]), 0, cache_path=cache)
        assert not r.from_cache and r.candidates > 20 and r.cached and r.ms > 0
        t = hip.get_tiling(h)
        assert (t.mode, t.chunk_records, t.block_threads, t.group_partitions) == (
            r.tiling.mode, r.tiling.chunk_records, r.tiling.block_threads, r.tiling.group_partitions)
        got = hip.count_block(h, R, qry.keys, qry.index)
        want = oracle.count_block(rec, R, ref.obj_bits, qry.keys, qry.index)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        entries = json.loads(cache.read_text())["entries"]
        assert len(entries) == 1
        (key, e), = entries.items()
        assert f"|R{1 << (R.bit_length() - 1)}|" in key and e["mode"] == r.tiling.mode
        # cached: applied again without timing
        hip.set_tiling(h, mode=1)
        r2 = hip.autotune(h, d_k, d_i, Q, int(qry.index[-1]), 0, cache_path=cache)
        assert r2.from_cache and r2.candidates == 0 and hip.get_tiling(h).mode == r.tiling.mode
        hip.set_tiling(h, mode=1)
        applied, r3 = hip.tiling_from_cache(h, Q, int(qry.index[-1]), 0, cache_path=cache)
        assert applied and hip.get_tiling(h).mode == r.tiling.mode
        applied, _ = hip.tiling_from_cache(h, Q, int(qry.index[-1]), kt.UNIQUE, cache_path=cache)
        assert not applied
    finally:
        hip.close_index(h)
        rt.free(d_k)
        rt.free(d_i)
    # a new handle on an index of this shape starts from the cached parameters
    old = {k: os.environ.get(k) for k in ("KSSDHIP_NO_TUNE_CACHE", "KSSDHIP_TUNE_CACHE")}
    os.environ.pop("KSSDHIP_NO_TUNE_CACHE", None)
    os.environ["KSSDHIP_TUNE_CACHE"] = str(cache)
    try:
        h2 = hip.open_index(rec, R, ref.obj_bits)
        t2 = hip.get_tiling(h2)
        hip.close_index(h2)
        assert t2.mode == 0
        assert (t2.chunk_records, t2.block_threads, t2.group_partitions) == (
            r.tiling.chunk_records, r.tiling.block_threads, r.tiling.group_partitions)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_autotune_host_block_with_unique_flags(hip, oracle, tmp_path):
    ref, qry, rec = _c2()
    h = hip.open_index(rec, ref.n, ref.obj_bits)
    try:
        r = hip.autotune_host(h, qry.keys, qry.index, kt.UNIQUE | kt.IGNORECONFLICT, cache_path=tmp_path / "t.json")
        assert r.candidates > 20 and not r.from_cache
        got = hip.count_block(h, ref.n, qry.keys, qry.index, flags=kt.UNIQUE | kt.IGNORECONFLICT)
        want = oracle.count_block(rec, ref.n, ref.obj_bits, qry.keys, qry.index, flags=kt.UNIQUE | kt.IGNORECONFLICT)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
    finally:
        hip.close_index(h)


def test_cli_tunes_then_reads_the_cache(tmp_path):
    """`ani` with KSSD3A_FORCE_REF_INDEX on the C2 slice: KSSDHIP_TUNE=1
    writes the cache, a second run reads it; both print the reference's
    bytes."""
    shutil.copytree(SCALE / "c2", tmp_path / "c2")
    want = (SCALE / "c2_best_index.tsv").read_text()
    cache = tmp_path / "cache" / "tuning.json"
    env = dict(os.environ, KSSDHIP_TUNE_CACHE=str(cache))
    env.pop("KSSDHIP_NO_TUNE_CACHE", None)
    argv = [str(kt.PKG / "kssd3a-hip"), "ani", "-r", "c2/ref", "-q", "c2/qry"]
    p = subprocess.run(argv, capture_output=True, text=True, cwd=tmp_path, env=dict(env, KSSDHIP_TUNE="1"),
                       timeout=300)
    assert p.returncode == 0, p.stderr
    assert p.stdout == want
    assert len(json.loads(cache.read_text())["entries"]) >= 1
    p = subprocess.run(argv, capture_output=True, text=True, cwd=tmp_path, env=env, timeout=300)
    assert p.returncode == 0 and p.stdout == want
