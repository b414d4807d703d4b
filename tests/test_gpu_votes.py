"""Votes of the sparse-edge nomination (kssdhip_count_votes) against a
numpy restatement of the vote loop of pairwise_indexed_self_scan
(pairwise_graph.c:1200-1235): for query q (genome qgid), every
sample_step-th key from its first; a key whose context has more than
max_ctx_freq index records (when > 0) is skipped; otherwise each record of
the context with gid < qgid adds one vote to (q, gid).  Non-conflict
sketches only, as that path requires.  Lookup and partitioned kernels,
with and without partition groups."""
import numpy as np
import pytest

import kssdtest as kt

pytestmark = pytest.mark.gpu
OBJ_BITS = 20


def _block(samples):
    keys = np.concatenate(samples) if samples else np.zeros(0, np.uint64)
    idx = np.zeros(len(samples) + 1, np.uint64)
    idx[1:] = np.cumsum([len(s) for s in samples])
    return keys, idx


def _want_votes(rec, keys, idx, q0, nq, n_ref, max_freq, step):
    ctx = rec["ctxgid"] >> np.uint64(20)
    gid = (rec["ctxgid"] & np.uint64((1 << 20) - 1)).astype(np.int64)
    votes = np.zeros((nq, n_ref), np.uint32)
    for q in range(nq):
        qgid = q0 + q
        for key in keys[idx[qgid]:idx[qgid + 1]][::step]:
            c = key >> np.uint64(OBJ_BITS)
            lo, hi = np.searchsorted(ctx, c, "left"), np.searchsorted(ctx, c, "right")
            if lo == hi or (max_freq and hi - lo > max_freq):
                continue
            for g in gid[lo:hi]:
                if g < qgid:
                    votes[q, g] += 1
    return votes


@pytest.mark.parametrize("mode,group", [(1, 1), (2, 1), (2, 4)])
@pytest.mark.parametrize("max_freq,step", [(0, 1), (3, 1), (0, 3), (5, 2)])
def test_votes_match_reference_loop(hip, mode, group, max_freq, step):
    rng = np.random.default_rng(100 + max_freq * 7 + step)
    pool = rng.integers(1, 1 << 40, size=1500, dtype=np.uint64)
    samples = kt.random_samples(rng, 40, 1500, 250, shared_pool=pool, conflict_p=0.0)
    samples[5] = np.zeros(0, np.uint64)
    keys, idx = _block(samples)
    rec = kt.build_index_np(keys, idx, OBJ_BITS)
    n = len(samples)
    h = hip.open_index(rec, n, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=mode, chunk_records=256, block_threads=256, group_partitions=group)
        for q0, nq in ((0, n), (7, 20)):
            got = hip.count_votes(h, n, keys[idx[q0]:], idx[q0:q0 + nq + 1], qry_gid_offset=q0,
                                  max_ctx_freq=max_freq, sample_step=step)
            want = _want_votes(rec, keys, idx, q0, nq, n, max_freq, step)
            assert np.array_equal(got, want), (q0, nq)
            assert want.any()
    finally:
        hip.close_index(h)
