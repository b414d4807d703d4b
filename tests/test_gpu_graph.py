"""The batched step (counter fill kernels + partitioned join + FP64 epilogue)
captured in a HIP graph: every replay must reproduce the eager result,
with eager work and synchronisations between the replays and the outputs
overwritten with garbage before each replay (so a replay that skipped its
zeroing or its join would show)."""
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


@pytest.mark.parametrize("mode,group", [(1, 1), (2, 1), (2, 4)])
def test_graph_replays_match_eager(hip, oracle, mode, group):
    rt = kt.Hip(0)
    rng = np.random.default_rng(21)
    pool = rng.integers(1, 1 << 40, size=900, dtype=np.uint64)
    refs = kt.random_samples(rng, 30, 900, 200, shared_pool=pool, conflict_p=0.2)
    qrys = kt.random_samples(rng, 25, 900, 200, shared_pool=pool, conflict_p=0.2)
    rk, ri = _block(refs)
    qk, qi = _block(qrys)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    R, Q = len(refs), len(qrys)
    h = hip.open_index(rec, R, OBJ_BITS)
    s = rt.stream()
    bufs = []
    try:
        hip.set_tiling(h, mode=mode, chunk_records=256, block_threads=256, group_partitions=group)
        d_k, d_i = rt.upload(qk), rt.upload(qi)
        qctx = np.diff(qi).astype(np.uint32)
        rctx = np.diff(ri).astype(np.uint32)
        d_qc, d_rc = rt.upload(qctx), rt.upload(rctx)
        cells = Q * R
        d_ctx, d_obj, d_out = rt.malloc(8 * cells), rt.malloc(8 * cells), rt.malloc(8 * cells)
        bufs = [d_k, d_i, d_qc, d_rc, d_ctx, d_obj, d_out]
        opts = kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0)
        total = int(qi[-1])

        def step():
            hip.count_block_dev(h, d_k, d_i, Q, total, 0, 0, d_ctx, d_obj, None, None, s)
            hip.epilogue_dev(d_ctx, d_obj, Q, R, d_qc, d_rc, opts, d_out, s)

        def results():
            rt.sync(s)
            return (rt.download(d_ctx, np.uint32, 2 * cells).copy(), rt.download(d_obj, np.uint32, 2 * cells).copy(),
                    rt.download(d_out, np.float64, cells).copy())

        step()  # eager; also sizes every workspace before capture
        eager = results()
        want = oracle.count_block(rec, R, OBJ_BITS, qk, qi)
        assert np.array_equal(eager[0].reshape(Q, R, 2), want[0])
        assert np.array_equal(eager[1].reshape(Q, R, 2), want[1])
        g, ex = rt.capture(s, step)
        try:
            for rep in range(4):
                # garbage in the outputs, eager work and a sync between replays
                for d in (d_ctx, d_obj, d_out):
                    rt.check(rt.lib.hipMemsetAsync(d, 0x5A + rep, 8 * cells, s))
                hip.count_block_dev(h, d_k, d_i, Q, total, 0, 0, d_ctx, d_obj, None, None, s)
                rt.sync(s)
                rt.graph_launch(ex, s)
                got = results()
                assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1]), rep
                assert np.array_equal(got[2], eager[2], equal_nan=True), rep
        finally:
            rt.graph_destroy(g, ex)
    finally:
        for b in bufs:
            rt.free(b)
        hip.close_index(h)
