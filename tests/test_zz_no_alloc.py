"""kssdhip_select_block_dev_noalloc (workspace in the library's static
per-device scratch array) against kssdhip_select_block_dev (the handle's
grow-only workspace), through the C ABI: the same passing rows, bit for bit,
on two calls in a row, the second with a smaller block; with per-genome LDS
rows, with spill rows, and with so many references that the kernel runs
fewer workgroups than spill rows would need; captured in a graph on a handle
that never ran an eager call; sizes beyond the array refused."""
import numpy as np
import pytest

import kssdtest as kt

pytestmark = pytest.mark.gpu
OBJ_BITS = 20
ROW = 56


def _samples(rng, n, pool, base, mean, empty=()):
    """n sorted samples of about `mean` contexts drawn from `pool` (with
    replacement), 30% of the objects changed in one lane, 20% of the entries
    with a second object; samples listed in `empty` left empty.  Returns the
    concatenated keys and the n + 1 offsets."""
    k = rng.integers(0, 2 * mean + 1, size=n)
    k[list(empty)] = 0
    owner = np.repeat(np.arange(n), k)
    ci = rng.integers(0, len(pool), size=len(owner))
    obj = base[ci] ^ (rng.integers(1, 4, size=len(ci)).astype(np.uint64) * (rng.random(len(ci)) < 0.3))
    keys = (pool[ci] << np.uint64(OBJ_BITS)) | obj
    conf = rng.random(len(ci)) < 0.2
    keys = np.concatenate([keys, (pool[ci[conf]] << np.uint64(OBJ_BITS)) | (obj[conf] ^ np.uint64(12))])
    owner = np.concatenate([owner, owner[conf]])
    order = np.lexsort((keys, owner))
    keys, owner = keys[order], owner[order]
    keep = np.ones(len(keys), bool)
    keep[1:] = (keys[1:] != keys[:-1]) | (owner[1:] != owner[:-1])
    keys, owner = keys[keep], owner[keep]
    idx = np.zeros(n + 1, np.uint64)
    idx[1:] = np.cumsum(np.bincount(owner, minlength=n))
    return keys, idx


# (references, queries, context pool, mean reference entries, mean query
# entries, LDS slots): per-genome LDS rows; an LDS table of 256 slots that
# most queries overflow (spill rows); 2^17 references, whose spill rows
# (2 or 4 MiB each) fit the array only for a few dozen workgroups
CASES = {
    "direct": (40, 90, 3000, 200, 800, 0),
    "spill": (4400, 90, 2500, 40, 160, 256),
    "many_refs": (1 << 17, 300, 20000, 3, 40, 256),
}


@pytest.fixture(scope="module")
def cases():
    cache = {}

    def get(name):
        if name not in cache:
            n_ref, n_qry, pool_n, rmean, qmean, slots = CASES[name]
            rng = np.random.default_rng({"direct": 31, "spill": 32, "many_refs": 33}[name])
            pool = rng.integers(1, 1 << 43, size=pool_n, dtype=np.uint64)
            base = rng.integers(0, 1 << OBJ_BITS, size=pool_n, dtype=np.uint64)
            rk, ri = _samples(rng, n_ref, pool, base, rmean)
            qk, qi = _samples(rng, n_qry, pool, base, qmean, empty=(2,))
            cache[name] = (rk, ri, qk, qi, kt.build_index_np(rk, ri, OBJ_BITS), slots)
        return cache[name]

    return get


class _Block:
    """A block's device inputs and outputs."""

    def __init__(self, rt, qk, qi, rctx, cap):
        self.rt = rt
        self.nq, self.total = len(qi) - 1, int(qi[-1])
        self.bufs = [rt.upload(qk if len(qk) else np.zeros(1, np.uint64)), rt.upload(qi),
                     rt.upload(np.diff(qi).astype(np.uint32)), rt.upload(rctx), rt.malloc(ROW * cap), rt.malloc(4),
                     rt.malloc(4 * max(self.nq, 1)), rt.malloc(4 * max(self.nq, 1))]
        self.cap = cap

    def args(self):
        d_k, d_i, d_qc, d_rc, d_rows, d_n, d_f, d_c = self.bufs
        return (d_k, d_i, self.nq, self.total, d_qc, d_rc, d_rows, d_n, d_f, d_c)

    def result(self, stream):
        rt = self.rt
        d_k, d_i, d_qc, d_rc, d_rows, d_n, d_f, d_c = self.bufs
        rt.sync(stream)
        n = int(rt.download(d_n, np.uint32, 1)[0])
        assert n <= self.cap
        rows = rt.download(d_rows, np.uint8, ROW * n).view(kt.ROW_DTYPE).copy()
        first = rt.download(d_f, np.uint32, self.nq)
        count = rt.download(d_c, np.uint32, self.nq)
        for q in range(self.nq):  # each query's rows are where its first/count say
            assert np.all(rows["query"][first[q]:first[q] + count[q]] == q)
        return np.sort(rows, order=["query", "ref"]), count

    def free(self):
        for b in self.bufs:
            self.rt.free(b)


def _select(hip, h, blk, flags, opts, stream, noalloc):
    d_k, d_i, nq, total, d_qc, d_rc, d_rows, d_n, d_f, d_c = blk.args()
    fn = hip.select_block_dev_noalloc if noalloc else hip.select_block_dev
    fn(h, d_k, d_i, nq, total, flags, d_qc, d_rc, opts, d_rows, blk.cap, d_n, d_f, d_c, stream)


def _opts(unique_rank):
    return kt.SelectOpts(1, 0.0, 0.0, int(unique_rank), kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0))


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("flags,unique_rank", [(0, False), (6, True)])
def test_noalloc_matches_select_block_dev(hip, cases, case, flags, unique_rank):
    rt = kt.Hip(0)
    rk, ri, qk, qi, rec, slots = cases(case)
    R, Q = len(ri) - 1, len(qi) - 1
    rctx = np.diff(ri).astype(np.uint32)
    opts = _opts(unique_rank)
    small = Q // 3  # the second call: the first third of the block
    big_blk = _Block(rt, qk, qi, rctx, 1 << 21)
    small_blk = _Block(rt, qk[: int(qi[small])], qi[: small + 1], rctx, 1 << 21)
    h = hip.open_index(rec, R, OBJ_BITS)
    s = rt.stream()
    try:
        hip.set_tiling(h, mode=0, query_slots=slots)
        want = []
        for blk in (big_blk, small_blk):
            _select(hip, h, blk, flags, opts, s, noalloc=False)
            want.append(blk.result(s))
        for blk, (w_rows, w_count) in zip((big_blk, small_blk), want):
            _select(hip, h, blk, flags, opts, s, noalloc=True)
            rows, count = blk.result(s)
            assert len(w_rows) > 0
            assert np.array_equal(count, w_count), (case, blk.nq)
            assert rows.tobytes() == w_rows.tobytes(), (case, blk.nq)
    finally:
        big_blk.free()
        small_blk.free()
        hip.close_index(h)


def test_noalloc_graph_capture_on_fresh_handle(hip, cases):
    """The no-allocation call can be captured before the handle ever ran a
    call (the grow-only entry point refuses that); every replay equals the
    grow-only entry point's eager rows."""
    rt = kt.Hip(0)
    rk, ri, qk, qi, rec, slots = cases("spill")
    R = len(ri) - 1
    rctx = np.diff(ri).astype(np.uint32)
    opts = _opts(False)
    blk = _Block(rt, qk, qi, rctx, 1 << 20)
    h = hip.open_index(rec, R, OBJ_BITS)
    s = rt.stream()
    try:
        hip.set_tiling(h, query_slots=slots)
        g, ex = rt.capture(s, lambda: _select(hip, h, blk, 2, opts, s, noalloc=True))
        try:
            got = []
            for rep in range(3):
                rt.check(rt.lib.hipMemsetAsync(blk.bufs[4], 0x21 + rep, ROW * blk.cap, s))
                rt.graph_launch(ex, s)
                got.append(blk.result(s))
        finally:
            rt.graph_destroy(g, ex)
        _select(hip, h, blk, 2, opts, s, noalloc=False)
        w_rows, w_count = blk.result(s)
        assert len(w_rows) > 0
        for rows, count in got:
            assert rows.tobytes() == w_rows.tobytes() and np.array_equal(count, w_count)
    finally:
        blk.free()
        hip.close_index(h)


def test_noalloc_refuses_blocks_beyond_the_array(hip, cases):
    """More queries than the array's work order holds: KSSDHIP_EINVAL with a
    message naming the limit, before anything is launched."""
    rt = kt.Hip(0)
    rk, ri, qk, qi, rec, _ = cases("direct")
    R = len(ri) - 1
    blk = _Block(rt, qk, qi, np.diff(ri).astype(np.uint32), 16)
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        d_k, d_i, nq, total, d_qc, d_rc, d_rows, d_n, d_f, d_c = blk.args()
        with pytest.raises(RuntimeError, match="at most"):
            hip.select_block_dev_noalloc(h, d_k, d_i, kt.NOALLOC_MAX_QUERIES + 1, total, 0, d_qc, d_rc, _opts(False),
                                         d_rows, blk.cap, d_n, d_f, d_c, None)
    finally:
        blk.free()
        hip.close_index(h)
