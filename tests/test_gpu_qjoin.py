"""GPU parity of the query-major join (tiling mode 3) and of the device-side
pass gate (kssdhip_select_block), through the C ABI, against the CPU oracle:
dense counts bit-exact for every flag set, with the LDS table in its direct
form (few references), its hashed form, and with queries that touch more
genomes than the table holds (spill rows); the gate's passing pairs equal the
oracle's restatement of command_ani.h:585-606 pair for pair."""
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


def fast_samples(rng, n, pool, base_obj, mean, conflict_p=0.2, obj_bits=OBJ_BITS):
    """Samples of about `mean` contexts drawn from `pool` (vectorised);
    objects are the pool's base object with 1-3 lanes changed for 30% of the
    entries, and a share of the contexts carries a second object."""
    out = []
    for _ in range(n):
        k = min(int(rng.integers(0, 2 * mean + 1)), len(pool))
        ci = rng.choice(len(pool), size=k, replace=False)
        obj = base_obj[ci].copy()
        noisy = rng.random(k) < 0.3
        lanes = rng.integers(0, obj_bits // 2, size=k).astype(np.uint64)
        obj[noisy] ^= rng.integers(1, 4, size=int(noisy.sum())).astype(np.uint64) << (np.uint64(2) * lanes[noisy])
        keys = (pool[ci] << np.uint64(obj_bits)) | obj
        conf = rng.random(k) < conflict_p
        extra = (pool[ci[conf]] << np.uint64(obj_bits)) | (obj[conf] ^ np.uint64(3))
        out.append(np.unique(np.concatenate([keys, extra])))
    return out


def _wide(seed, n_ref, n_qry, pool_n, mean):
    rng = np.random.default_rng(seed)
    pool = rng.integers(1, 1 << 43, size=pool_n, dtype=np.uint64)
    base = rng.integers(0, 1 << OBJ_BITS, size=pool_n, dtype=np.uint64)
    refs = fast_samples(rng, n_ref, pool, base, mean)
    qrys = fast_samples(rng, n_qry, pool, base, 4 * mean)
    qrys[1] = np.zeros(0, np.uint64)  # an empty query
    return refs, qrys


# (references, queries, context pool, mean entries, query_slots): the LDS
# table in direct form; hashed for few references (slots asked for where a
# direct row would fit); hashed without spilling; hashed with most queries
# touching more genomes than its 256 slots
SHAPES = {
    "direct": (40, 70, 3000, 200, 0),
    "small_hashed": (40, 70, 3000, 200, 256),
    "hashed": (4400, 40, 200000, 60, 0),
    "spill": (4400, 24, 2500, 40, 256),
}
SEEDS = {"direct": 11, "small_hashed": 14, "hashed": 12, "spill": 13}


@pytest.fixture(scope="module")
def wide_cases():
    cache = {}

    def get(name):
        if name not in cache:
            n_ref, n_qry, pool_n, mean, slots = SHAPES[name]
            refs, qrys = _wide(SEEDS[name], n_ref, n_qry, pool_n, mean)
            rk, ri = _block(refs)
            qk, qi = _block(qrys)
            cache[name] = (rk, ri, qk, qi, kt.build_index_np(rk, ri, OBJ_BITS), slots)
        return cache[name]

    return get


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("flags", [0, 2, 4, 6])
def test_query_major_dense_matches_oracle(hip, oracle, wide_cases, shape, flags):
    rk, ri, qk, qi, rec, slots = wide_cases(shape)
    R = len(ri) - 1
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=3, query_slots=slots, query_threads=256)
        got = hip.count_block(h, R, qk, qi, flags=flags)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, R, OBJ_BITS, qk, qi, flags=flags)
    for name, g, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), got, want):
        assert np.array_equal(g, w), (shape, flags, name)
    if shape == "spill":  # the shape must really overflow the table
        assert (got[0][:, :, 0] > 0).sum(1).max() > 256


@pytest.mark.parametrize("shape", ["direct", "spill"])
def test_query_major_lower_blocks(hip, oracle, wide_cases, shape):
    rk, ri, _, _, rec, slots = wide_cases(shape)
    R = len(ri) - 1
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=3, query_slots=slots)
        for off in (0, 7, R - 9):
            end = min(off + 9, R)
            got = hip.count_block(h, R, rk[ri[off]:ri[end]], ri[off:end + 1], qry_gid_offset=off, flags=1 | 4)
            want = oracle.count_block(rec, R, OBJ_BITS, rk[ri[off]:ri[end]], ri[off:end + 1],
                                      qry_gid_offset=off, flags=1 | 4)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), off
    finally:
        hip.close_index(h)


def test_query_major_accumulate_and_auto(hip, oracle, wide_cases):
    """ACCUMULATE adds into the outputs (twice the counts after two calls);
    a plain call overwrites garbage in the outputs; auto mode agrees."""
    rt = kt.Hip(0)
    rk, ri, qk, qi, rec, _ = wide_cases("hashed")
    R, Q = len(ri) - 1, len(qi) - 1
    h = hip.open_index(rec, R, OBJ_BITS)
    s = rt.stream()
    d_k, d_i = rt.upload(qk), rt.upload(qi)
    cells = Q * R
    d_c, d_o = rt.malloc(8 * cells), rt.malloc(8 * cells)
    try:
        want = oracle.count_block(rec, R, OBJ_BITS, qk, qi)
        hip.set_tiling(h, mode=3)
        rt.check(rt.lib.hipMemsetAsync(d_c, 0, 8 * cells, s))
        rt.check(rt.lib.hipMemsetAsync(d_o, 0, 8 * cells, s))
        for _ in range(2):
            hip.count_block_dev(h, d_k, d_i, Q, int(qi[-1]), 0, kt.ACCUMULATE, d_c, d_o, None, None, s)
        rt.sync(s)
        c = rt.download(d_c, np.uint32, 2 * cells).reshape(Q, R, 2)
        o = rt.download(d_o, np.uint32, 2 * cells).reshape(Q, R, 2)
        assert np.array_equal(c, 2 * want[0]) and np.array_equal(o, 2 * want[1])
        hip.set_tiling(h, mode=0)
        rt.check(rt.lib.hipMemsetAsync(d_c, 0x7F, 8 * cells, s))  # garbage: the kernel must overwrite
        hip.count_block_dev(h, d_k, d_i, Q, int(qi[-1]), 0, 0, d_c, d_o, None, None, s)
        rt.sync(s)
        assert np.array_equal(rt.download(d_c, np.uint32, 2 * cells).reshape(Q, R, 2), want[0])
    finally:
        for d in (d_k, d_i, d_c, d_o):
            rt.free(d)
        hip.close_index(h)


def _ctx_totals(keys, idx):
    return np.diff(idx).astype(np.uint32)


def _check_select(hip, oracle, rec, R, qk, qi, flags, unique_rank, ctxcut, afcut, anicut, slots=0, mode=3):
    qctx = _ctx_totals(qk, qi)
    rctx = np.maximum(np.bincount((rec["ctxgid"] & np.uint64((1 << 20) - 1)).astype(np.int64),
                                  minlength=R).astype(np.uint32), 0)
    rctx[::7] += 3  # totals need not equal the index's entries
    opts = kt.SelectOpts(ctxcut, afcut, anicut, int(unique_rank),
                         kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0))
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=mode, query_slots=slots)
        rows = hip.select_block(h, qk, qi, qctx, rctx, opts, flags=flags)
    finally:
        hip.close_index(h)
    ctx, obj, uctx, uobj = oracle.count_block(rec, R, OBJ_BITS, qk, qi, flags=flags)
    ranked = oracle.index_rank(ctx, obj, uctx, uobj, rctx, qctx, ctxcut=ctxcut, afcut=afcut, anicut=anicut,
                               unique_ctx=unique_rank)
    ofs = 0
    for q, (pass_ids, rank, _) in enumerate(ranked):
        mine = rows[rows["query"] == q]
        if unique_rank:
            # the device keeps pairs up to 1e-9 below the cut; the oracle's set
            # must be contained, and the extras must lie in that margin
            assert set(pass_ids.tolist()) <= set(mine["ref"].tolist()), q
            extra = ~np.isin(mine["ref"], pass_ids)
            assert np.all(mine["rank_ani"][extra] >= np.float64(np.float32(anicut)) - 1e-9)
            keep = ~extra
            mine = mine[keep]
        assert np.array_equal(mine["ref"], pass_ids), q
        r = mine["ref"].astype(np.int64)
        assert np.array_equal(mine["num_ctx"], ctx[q, r, 0]) and np.array_equal(mine["num_mut2_ctx"], ctx[q, r, 1])
        assert np.array_equal(mine["diff_obj"], obj[q, r, 0])
        assert np.array_equal(mine["diff_obj_section"], obj[q, r, 1])
        if flags & kt.UNIQUE:
            assert np.array_equal(mine["u_num_ctx"], uctx[q, r, 0])
            assert np.array_equal(mine["u_diff_obj_section"], uobj[q, r, 1])
        if unique_rank:
            assert np.allclose(mine["rank_ani"], rank, rtol=1e-12, atol=0)
        else:
            assert np.array_equal(mine["rank_ani"], rank), q
        ofs += len(mine)
    # values: the FP64 CtxMoE value of each kept pair
    if len(rows):
        qsel, rsel = rows["query"].astype(np.int64), rows["ref"].astype(np.int64)
        out = np.zeros((len(qi) - 1) * R, np.float64)
        oracle.lib.orc_matrix_values(kt._p(ctx), kt._p(obj), len(qi) - 1, R, kt._p(qctx), kt._p(rctx), 3, 1,
                                     0.5, 1.0, 8, OBJ_BITS, 44, kt._p(out), 4)
        want = out.reshape(-1, R)[qsel, rsel]
        rel = np.abs(rows["value"] - want) / np.maximum(np.abs(want), 1e-300)
        assert float(rel.max()) <= 1e-12
    return rows


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("flags,unique_rank", [(0, False), (4, False), (2, True), (6, True), (2, False)])
@pytest.mark.parametrize("mode", [3, 0])
def test_select_matches_oracle_gate(hip, oracle, wide_cases, shape, flags, unique_rank, mode):
    """mode 3: the query-major join in select mode; mode 0: these blocks have
    fewer queries than compute units, so they are counted into dense rows
    by the lookup kernel and gated row by row."""
    rk, ri, qk, qi, rec, slots = wide_cases(shape)
    R = len(ri) - 1
    rows = _check_select(hip, oracle, rec, R, qk, qi, flags, unique_rank, 1, 0.0, 0.0, slots, mode)
    assert len(rows) > 0


@pytest.mark.parametrize("ctxcut,afcut,anicut", [(3, 0.0, 0.0), (1, 0.05, 0.0), (2, 0.0, 0.93), (1, 0.02, 0.97)])
def test_select_gate_cuts(hip, oracle, wide_cases, ctxcut, afcut, anicut):
    rk, ri, qk, qi, rec, slots = wide_cases("direct")
    for mode in (3, 0):
        _check_select(hip, oracle, rec, len(ri) - 1, qk, qi, 2, True, ctxcut, afcut, anicut, mode=mode)
        _check_select(hip, oracle, rec, len(ri) - 1, qk, qi, 0, False, ctxcut, afcut, anicut, mode=mode)


def test_select_capacity_and_empty(hip, oracle, wide_cases):
    """A buffer too small reports the count needed (KSSDHIP_ENOSPC) and the
    binding retries; an all-empty block passes nothing; ctxcut 0 is refused."""
    rk, ri, qk, qi, rec, _ = wide_cases("direct")
    R = len(ri) - 1
    qctx, rctx = _ctx_totals(qk, qi), np.diff(ri).astype(np.uint32)
    opts = kt.SelectOpts(1, 0.0, 0.0, 0, kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0))
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        full = hip.select_block(h, qk, qi, qctx, rctx, opts)
        small = hip.select_block(h, qk, qi, qctx, rctx, opts, cap=3)
        assert len(full) > 3 and np.array_equal(full, small)
        none = hip.select_block(h, np.zeros(0, np.uint64), np.zeros(4, np.uint64), np.zeros(3, np.uint32), rctx,
                                opts)
        assert len(none) == 0
        bad = kt.SelectOpts(0, 0.0, 0.0, 0, opts.value)
        with pytest.raises(RuntimeError):
            hip.select_block(h, qk, qi, qctx, rctx, bad)
    finally:
        hip.close_index(h)


@pytest.mark.parametrize("mode", [3, 0])
def test_select_graph_capture(hip, oracle, wide_cases, mode):
    """Capturing the select before its workspace exists is refused (no
    allocation inside a capture); after one eager call every replay of the
    captured call equals the eager rows, with eager work in between."""
    rt = kt.Hip(0)
    rk, ri, qk, qi, rec, _ = wide_cases("spill")
    R, Q = len(ri) - 1, len(qi) - 1
    qctx, rctx = _ctx_totals(qk, qi), np.diff(ri).astype(np.uint32)
    opts = kt.SelectOpts(1, 0.0, 0.0, 0, kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0))
    h = hip.open_index(rec, R, OBJ_BITS)
    s = rt.stream()
    cap = Q * R
    bufs = [rt.upload(qk), rt.upload(qi), rt.upload(qctx), rt.upload(rctx), rt.malloc(56 * cap), rt.malloc(4),
            rt.malloc(4 * Q), rt.malloc(4 * Q)]
    d_k, d_i, d_qc, d_rc, d_rows, d_n, d_first, d_cnt = bufs
    try:
        hip.set_tiling(h, mode=mode, query_slots=256)

        def call():
            hip.select_block_dev(h, d_k, d_i, Q, int(qi[-1]), 0, d_qc, d_rc, opts, d_rows, cap, d_n, d_first, d_cnt,
                                 s)

        def rows_now():
            rt.sync(s)
            n = int(rt.download(d_n, np.uint32, 1)[0])
            r = rt.download(d_rows, np.uint8, 56 * n).view(kt.ROW_DTYPE).copy()
            return np.sort(r, order=["query", "ref"])

        with pytest.raises(RuntimeError, match="captured"):
            rt.capture(s, call)
        call()
        eager = rows_now()
        assert len(eager) > 0
        g, ex = rt.capture(s, call)
        try:
            for rep in range(3):
                rt.check(rt.lib.hipMemsetAsync(d_rows, 0x33 + rep, 56 * cap, s))
                hip.count_block(h, R, qk[: int(qi[3])], qi[:4])  # eager work on the handle between replays
                rt.graph_launch(ex, s)
                assert np.array_equal(rows_now(), eager), rep
        finally:
            rt.graph_destroy(g, ex)
    finally:
        for b in bufs:
            rt.free(b)
        hip.close_index(h)


@pytest.mark.parametrize("shape", ["direct", "hashed"])
def test_matched_records(hip, wide_cases, shape):
    """kssdhip_index_matched_records equals the count of index records whose
    context occurs in some query of the block (numpy)."""
    rt = kt.Hip(0)
    rk, ri, qk, qi, rec, _ = wide_cases(shape)
    R, Q = len(ri) - 1, len(qi) - 1
    want = int(np.isin(rec["ctxgid"] >> np.uint64(20), np.unique(qk >> np.uint64(OBJ_BITS))).sum())
    h = hip.open_index(rec, R, OBJ_BITS)
    d_k, d_i = rt.upload(qk), rt.upload(qi)
    try:
        assert hip.matched_records(h, d_k, d_i, Q, int(qi[-1])) == want
        assert hip.matched_records(h, d_k, d_i, 0, 0) == 0
    finally:
        rt.free(d_k)
        rt.free(d_i)
        hip.close_index(h)


@pytest.mark.parametrize("select", [False, True])
def test_smaller_block_after_larger_on_one_handle(hip, oracle, wide_cases, select):
    """The handle's workspace keeps its spill rows where they were when a
    later block has fewer queries (a shorter work order): a 96-query block
    (the spill shape's queries four times over), then its first 24 queries,
    then 7, each against the oracle."""
    rk, ri, qk, qi, rec, slots = wide_cases("spill")
    R, Q = len(ri) - 1, len(qi) - 1
    sizes = [np.diff(qi)] * 4
    big_k = np.concatenate([qk] * 4)
    big_i = np.zeros(4 * Q + 1, np.uint64)
    big_i[1:] = np.cumsum(np.concatenate(sizes))
    rctx = np.diff(ri).astype(np.uint32)
    opts = kt.SelectOpts(1, 0.0, 0.0, 0, kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0))
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=3, query_slots=slots, query_threads=256)
        for nq in (4 * Q, Q, 7):
            bk, bi = big_k[: int(big_i[nq])], big_i[: nq + 1]
            ctx, obj, _, _ = oracle.count_block(rec, R, OBJ_BITS, bk, bi)
            if select:
                qctx = np.diff(bi).astype(np.uint32)
                rows = hip.select_block(h, bk, bi, qctx, rctx, opts)
                ranked = oracle.index_rank(ctx, obj, ctx, obj, rctx, qctx, ctxcut=1, afcut=0.0, anicut=0.0)
                q = np.concatenate([np.full(len(p[0]), i, np.uint32) for i, p in enumerate(ranked)])
                r = np.concatenate([p[0] for p in ranked]).astype(np.int64)
                assert np.array_equal(rows["query"], q) and np.array_equal(rows["ref"], r), nq
                assert np.array_equal(rows["num_ctx"], ctx[q, r, 0]) and np.array_equal(rows["diff_obj"], obj[q, r, 0])
            else:
                got = hip.count_block(h, R, bk, bi)
                assert np.array_equal(got[0], ctx) and np.array_equal(got[1], obj), nq
    finally:
        hip.close_index(h)


# ---------------------------------------------------------------- record table
# Home line of a context in the index's record table: the top home_bits bits
# of a bijection of its ctx_bits-bit value (an odd multiplier modulo 2^w,
# then an xor-shift).  Inverted here to place chosen contexts in one home
# line.
_M1 = 0x9E3779B97F4A7C15


def _xs_inv(y, s, w):
    x = y
    for _ in range(w // s + 1):
        x = y ^ (x >> s)
    return x


def _perm_inv(p, w):
    m, s = (1 << w) - 1, (w + 1) // 2
    x = _xs_inv(p, s, w)
    return (x * pow(_M1, -1, 1 << w)) & m


def _perm(x, w):
    m, s = (1 << w) - 1, (w + 1) // 2
    x = (x * _M1) & m
    return x ^ (x >> s)


def test_perm_inverse():
    rng = np.random.default_rng(5)
    for w in (32, 36, 44):
        for x in rng.integers(0, 1 << w, size=200, dtype=np.uint64):
            assert _perm_inv(_perm(int(x), w), w) == int(x)


def _crowded_case(seed, obj_bits, n_ref=48):
    """References whose contexts crowd a few home lines: one home with ~40
    contexts (runs of 1-12 records and long runs of 13-40, so the home spills
    over a chain of overflow lines), one with exactly 16 slots (one full line,
    no continuation), one with 17 (a continuation for one record), plus
    scattered contexts; queries hold most of them (objects kept or changed)
    and absent contexts of the crowded homes (looked up through the whole
    chain), some with a second object."""
    rng = np.random.default_rng(seed)
    w = 64 - obj_bits if obj_bits >= 20 else 44
    bb = 21
    rb = w - bb
    omask = (1 << obj_bits) - 1

    def ctx_in(home, count):
        rems = rng.choice(1 << min(rb, 20), size=count, replace=False)
        return [_perm_inv((home << rb) | int(r), w) for r in rems]

    runs = {}  # context -> run length (genomes holding it)
    for c in ctx_in(12345, 40):
        runs[c] = int(rng.integers(13, 41)) if rng.random() < 0.2 else int(rng.integers(1, 13))
    for c, ln in zip(ctx_in(777, 4), (4, 4, 4, 4)):  # 16 slots: one line exactly
        runs[c] = ln
    for c, ln in zip(ctx_in(4242, 3), (6, 6, 5)):  # 17 slots: a continuation
        runs[c] = ln
    for c in rng.integers(1, 1 << w, size=300, dtype=np.uint64):
        runs[int(c)] = int(rng.integers(1, 6))
    per_ref = [[] for _ in range(n_ref)]
    base = {}
    for c, ln in runs.items():
        base[c] = int(rng.integers(0, omask + 1))
        for g in rng.choice(n_ref, size=min(ln, n_ref), replace=False):
            o = base[c] ^ (int(rng.integers(0, 4)) << (2 * int(rng.integers(0, obj_bits // 2))))
            per_ref[g].append((c << obj_bits) | (o & omask))
            if rng.random() < 0.1:  # a conflict: a second object in this genome
                per_ref[g].append((c << obj_bits) | ((o ^ 3) & omask))
    refs = [np.unique(np.array(k, dtype=np.uint64)) for k in per_ref]
    absent = ctx_in(12345, 30) + ctx_in(4242, 5)
    absent = [c for c in absent if c not in runs]
    ctxs = list(runs)
    qrys = []
    for _ in range(300):
        pick = rng.choice(len(ctxs), size=int(rng.integers(0, 60)), replace=False)
        ks = []
        for i in pick:
            c = ctxs[i]
            o = base[c] if rng.random() < 0.5 else int(rng.integers(0, omask + 1))
            ks.append((c << obj_bits) | o)
            if rng.random() < 0.1:
                ks.append((c << obj_bits) | (o ^ 1))
        for c in rng.choice(absent, size=min(5, len(absent)), replace=False):
            ks.append((int(c) << obj_bits) | int(rng.integers(0, omask + 1)))
        qrys.append(np.unique(np.array(ks, dtype=np.uint64)))
    qrys[3] = np.zeros(0, np.uint64)
    return refs, qrys


@pytest.mark.parametrize("obj_bits", [8, 16, 20, 28, 32])
def test_record_table_overflow_and_long_contexts(hip, oracle, obj_bits):
    """Home lines that overflow into chains of overflow lines, a home of
    exactly one full line, long contexts (records walked in the sorted
    index), absent contexts that walk a whole chain: query-major counts
    equal the oracle's for every flag set, also below the diagonal, and the
    table really has overflow lines.  Object widths below 20 keep 44-bit
    contexts; below 12 the slot test compares 64-bit values."""
    refs, qrys = _crowded_case(31 + obj_bits, obj_bits)
    rk, ri = _block(refs)
    qk, qi = _block(qrys)
    R = len(refs)
    rec = kt.build_index_np(rk, ri, obj_bits)
    h = hip.open_index(rec, R, obj_bits)
    try:
        lines, ovf, hb, cb = hip.table_info(h)
        assert (hb, cb) == (21, 64 - obj_bits if obj_bits >= 20 else 44) and ovf >= 3 and lines == (1 << 21) + ovf
        hip.set_tiling(h, mode=3, query_threads=256)
        for flags in (0, 2, 4, 6):
            got = hip.count_block(h, R, qk, qi, flags=flags)
            want = oracle.count_block(rec, R, obj_bits, qk, qi, flags=flags)
            for name, g, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), got, want):
                assert np.array_equal(g, w), (obj_bits, flags, name)
        for off in (0, 20, R - 9):
            got = hip.count_block(h, R, rk[ri[off]:ri[off + 9]], ri[off:off + 10], qry_gid_offset=off, flags=1)
            want = oracle.count_block(rec, R, obj_bits, rk[ri[off]:ri[off + 9]], ri[off:off + 10],
                                      qry_gid_offset=off, flags=1)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), off
    finally:
        hip.close_index(h)


@pytest.mark.parametrize("flags", [0, 1, 2, 4, 6])
def test_query_major_single_record_contexts(hip, oracle, flags):
    """Sketches without second objects over a large context pool, so that most
    contexts an index holds have one record and most waves find only such
    contexts (each key's lane then counts its own record); a share of the
    pool shared more widely keeps some waves on the wave-wide walk.  LOWER
    calls count the block against itself below each query's gid."""
    rng = np.random.default_rng(123)
    pool = rng.integers(1, 1 << 43, size=4000000, dtype=np.uint64)  # 4M contexts: about 7 of 8 waves take the per-lane path
    base = rng.integers(0, 1 << OBJ_BITS, size=len(pool), dtype=np.uint64)
    refs = fast_samples(rng, 300, pool, base, 900, conflict_p=0.0)
    hot = pool[:300]
    refs = [np.unique(np.concatenate([r, (hot[rng.integers(0, 300, 20)] << np.uint64(OBJ_BITS)) |
                                      base[:20]])) for r in refs]
    qrys = refs[:40] if flags & 1 else fast_samples(rng, 600, pool, base, 900, conflict_p=0.0)
    rk, ri = _block(refs)
    qk, qi = _block(qrys)
    R = len(refs)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=3)
        got = hip.count_block(h, R, qk, qi, flags=flags)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, R, OBJ_BITS, qk, qi, flags=flags)
    for name, g, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), got, want):
        assert np.array_equal(g, w), (flags, name)
    assert (got[0][:, :, 0] > 0).sum() > 0


@pytest.mark.parametrize("mode_items", [False, True])
@pytest.mark.parametrize("flags", [0, 2, 4, 6])
def test_query_major_groups_longer_than_a_wave(hip, oracle, flags, mode_items):
    """Query-side context groups of 1 to 300 keys (one context, many
    objects): groups that fill whole 64-key waves, run past a wave's last
    lane and, with key-range items (few queries), past item ends; the index
    holds the group's contexts with objects near the group's, so the
    minimum over the group matters.  Counts against the oracle."""
    rng = np.random.default_rng(91)
    pool = rng.integers(1, 1 << 43, size=6000, dtype=np.uint64)
    base = rng.integers(0, 1 << OBJ_BITS, size=len(pool), dtype=np.uint64)
    refs = fast_samples(rng, 120, pool, base, 700, conflict_p=0.3)
    big = rng.choice(len(pool), size=24, replace=False)
    qrys = []
    nq = 6 if mode_items else 1100  # 1,100 queries: more than twice the resident workgroups, whole queries
    for q in range(nq):
        keys = [fast_samples(rng, 1, pool, base, 400, conflict_p=0.4)[0]]
        for c in rng.choice(big, size=3, replace=False):
            n = int(rng.choice([2, 63, 64, 65, 130, 300]))
            objs = np.unique((base[c] ^ rng.integers(0, 1 << OBJ_BITS, size=n, dtype=np.uint64)
                              & np.uint64((1 << OBJ_BITS) - 1)))
            keys.append((pool[c] << np.uint64(OBJ_BITS)) | objs)
        qrys.append(np.unique(np.concatenate(keys)))
    rk, ri = _block(refs)
    qk, qi = _block(qrys)
    R = len(refs)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=3)
        got = hip.count_block(h, R, qk, qi, flags=flags)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, R, OBJ_BITS, qk, qi, flags=flags)
    for name, g, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), got, want):
        assert np.array_equal(g, w), (flags, mode_items, name)


@pytest.mark.parametrize("qconf", [0.5, 1.0])
@pytest.mark.parametrize("flags", [0, 2, 4, 6, 8 | 2])
def test_query_major_key_range_items(hip, oracle, flags, qconf):
    """Blocks of few large queries are cut into key-range items (rounds of
    halving sizes, then 1,024 keys) that may span queries and add into the
    rows: 6 queries of about 2,000-9,000 keys, one empty, whose context
    groups (a second object for a share qconf of the contexts; every context
    at 1.0) straddle the item boundaries, so that a group headed before an
    item's end reads its keys past it; against the oracle for every flag
    set; with ACCUMULATE the rows add to what they held."""
    rng = np.random.default_rng(77)
    pool = rng.integers(1, 1 << 43, size=40000, dtype=np.uint64)
    base = rng.integers(0, 1 << OBJ_BITS, size=len(pool), dtype=np.uint64)
    refs = fast_samples(rng, 300, pool, base, 600, conflict_p=0.3)
    qrys = fast_samples(rng, 6, pool, base, 2500, conflict_p=qconf)
    qrys[2] = np.zeros(0, np.uint64)
    rk, ri = _block(refs)
    qk, qi = _block(qrys)
    R = len(refs)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    assert max(np.diff(qi)) > 4 * 1024
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        hip.set_tiling(h, mode=3)
        want = oracle.count_block(rec, R, OBJ_BITS, qk, qi, flags=flags & ~8)
        got = hip.count_block(h, R, qk, qi, flags=flags & ~8)
        for name, g, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), got, want):
            assert np.array_equal(g, w), (flags, name)
        if flags & 8:  # twice into the same device rows
            rt = kt.Hip(0)
            s = rt.stream()
            d_k, d_i = rt.upload(qk), rt.upload(qi)
            cells = (len(qi) - 1) * R
            bufs = [rt.malloc(8 * cells) for _ in range(4)]
            try:
                for b in bufs:
                    rt.check(rt.lib.hipMemsetAsync(b, 0, 8 * cells, s))
                for _ in range(2):
                    hip.count_block_dev(h, d_k, d_i, len(qi) - 1, int(qi[-1]), 0, flags, *bufs, stream=s)
                rt.sync(s)
                for name, b, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), bufs, want):
                    assert np.array_equal(rt.download(b, np.uint32, 2 * cells).reshape(w.shape), 2 * w), name
            finally:
                for b in bufs:
                    rt.free(b)
                rt.free(d_k)
                rt.free(d_i)
        for off in (0, R - 40):  # below the diagonal
            sub_k, sub_i = rk[ri[off]:ri[off + 40]], ri[off:off + 41]
            got = hip.count_block(h, R, sub_k, sub_i, qry_gid_offset=off, flags=1)
            want = oracle.count_block(rec, R, OBJ_BITS, sub_k, sub_i, qry_gid_offset=off, flags=1)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), off
    finally:
        hip.close_index(h)
