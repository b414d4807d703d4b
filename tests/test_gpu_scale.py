"""GPU parity on down-scaled benchmark workloads (tests/golden/scale: sketches
made by the reference from the C2 and C3 recipes) and on blocks of more than
1024 queries: every counting kernel and tiling against the oracle, and the
device-side pass gate, eager and graph-captured, against the oracle's
restatement of the reference's gate."""
import numpy as np
import pytest

import kssdtest as kt

pytestmark = pytest.mark.gpu
SCALE = kt.GOLDEN / "scale"
OBJ_BITS = 20

TILINGS = [  # (name, set_tiling arguments)
    ("lookup", dict(mode=1)),
    ("partitioned_c5696_g4", dict(mode=2, chunk_records=5696, block_threads=1024, group_partitions=4)),
    ("partitioned_c256_g1", dict(mode=2, chunk_records=256, block_threads=256, group_partitions=1)),
    ("query_major", dict(mode=3)),
    ("query_major_slots256", dict(mode=3, query_slots=256, query_threads=256)),
    ("auto", dict(mode=0)),
]


def _load(name):
    ref = kt.read_sketch_dir(SCALE / name / "ref")
    qry = kt.read_sketch_dir(SCALE / name / "qry")
    return ref, qry, kt.build_index_np(ref.keys, ref.index, ref.obj_bits)


@pytest.mark.parametrize("case", ["c2", "c3"])
@pytest.mark.parametrize("tiling", [t[0] for t in TILINGS])
@pytest.mark.parametrize("flags", [0, 2, 6])
def test_scale_counts_match_oracle(hip, oracle, case, tiling, flags):
    ref, qry, rec = _load(case)
    R = ref.n
    h = hip.open_index(rec, R, ref.obj_bits)
    try:
        hip.set_tiling(h, **dict(TILINGS)[tiling])
        got = hip.count_block(h, R, qry.keys, qry.index, flags=flags)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, R, ref.obj_bits, qry.keys, qry.index, flags=flags)
    for name, g, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), got, want):
        assert np.array_equal(g, w), (case, tiling, flags, name)


@pytest.mark.parametrize("graph", [False, True])
def test_scale_c3_select_matches_oracle_gate(hip, oracle, graph):
    """The C3 slice through kssdhip_select_block_dev with the reference's
    default gate and a looser one; with `graph` the call is captured once
    and replayed, with eager work between the replays."""
    rt = kt.Hip(0)
    ref, qry, rec = _load("c3")
    R, Q = ref.n, qry.n
    qctx = np.diff(qry.index).astype(np.uint32)
    rctx = np.diff(ref.index).astype(np.uint32)
    ctx, obj, _, _ = oracle.count_block(rec, R, OBJ_BITS, qry.keys, qry.index)
    h = hip.open_index(rec, R, OBJ_BITS)
    s = rt.stream()
    cap = 1 << 16
    bufs = [rt.upload(qry.keys), rt.upload(qry.index), rt.upload(qctx), rt.upload(rctx), rt.malloc(56 * cap),
            rt.malloc(4), rt.malloc(4 * Q), rt.malloc(4 * Q)]
    d_k, d_i, d_qc, d_rc, d_rows, d_n, d_f, d_c = bufs
    try:
        for ctxcut, afcut, anicut in ((3, 0.5, 0.95), (1, 0.1, 0.9)):
            opts = kt.SelectOpts(ctxcut, afcut, anicut, 0, kt.EpilogueOpts(3, 1, 0, 1, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0))

            def call():
                hip.select_block_dev(h, d_k, d_i, Q, int(qry.index[-1]), 0, d_qc, d_rc, opts, d_rows, cap, d_n, d_f,
                                     d_c, s)

            call()  # sizes the workspace
            reps = 1
            if graph:
                g, ex = rt.capture(s, call)
                reps = 3
            for rep in range(reps):
                if graph:
                    rt.check(rt.lib.hipMemsetAsync(d_rows, 0x40 + rep, 56 * cap, s))
                    hip.count_block(h, R, qry.keys[: int(qry.index[2])], qry.index[:3])  # eager work between
                    rt.graph_launch(ex, s)
                rt.sync(s)
                n = int(rt.download(d_n, np.uint32, 1)[0])
                rows = np.sort(rt.download(d_rows, np.uint8, 56 * n).view(kt.ROW_DTYPE), order=["query", "ref"])
                ranked = oracle.index_rank(ctx, obj, ctx, obj, rctx, qctx, ctxcut=ctxcut, afcut=afcut, anicut=anicut)
                want_q = np.concatenate([np.full(len(p[0]), q, np.uint32) for q, p in enumerate(ranked)])
                want_r = np.concatenate([p[0] for p in ranked])
                assert len(want_r) > 0
                assert np.array_equal(rows["query"], want_q) and np.array_equal(rows["ref"], want_r), (ctxcut, rep)
                assert np.array_equal(rows["rank_ani"], np.concatenate([p[1] for p in ranked]))
                qq, rr = rows["query"].astype(np.int64), rows["ref"].astype(np.int64)
                assert np.array_equal(rows["num_ctx"], ctx[qq, rr, 0]) and np.array_equal(rows["diff_obj"],
                                                                                         obj[qq, rr, 0])
                counts = rt.download(d_c, np.uint32, Q)
                assert np.array_equal(counts, np.array([len(p[0]) for p in ranked], np.uint32))
            if graph:
                rt.graph_destroy(g, ex)
    finally:
        for b in bufs:
            rt.free(b)
        hip.close_index(h)


def _many_small(seed, n_ref, n_qry):
    rng = np.random.default_rng(seed)
    pool = rng.integers(1, 1 << 40, size=4000, dtype=np.uint64)
    base = rng.integers(0, 1 << OBJ_BITS, size=4000, dtype=np.uint64)

    def sample(mean):
        k = int(rng.integers(0, 2 * mean + 1))
        ci = rng.choice(len(pool), size=k, replace=False)
        obj = base[ci] ^ (rng.integers(0, 4, size=k).astype(np.uint64) * (rng.random(k) < 0.3))
        keys = (pool[ci] << np.uint64(OBJ_BITS)) | obj
        conf = rng.random(k) < 0.15
        extra = (pool[ci[conf]] << np.uint64(OBJ_BITS)) | (obj[conf] ^ np.uint64(12))
        return np.unique(np.concatenate([keys, extra]))

    refs = [sample(120) for _ in range(n_ref)]
    qrys = [sample(6) if i % 17 else np.zeros(0, np.uint64) for i in range(n_qry)]
    return refs, qrys


def _block(samples):
    keys = np.concatenate(samples) if samples else np.zeros(0, np.uint64)
    idx = np.zeros(len(samples) + 1, np.uint64)
    idx[1:] = np.cumsum([len(s) for s in samples])
    return keys, idx


@pytest.mark.parametrize("nq", [1100, 2100])
@pytest.mark.parametrize("group", [1, 4])
@pytest.mark.parametrize("flags", [0, 2, 4, 6])
def test_partitioned_join_query_slices(hip, oracle, nq, group, flags):
    """Blocks of more than 1024 queries take the partitioned join's query
    slices (grid.y > 0): slice offsets, unique slots and empty queries."""
    refs, qrys = _many_small(90 + nq + group, 40, nq)
    rk, ri = _block(refs)
    qk, qi = _block(qrys)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        hip.set_tiling(h, mode=2, chunk_records=512, block_threads=256, group_partitions=group)
        got = hip.count_block(h, len(refs), qk, qi, flags=flags)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, qk, qi, flags=flags)
    for name, g, w in zip(("ctx", "obj", "unique_ctx", "unique_obj"), got, want):
        assert np.array_equal(g, w), (nq, group, flags, name)


@pytest.mark.parametrize("group", [1, 4])
def test_partitioned_join_lower_with_offset(hip, oracle, group):
    """LOWER over a 1100-query block that starts at gid 300 of a 1500-genome
    self comparison."""
    refs, _ = _many_small(7 + group, 1500, 0)
    rk, ri = _block(refs)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    off, end = 300, 1400
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        hip.set_tiling(h, mode=2, chunk_records=512, block_threads=256, group_partitions=group)
        got = hip.count_block(h, len(refs), rk[ri[off]:ri[end]], ri[off:end + 1], qry_gid_offset=off, flags=1)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, rk[ri[off]:ri[end]], ri[off:end + 1], qry_gid_offset=off,
                              flags=1)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
