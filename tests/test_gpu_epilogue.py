"""GPU: the fused FP64 epilogue (kssdhip_epilogue_dev) against an FP64 host
evaluation of the reference formulas (the oracle) to 1e-12 relative, and
against the reference's printed self matrices; the device-resident plain
kernel against the oracle over many back-to-back launches sharing one
workspace."""
import numpy as np
import pytest

import kssdtest as kt

pytestmark = pytest.mark.gpu
GEN = kt.GOLDEN / "genomes"
REL_TOL = 1e-12


@pytest.fixture(scope="module")
def rt():
    return kt.Hip(0)


def host_value(oracle, s, values_ani, qc, rc, xny, dobj, dsec, mut2, afcut=0.5, e=1.0,
               float_af=False, obj_bits=20, ctx_bits=44):
    """make_selected_output_row for metrics 3..9 (command_ani.c:1321-1379)."""
    if qc == 0 or rc == 0:
        return 1.0 - e if values_ani else e
    if float_af:
        afq = float(np.float32(np.float32(xny) / np.float32(qc)))
        afr = float(np.float32(np.float32(xny) / np.float32(rc)))
    else:
        afq, afr = xny / qc, xny / rc
    af_pass = max(afq, afr) >= afcut
    f = kt.Features(xny, 0, dsec, mut2, dobj)
    k = ctx_bits // 2
    moe = oracle.lib.orc_lm3ways_dist(f, 8) if af_pass else e
    naive = oracle.lib.orc_naive_dist(f) if af_pass else e
    pd = oracle.lib.orc_p_dist(f, obj_bits) if af_pass else e
    mash = oracle.lib.orc_mashD(k, rc, qc, xny)
    aaf = oracle.lib.orc_aafD(k, rc, qc, xny)
    d = {3: moe, 4: naive, 5: mash, 6: aaf, 7: moe if af_pass else mash,
         8: moe if af_pass else aaf, 9: pd}[s]
    return 1.0 - d if values_ani else d


def close(a, b):
    if np.isinf(a) or np.isinf(b) or np.isnan(a) or np.isnan(b):
        return (np.isinf(a) and np.isinf(b) and np.sign(a) == np.sign(b)) or (np.isnan(a) and np.isnan(b))
    return abs(a - b) <= REL_TOL * max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("float_af", [0, 1])
def test_epilogue_matches_fp64_host(hip, oracle, rt, float_af):
    ref = kt.read_sketch_dir(GEN / "ref")
    qry = kt.read_sketch_dir(GEN / "qry")
    h = hip.open_index(ref.records, ref.n, ref.obj_bits)
    try:
        ctx, obj, _, _ = hip.count_block(h, ref.n, qry.keys, qry.index)
    finally:
        hip.close_index(h)
    qc = np.diff(qry.index).astype(np.uint32)
    rc = np.diff(ref.index).astype(np.uint32)
    qc[1] = 0  # exception cells
    d_ctx, d_obj = rt.upload(ctx), rt.upload(obj)
    d_qc, d_rc = rt.upload(qc), rt.upload(rc)
    d_out = rt.malloc(8 * qry.n * ref.n)
    try:
        for s in range(3, 10):
            for va in (0, 1):
                o = kt.EpilogueOpts(s, va, 0, float_af, 0, 0, 8, 20, 44, 0.5, 1.0)
                hip.epilogue_dev(d_ctx, d_obj, qry.n, ref.n, d_qc, d_rc, o, d_out)
                got = rt.download(d_out, np.float64, qry.n * ref.n).reshape(qry.n, ref.n)
                for q in range(qry.n):
                    for r in range(ref.n):
                        want = host_value(oracle, s, va, int(qc[q]), int(rc[r]), int(ctx[q, r, 0]),
                                          int(obj[q, r, 0]), int(obj[q, r, 1]), int(ctx[q, r, 1]),
                                          float_af=bool(float_af))
                        assert close(got[q, r], want), (s, va, q, r, got[q, r], want)
    finally:
        for p in (d_ctx, d_obj, d_qc, d_rc, d_out):
            rt.free(p)


def test_epilogue_reproduces_reference_self_matrix(hip, rt):
    """Indexed lower-triangle counts + device epilogue print the reference's
    own --format matrix cells (-s3 ANI and -s9 distance)."""
    s = kt.read_sketch_dir(GEN / "all")
    rec = kt.build_index_np(s.keys, s.index, s.obj_bits)
    h = hip.open_index(rec, s.n, s.obj_bits)
    try:
        ctx, obj, _, _ = hip.count_block(h, s.n, s.keys, s.index, flags=kt.FLAGS["LOWER"])
    finally:
        hip.close_index(h)
    cnt = np.diff(s.index).astype(np.uint32)
    d_ctx, d_obj, d_c = rt.upload(ctx), rt.upload(obj), rt.upload(cnt)
    d_out = rt.malloc(8 * s.n * s.n)
    try:
        for metric, va, golden in ((3, 1, "self_matrix_direct_s3.tsv"), (9, 0, "self_matrix_direct_s9.tsv")):
            o = kt.EpilogueOpts(metric, va, 0, 0, 1, 0, 8, 20, 44, 0.5, 1.0)
            hip.epilogue_dev(d_ctx, d_obj, s.n, s.n, d_c, d_c, o, d_out)
            got = rt.download(d_out, np.float64, s.n * s.n).reshape(s.n, s.n)
            lines = (GEN / golden).read_text().splitlines()[1:]
            for q in range(s.n):
                cells = lines[q].split("\t")[1:]
                for r in range(q):
                    assert "%f" % got[q, r] == cells[r], (golden, q, r)
    finally:
        for p in (d_ctx, d_obj, d_c, d_out):
            rt.free(p)


def test_pair_dev_back_to_back_launches(hip, oracle, rt):
    ref = kt.read_sketch_dir(GEN / "ref")
    qry = kt.read_sketch_dir(GEN / "qry")
    d_r, d_q = rt.upload(ref.keys), rt.upload(qry.keys)
    npairs = qry.n * ref.n
    d_out = rt.malloc(20 * npairs)
    d_ws = rt.malloc(16400)  # KSSDHIP_PAIR_WORKSPACE_BYTES
    stream = rt.stream()
    try:
        hip.check(hip.lib.kssdhip_pair_workspace_init_dev(d_ws, stream))
        for q in range(qry.n):
            for r in range(ref.n):
                p = q * ref.n + r
                hip.pair_dev(d_q.value + 8 * int(qry.index[q]), int(qry.index[q + 1] - qry.index[q]),
                             d_r.value + 8 * int(ref.index[r]), int(ref.index[r + 1] - ref.index[r]),
                             20, False, d_out.value + 20 * p, d_ws, stream)
        rt.sync(stream)
        got = rt.download(d_out, np.uint32, 5 * npairs).reshape(npairs, 5)
        for q in range(qry.n):
            for r in range(ref.n):
                g = got[q * ref.n + r]
                assert (g[0], g[4], g[2], g[3]) == oracle.pair(qry.sample(q), ref.sample(r), 20)
    finally:
        for p in (d_r, d_q, d_out, d_ws):
            rt.free(p)


@pytest.mark.parametrize("cut", [0.0, 0.95, 2.0])
def test_select_hits_matches_numpy(hip, rt, cut):
    """kssdhip_select_hits_dev: every cell >= the cut, each once, with its
    (query, ref) position; the count past the capacity, only `cap` stored;
    a second call on the same counter starts again from zero."""
    rng = np.random.default_rng(int(cut * 100) + 5)
    nq, nr = 37, 1013
    v = rng.random((nq, nr))
    v[3, :] = 0.95  # ties at the cut count as hits
    d_v = rt.upload(v)
    want = np.argwhere(v >= cut)
    cap = len(want) + 10
    d_hits, d_n = rt.malloc(16 * cap), rt.malloc(4)
    s = rt.stream()
    for _ in range(2):
        hip.select_hits_dev(d_v, nq, nr, cut, cap, d_hits, d_n, s)
        rt.sync(s)
        n = int(rt.download(d_n, np.uint32, 1)[0])
        assert n == len(want)
        got = rt.download(d_hits, np.uint8, 16 * n).view(kt.HIT_DTYPE)
        order = np.lexsort((got["ref"], got["query"]))
        assert np.array_equal(np.stack([got["query"][order], got["ref"][order]], 1), want)
        assert np.array_equal(got["value"][order], v[want[:, 0], want[:, 1]])
    small = max(len(want) // 3, 1)
    hip.select_hits_dev(d_v, nq, nr, cut, small, d_hits, d_n, s)
    rt.sync(s)
    assert int(rt.download(d_n, np.uint32, 1)[0]) == len(want)
    part = rt.download(d_hits, np.uint8, 16 * min(small, len(want))).view(kt.HIT_DTYPE)
    assert all(v[q, r] >= cut for q, r in zip(part["query"], part["ref"]))
    for p in (d_v, d_hits, d_n):
        rt.free(p)


def _random_tuples(rng, nq, nr):
    """Dense count tuples that honour the counters' invariants (mut2 <=
    diff_obj <= num_ctx <= min(qry_ctx, ref_ctx); diff_obj_section between
    diff_obj + mut2 and 11 diff_obj), with edge cases: empty sketches,
    totals of 1 and 2, no shared context, all contexts shared, no or all
    objects different."""
    qc = rng.integers(0, 300000, size=nq).astype(np.uint32)
    rc = rng.integers(0, 300000, size=nr).astype(np.uint32)
    qc[:8] = [0, 1, 2, 3, 0, 1, 50, 7]
    rc[:8] = [0, 1, 2, 0, 5, 1, 50, 7]
    top = np.minimum(qc[:, None], rc[None, :]).astype(np.int64)
    frac = rng.random((nq, nr)) ** 3
    xny = np.floor(frac * (top + 1)).astype(np.int64)
    xny = np.minimum(xny, top)
    edge = rng.integers(0, 6, size=(nq, nr))
    xny = np.where(edge == 0, 0, np.where(edge == 1, top, xny))
    dobj = np.floor(rng.random((nq, nr)) * (xny + 1)).astype(np.int64)
    dobj = np.where(edge == 2, 0, np.where(edge == 3, xny, np.minimum(dobj, xny)))
    mut2 = np.floor(rng.random((nq, nr)) * (dobj + 1)).astype(np.int64)
    mut2 = np.minimum(mut2, dobj)
    lo, hi = dobj + mut2, 11 * dobj
    dsec = lo + np.floor(rng.random((nq, nr)) * (hi - lo + 1)).astype(np.int64)
    dsec = np.minimum(dsec, hi)
    ctx = np.stack([xny, mut2], -1).astype(np.uint32)
    obj = np.stack([dobj, dsec], -1).astype(np.uint32)
    return ctx, obj, qc, rc


def test_epilogue_one_million_tuples(hip, oracle, rt):
    """The device epilogue over 1,000 x 1,000 random count tuples for every
    metric 3..9, distance and ANI, against the FP64 host epilogue of the
    oracle (orc_matrix_values, itself equal to the reference's printed
    matrices): within 1e-12 relative, and infinities / NaNs in the same
    cells."""
    nq = nr = 1000
    ctx, obj, qc, rc = _random_tuples(np.random.default_rng(41), nq, nr)
    d_ctx, d_obj, d_qc, d_rc = rt.upload(ctx), rt.upload(obj), rt.upload(qc), rt.upload(rc)
    d_out = rt.malloc(8 * nq * nr)
    try:
        for s in range(3, 10):
            for va in (0, 1):
                want = np.zeros(nq * nr, np.float64)
                oracle.lib.orc_matrix_values(kt._p(ctx), kt._p(obj), nq, nr, kt._p(qc), kt._p(rc), s, va, 0.5, 1.0,
                                             8, 20, 44, kt._p(want), 8)
                o = kt.EpilogueOpts(s, va, 0, 0, 0, 0, 8, 20, 44, 0.5, 1.0)
                hip.epilogue_dev(d_ctx, d_obj, nq, nr, d_qc, d_rc, o, d_out)
                got = rt.download(d_out, np.float64, nq * nr)
                fin = np.isfinite(want)
                assert np.array_equal(np.isnan(got), np.isnan(want)), (s, va)
                assert np.array_equal(np.isposinf(got), np.isposinf(want)), (s, va)
                assert np.array_equal(np.isneginf(got), np.isneginf(want)), (s, va)
                rel = np.abs(got[fin] - want[fin]) / np.maximum(np.abs(want[fin]), 1e-300)
                assert float(rel.max(initial=0.0)) <= REL_TOL, (s, va, float(rel.max()))
                assert fin.sum() > nq * nr // 2
    finally:
        for p in (d_ctx, d_obj, d_qc, d_rc, d_out):
            rt.free(p)
