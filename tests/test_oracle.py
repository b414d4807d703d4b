"""CPU tests that pin the oracle (oracle/kssd_oracle.c) before it is trusted as
the checker of the HIP path.  Each check is independent of the oracle's own
code: exhaustive enumeration, brute force on small inputs (the closed form of
the reference's tests/test_unique_context.py:99-104), structural invariants,
and the reference binary's outputs committed under tests/golden/."""
import numpy as np
import pytest

import kssdtest as kt

OBJ_BITS = 20


def _lanes(x):
    """Number of nonzero 2-bit lanes of each u32, lane by lane."""
    lanes = np.zeros(len(x), np.uint8)
    for bit in range(0, 32, 2):
        lanes += (((x >> np.uint32(bit)) & np.uint32(3)) != 0).astype(np.uint8)
    return lanes


def test_dna_popcount_exhaustive(oracle):
    """Every 20-bit XOR value (all 2^20), then 2^20 random 32-bit values and
    the top-lane edge cases: the number of 2-bit lanes that are nonzero."""
    rng = np.random.default_rng(0)
    for x in (np.arange(1 << 20, dtype=np.uint32),
              rng.integers(0, 1 << 32, size=1 << 20, dtype=np.uint64).astype(np.uint32),
              np.array([0xFFFFFFFF, 0x80000000, 0x40000000, 0xC0000001, 0xAAAAAAAA, 0x55555555], np.uint32)):
        got = np.zeros(len(x), np.uint8)
        oracle.lib.orc_dna_popcount_many(kt._p(x), len(x), kt._p(got))
        assert np.array_equal(got, _lanes(x))
    assert oracle.lib.orc_dna_popcount(0xFFFFFFFF) == 16


@pytest.mark.parametrize("seed", range(6))
def test_find_first_is_lower_bound(oracle, seed):
    rng = np.random.default_rng(seed)
    samples = kt.random_samples(rng, 12, 300, 80)
    keys = np.concatenate(samples)
    idx = np.zeros(len(samples) + 1, np.uint64)
    idx[1:] = np.cumsum([len(s) for s in samples])
    rec = kt.build_index_np(keys, idx, OBJ_BITS)
    q = np.unique(np.concatenate([samples[0], rng.integers(1, 1 << 40, 50, dtype=np.uint64) << np.uint64(OBJ_BITS)]))
    got = oracle.find_first(rec, q, OBJ_BITS)
    ctx = rec["ctxgid"] >> np.uint64(20)
    qc = q >> np.uint64(OBJ_BITS)
    pos = np.searchsorted(ctx, qc, side="left")
    found = (pos < len(ctx)) & (ctx[np.minimum(pos, len(ctx) - 1)] == qc)
    want = np.where(found, pos, np.iinfo(np.uint64).max).astype(np.uint64)
    assert np.array_equal(got, want)


def test_build_index_matches_numpy(oracle):
    rng = np.random.default_rng(7)
    samples = kt.random_samples(rng, 9, 200, 60)
    keys = np.concatenate(samples)
    idx = np.zeros(len(samples) + 1, np.uint64)
    idx[1:] = np.cumsum([len(s) for s in samples])
    out = np.zeros(len(keys), dtype=kt.REC_DTYPE)
    oracle.lib.orc_build_index(kt._p(keys), kt._p(idx), len(samples), OBJ_BITS, kt._p(out))
    assert out.tobytes() == kt.build_index_np(keys, idx, OBJ_BITS).tobytes()


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("ignore", [False, True])
def test_pair_features_brute_force(oracle, seed, ignore):
    rng = np.random.default_rng(100 + seed)
    a, b = kt.random_samples(rng, 2, 120, 60, conflict_p=0.25)
    assert oracle.pair(a, b, OBJ_BITS, ignore) == kt.brute_counts(a, b, OBJ_BITS, ignore)


def _block(samples):
    keys = np.concatenate(samples) if samples else np.zeros(0, np.uint64)
    idx = np.zeros(len(samples) + 1, np.uint64)
    idx[1:] = np.cumsum([len(s) for s in samples])
    return keys, idx


@pytest.mark.parametrize("seed", range(8))
def test_count_block_full_lower_unique(oracle, seed):
    rng = np.random.default_rng(200 + seed)
    pool = rng.integers(1, 1 << 40, size=150, dtype=np.uint64)
    refs = kt.random_samples(rng, 7, 150, 40, shared_pool=pool, conflict_p=0.2)
    qrys = kt.random_samples(rng, 5, 150, 50, shared_pool=pool, conflict_p=0.2)
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    qkeys, qidx = _block(qrys)
    for ignore in (0, kt.FLAGS["IGNORECONFLICT"]):
        ctx, obj, uctx, uobj = oracle.count_block(rec, len(refs), OBJ_BITS, qkeys, qidx,
                                                  flags=ignore | kt.FLAGS["UNIQUE"], threads=2)
        owners = {}
        for g, r in enumerate(refs):
            for c in set((r >> np.uint64(OBJ_BITS)).tolist()):
                owners.setdefault(c, set()).add(g)
        for qi, q in enumerate(qrys):
            for g, r in enumerate(refs):
                xny, dobj, dsec, mut2 = kt.brute_counts(q, r, OBJ_BITS, bool(ignore))
                assert (ctx[qi, g, 0], obj[qi, g, 0], obj[qi, g, 1], ctx[qi, g, 1]) == (xny, dobj, dsec, mut2)
                # unique-owner counts: contexts owned by exactly this reference
                # (ownership counts every reference, command_ani.h:564-574)
                sole = {c for c, o in owners.items() if o == {g}}
                rmask = np.isin(r >> np.uint64(OBJ_BITS), np.array(sorted(sole), dtype=np.uint64))
                want = kt.brute_counts(q, r[rmask], OBJ_BITS, bool(ignore))
                assert (uctx[qi, g, 0], uobj[qi, g, 0], uobj[qi, g, 1], uctx[qi, g, 1]) == want
    # lower triangle of a self comparison == full counts restricted to gid < qgid
    keys, idx = _block(refs)
    full = oracle.count_block(rec, len(refs), OBJ_BITS, keys, idx, threads=2)
    low = oracle.count_block(rec, len(refs), OBJ_BITS, keys, idx, flags=kt.FLAGS["LOWER"], threads=2)
    tri = np.tril(np.ones((len(refs), len(refs)), bool), -1)
    for a, b in ((full[0], low[0]), (full[1], low[1])):
        assert np.array_equal(np.where(tri[..., None], a, 0), b)
    # a block that starts mid-collection uses absolute offsets and gid offset
    low2 = oracle.count_block(rec, len(refs), OBJ_BITS, keys[idx[3]:], idx[3:], qry_gid_offset=3,
                              flags=kt.FLAGS["LOWER"], threads=3)
    assert np.array_equal(low2[0], low[0][3:]) and np.array_equal(low2[1], low[1][3:])


def test_count_block_thread_invariance(oracle):
    rng = np.random.default_rng(11)
    pool = rng.integers(1, 1 << 40, size=400, dtype=np.uint64)
    refs = kt.random_samples(rng, 30, 400, 100, shared_pool=pool)
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    a = oracle.count_block(rec, 30, OBJ_BITS, rkeys, ridx, threads=1)
    b = oracle.count_block(rec, 30, OBJ_BITS, rkeys, ridx, threads=4)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_formula_special_cases(oracle):
    F = kt.Features
    # command_ani.h:204-205: no shared context -> naive distance 1
    assert oracle.lib.orc_naive_dist(F(0, 0, 0, 0, 0)) == 1.0
    # identical sketches: zero mismatches -> 0 for naive, CtxMoE, p_dist
    same = F(1000, 0, 0, 0, 0)
    assert oracle.lib.orc_naive_dist(same) == 0.0
    assert oracle.lib.orc_lm3ways_dist(same, 8) == 0.0
    assert oracle.lib.orc_p_dist(same, 20) == 0.0
    # model_ani.h:292-293
    assert oracle.lib.orc_lm3ways_dist(F(0, 0, 5, 1, 3), 8) == 1.0
    # p_dist closed form: sections / (XnY * obj_bits/2)
    assert oracle.lib.orc_p_dist(F(200, 0, 50, 3, 40), 20) == 50 / (200 * 10)
    # Mash/AAF on identical sets: distance 0
    assert oracle.lib.orc_mashD(22, 500, 500, 500) == 0.0
    assert oracle.lib.orc_aafD(22, 500, 800, 500) == 0.0


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("ignore", [False, True])
def test_stream_one_query_equals_index_counts(oracle, seed, ignore):
    """The one-query stream restatement (hash of the query's contexts,
    references streamed) gives each reference the counters the index path
    gives the same pair (a different code path: fenceposts over the sorted
    index), on sketches with conflicting objects; with afcut 0 every pair
    with a shared context passes, and ref_ctx counts the walked contexts."""
    rng = np.random.default_rng(300 + seed)
    pool = rng.integers(1, 1 << 40, size=400, dtype=np.uint64)
    refs = kt.random_samples(rng, 30, 400, 120, shared_pool=pool, conflict_p=0.2)
    q = kt.random_samples(rng, 1, 400, 150, shared_pool=pool, conflict_p=0.2)[0]
    rk = np.concatenate(refs)
    ri = np.zeros(len(refs) + 1, np.uint64)
    ri[1:] = np.cumsum([len(s) for s in refs])
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    flags = kt.FLAGS["IGNORECONFLICT"] if ignore else 0
    want_c, want_o, _, _ = oracle.count_block(rec, len(refs), OBJ_BITS, q, np.array([0, len(q)], np.uint64),
                                              flags=flags)
    ctx, obj, rctx, ok, qc = oracle.stream_one_query(q, rk, ri, ignoreconflict=ignore)
    assert np.array_equal(ctx, want_c[0]) and np.array_equal(obj, want_o[0])
    assert qc == len(np.unique(q >> np.uint64(OBJ_BITS)))
    for r, s in enumerate(refs):
        c, n = np.unique(s >> np.uint64(OBJ_BITS), return_counts=True)
        assert rctx[r] == ((n == 1).sum() if ignore else len(c))
    assert np.array_equal(ok, ctx[:, 0] >= 1)


@pytest.mark.parametrize("afcut", [0.05, 0.2, 0.5])
def test_stream_one_query_af_gate(oracle, afcut):
    """With afcut > 0 the stream path gives up a reference once the shared
    contexts so far plus the contexts after the current one fall short of
    the needed count, checked before the current context is looked up
    (command_ani.c:2154-2156).  So a reference passes exactly when its full
    counts meet ctxcut, that count and afcut (command_ani.h:75-90) and it
    does not reach the count exactly on its last context (computed here from
    the index path's counts and the sketches)."""
    rng = np.random.default_rng(11)
    pool = rng.integers(1, 1 << 40, size=300, dtype=np.uint64)
    refs = kt.random_samples(rng, 40, 300, 100, shared_pool=pool)
    q = kt.random_samples(rng, 1, 300, 100, shared_pool=pool)[0]
    rk = np.concatenate(refs)
    ri = np.zeros(len(refs) + 1, np.uint64)
    ri[1:] = np.cumsum([len(s) for s in refs])
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    full, _, _, _ = oracle.count_block(rec, len(refs), OBJ_BITS, q, np.array([0, len(q)], np.uint64))
    _, _, _, ok, qc = oracle.stream_one_query(q, rk, ri, ctxcut=3, afcut=afcut)
    x = full[0, :, 0].astype(np.float64)
    tot = np.array([len(np.unique(s >> np.uint64(OBJ_BITS))) for s in refs], np.float64)
    need = np.ceil(np.float64(np.float32(afcut)) * np.minimum(qc, tot))
    want = (x >= 3) & (x >= need) & (np.maximum(x / qc, x / np.maximum(tot, 1)) >= np.float64(np.float32(afcut)))
    qctx = set((q >> np.uint64(OBJ_BITS)).tolist())
    last_shared = np.array([len(s) > 0 and int(s[-1] >> np.uint64(OBJ_BITS)) in qctx for s in refs])
    boundary = (x == need) & (need > 0) & last_shared
    assert np.array_equal(ok, want & ~boundary)
    assert 0 < (want & ~boundary).sum() < len(refs)


def test_stream_one_query_matches_reference_rows(oracle):
    """The reference binary's one-query stream output (stream_qry1_all_s3:
    `ani -r ref -q qry1 -s3 -n0 -f0 -t1`) lists exactly the references the
    restatement passes, with the same XnY_ctx, N_diff_obj,
    N_diff_obj_section and N_mut2_ctx."""
    g = kt.GOLDEN / "genomes"
    ref, qry = kt.read_sketch_dir(g / "ref"), kt.read_sketch_dir(g / "qry1")
    ctx, obj, _, ok, _ = oracle.stream_one_query(qry.keys, ref.keys, ref.index, ref_conflict=bool(ref.conflict))
    lines = (g / "stream_qry1_all_s3.tsv").read_text().splitlines()
    head = lines[0].split("\t")
    col = {k: head.index(k) for k in ("Ref", "XnY_ctx", "N_diff_obj", "N_diff_obj_section", "N_mut2_ctx")}
    rows = {c[col["Ref"]]: c for c in (ln.split("\t") for ln in lines[1:])}
    assert set(rows) == {ref.names[r] for r in np.flatnonzero(ok)}
    for name, c in rows.items():
        r = ref.names.index(name)
        assert [int(c[col[k]]) for k in ("XnY_ctx", "N_diff_obj", "N_diff_obj_section", "N_mut2_ctx")] == \
            [int(ctx[r, 0]), int(obj[r, 0]), int(obj[r, 1]), int(ctx[r, 1])]


@pytest.mark.parametrize("stem,ignore", [("afb_stream_one", False), ("afb_stream_one_ignoreconflict", True)])
def test_stream_one_query_af_boundary_matches_reference(oracle, stem, ignore):
    """References that share exactly the needed number of contexts: the
    reference binary's one-query stream path keeps those whose last walked
    context is not shared and drops the others (tests/golden/scale/afb_*);
    the restatement passes the same set."""
    g = kt.GOLDEN / "scale" / "afb"
    ref, qry = kt.read_sketch_dir(g / "ref"), kt.read_sketch_dir(g / "qry1")
    _, _, _, ok, _ = oracle.stream_one_query(qry.keys, ref.keys, ref.index, ignoreconflict=ignore,
                                             ref_conflict=bool(ref.conflict), ctxcut=1, afcut=0.5)
    lines = (kt.GOLDEN / "scale" / f"{stem}.tsv").read_text().splitlines()[1:]
    assert {ln.split("\t")[1] for ln in lines} == {ref.names[r] for r in np.flatnonzero(ok)}
    assert "r_shared40_last_shared" not in {ln.split("\t")[1] for ln in lines}
