"""CPU tests that pin the oracle (oracle/kssd_oracle.c) before it is trusted as
the checker of the HIP path.  Each check is independent of the oracle's own
code: exhaustive enumeration, brute force on small inputs (the closed form of
the reference's tests/test_unique_context.py:99-104), structural invariants,
and the reference binary's outputs committed under tests/golden/."""
import numpy as np
import pytest

import kssdtest as kt

OBJ_BITS = 20


def test_dna_popcount_exhaustive(oracle):
    # every 20-bit XOR value: number of 2-bit lanes that are nonzero
    x = np.arange(1 << 20, dtype=np.uint32)
    lanes = np.zeros_like(x)
    for bit in range(0, 32, 2):
        lanes += ((x >> bit) & 3) != 0
    sample = np.random.default_rng(0).choice(len(x), 4096, replace=False)
    got = np.array([oracle.lib.orc_dna_popcount(int(v)) for v in x[sample]])
    assert np.array_equal(got, lanes[sample])
    # full 32-bit lanes at the top end
    for v in (0xFFFFFFFF, 0x80000000, 0x40000000, 0xC0000001):
        assert oracle.lib.orc_dna_popcount(v) == sum(((v >> b) & 3) != 0 for b in range(0, 32, 2))


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
