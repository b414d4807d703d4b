"""Serving loop (kssdhip_serve_*): every request's counts are the same alone
and in the mix.  Many queries are submitted from several host threads at
once against a small device budget (so batches are cut by admission and
several slots run concurrently); each result must equal the query counted
alone through kssdhip_count_block, and the oracle's counts."""
import threading

import numpy as np
import pytest

import kssdtest as kt

pytestmark = pytest.mark.gpu
OBJ_BITS = 20


def _fixture(seed, n_ref=40, n_qry=90, pool_n=600, mean=120):
    rng = np.random.default_rng(seed)
    pool = rng.integers(1, 1 << 40, size=pool_n, dtype=np.uint64)
    refs = kt.random_samples(rng, n_ref, pool_n, mean, shared_pool=pool, conflict_p=0.2)
    qrys = kt.random_samples(rng, n_qry, pool_n, mean, shared_pool=pool, conflict_p=0.2)
    qrys[3] = np.zeros(0, np.uint64)  # an empty request
    return refs, qrys


def _block(samples):
    keys = np.concatenate(samples) if samples else np.zeros(0, np.uint64)
    idx = np.zeros(len(samples) + 1, np.uint64)
    idx[1:] = np.cumsum([len(s) for s in samples])
    return keys, idx


def _alone(hip, h, n_ref, q, flags):
    k, i = _block([q])
    ctx, obj, uctx, uobj = hip.count_block(h, n_ref, k, i, flags=flags)
    return ctx[0], obj[0], uctx[0], uobj[0]


@pytest.mark.parametrize("flags", [0, kt.FLAGS["UNIQUE"], kt.FLAGS["IGNORECONFLICT"]])
@pytest.mark.parametrize("slots,max_batch", [(3, 8), (1, 64), (4, 1)])
def test_alone_equals_in_mix(hip, oracle, flags, slots, max_batch):
    refs, qrys = _fixture(7 + flags)
    rk, ri = _block(refs)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    R = len(refs)
    h = hip.open_index(rec, R, OBJ_BITS)
    unique = bool(flags & kt.FLAGS["UNIQUE"])
    try:
        alone = [_alone(hip, h, R, q, flags) for q in qrys]
        qk, qi = _block(qrys)
        want = oracle.count_block(rec, R, OBJ_BITS, qk, qi, flags=flags)
        for n, a in enumerate(alone):
            assert np.array_equal(a[0], want[0][n]) and np.array_equal(a[1], want[1][n]), n
        # budget: a few requests per slot, so admission cuts the batches
        per_req = 2 * 256 + (4 if unique else 2) * ((R * 8 + 255) // 256) * 256 + 8 * 400
        srv = hip.serve_open(h, budget_bytes=slots * 5 * per_req, slots=slots, max_batch=max_batch,
                             flags=flags)
        got = [None] * len(qrys)
        errors = []

        def client(ids):
            try:
                tickets = [(i, hip.serve_submit(srv, qrys[i])) for i in ids]
                for i, t in tickets:
                    assert t >= 0, t
                    got[i] = hip.serve_wait(srv, t, R, unique=unique)
            except Exception as e:  # reported below
                errors.append(e)

        order = np.random.default_rng(11).permutation(len(qrys))
        threads = [threading.Thread(target=client, args=(order[k::4].tolist(),)) for k in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=60)
        assert not errors, errors
        info = hip.serve_info(srv)
        hip.serve_close(srv)
        assert info.queries == len(qrys)
        assert 1 <= info.max_batch_seen <= max_batch
        for n, (a, g) in enumerate(zip(alone, got)):
            assert g is not None, n
            assert np.array_equal(g[0], a[0]) and np.array_equal(g[1], a[1]), n
            if unique:
                assert np.array_equal(g[2], a[2]) and np.array_equal(g[3], a[3]), n
            assert 1 <= g[4].batch_queries <= max_batch
            assert g[4].total_ms >= g[4].device_ms >= 0.0
    finally:
        hip.close_index(h)


def test_oversized_request_is_rejected(hip):
    refs, qrys = _fixture(3, n_qry=5)
    rk, ri = _block(refs)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        srv = hip.serve_open(h, budget_bytes=64 * 1024, slots=1, max_batch=4)
        big = np.arange(1, 20000, dtype=np.uint64) << np.uint64(OBJ_BITS)
        assert hip.serve_submit(srv, big) == kt.KSSDHIP_EINVAL
        t = hip.serve_submit(srv, qrys[0])
        assert t >= 0
        ctx, obj, _, _, st = hip.serve_wait(srv, t, len(refs))
        alone = _alone(hip, h, len(refs), qrys[0], 0)
        assert np.array_equal(ctx, alone[0]) and np.array_equal(obj, alone[1])
        hip.serve_close(srv)
    finally:
        hip.close_index(h)


def test_close_with_queued_requests(hip):
    """Closing drains launched batches and fails queued ones without hanging."""
    refs, qrys = _fixture(5)
    rk, ri = _block(refs)
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        srv = hip.serve_open(h, budget_bytes=1 << 20, slots=1, max_batch=1)
        tickets = [hip.serve_submit(srv, q) for q in qrys]
        assert all(t >= 0 for t in tickets)
        hip.serve_close(srv)
    finally:
        hip.close_index(h)


# ------------------------------------------------------------------ results (select) mode
def _c5_fixture():
    """A C5-shaped slice: 60 references (12 clades of 5 members with
    U(1%, 15%) substitutions, 150-400 kbp ancestors) and two kinds of
    request: assemblies (new clade members, sketched like the references)
    and read-set-shaped sketches of new members (70% of the contexts kept,
    a quarter of those with a second, erroneous object, plus 5% error-only
    contexts), as `sketch --asone --conflict` leaves them."""
    host = kt.Host()
    rng = np.random.default_rng(55)
    nclade, per = 12, 5
    clen = rng.integers(150_000, 400_000, size=nclade)
    specs = [kt.SimSpec(900 + c, int(clen[c]), 5000 + per * c + m, float(rng.uniform(0.01, 0.15)), 0.0, 0, 0.0)
             for c in range(nclade) for m in range(per)]
    rk, ri = host.simsketch(specs, threads=8)
    clades = rng.integers(0, nclade, size=50)
    qspecs = [kt.SimSpec(900 + int(c), int(clen[c]), 70000 + i, float(rng.uniform(0.01, 0.15)), 0.0, 0, 0.0)
              for i, c in enumerate(clades)]
    qk, qi = host.simsketch(qspecs, threads=8)
    reqs = []
    for i in range(50):
        k = qk[int(qi[i]):int(qi[i + 1])]
        if i % 5 >= 3:  # read-set-shaped
            keep = k[rng.random(len(k)) < 0.7]
            errs = keep[rng.random(len(keep)) < 0.25] ^ np.uint64(1 << (2 * int(rng.integers(0, 10))))
            noise = (rng.integers(1, 1 << 40, size=len(k) // 20, dtype=np.uint64) << np.uint64(OBJ_BITS))
            k = np.unique(np.concatenate([keep, errs, noise]))
        reqs.append(k)
    reqs[7] = np.zeros(0, np.uint64)  # an empty request
    return rk, ri, reqs


@pytest.fixture(scope="module")
def c5():
    return _c5_fixture()


def _expected(oracle, rec, R, rctx, q, flags, ctxcut, afcut, anicut, vo, ntop):
    idx = np.array([0, len(q)], np.uint64)
    ctx, obj, uctx, uobj = oracle.count_block(rec, R, OBJ_BITS, q, idx, flags=flags)
    qctx = np.array([len(np.unique(q >> np.uint64(OBJ_BITS)))], np.uint32)
    (pass_ids, _, _), = oracle.index_rank(ctx, obj, uctx, uobj, rctx, qctx, ctxcut=ctxcut, afcut=afcut,
                                          anicut=anicut)
    vals = np.zeros(R, np.float64)
    oracle.lib.orc_matrix_values(kt._p(ctx), kt._p(obj), 1, R, kt._p(qctx), kt._p(rctx), vo.metric, vo.values_ani,
                                 vo.afcut, vo.exception, vo.drfold, OBJ_BITS, 44, kt._p(vals), 1)
    order = sorted(pass_ids.tolist(), key=lambda r: (-vals[r], r))
    if ntop:
        order = order[:ntop]
    return order, ctx[0], obj[0], vals


@pytest.mark.parametrize("flags,ntop", [(0, 0), (kt.FLAGS["IGNORECONFLICT"], 0), (0, 2)])
def test_serve_results_match_oracle(hip, oracle, c5, flags, ntop):
    """Requests of both kinds submitted from four threads to a server that
    returns results: each request's ranked rows are the oracle's passing
    pairs for that request alone (counts bit-exact, FP64 value within
    1e-12), ranked by value, cut to ntop."""
    rk, ri, reqs = c5
    R = len(ri) - 1
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    rctx = np.diff(ri).astype(np.uint32)
    vo = kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0)
    ctxcut, afcut, anicut = 3, 0.02, 0.8
    sel = kt.SelectOpts(ctxcut, afcut, anicut, 0, vo)
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        srv = hip.serve_open(h, budget_bytes=4 << 20, slots=3, max_batch=8, flags=flags, select_opts=sel,
                             ref_ctx=rctx, ntop=ntop)
        got = [None] * len(reqs)
        errors = []

        def client(ids):
            try:
                tickets = [(i, hip.serve_submit(srv, reqs[i])) for i in ids]
                for i, t in tickets:
                    assert t >= 0, t
                    got[i] = hip.serve_wait_rows(srv, t, cap=2)
            except Exception as e:  # reported below
                errors.append(e)

        order = np.random.default_rng(3).permutation(len(reqs))
        threads = [threading.Thread(target=client, args=(order[k::4].tolist(),)) for k in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not errors, errors
        info = hip.serve_info(srv)
        hip.serve_close(srv)
        assert info.queries == len(reqs) and info.max_batch_seen >= 2
        n_rows = 0
        for i, q in enumerate(reqs):
            rows, st = got[i]
            want, ctx, obj, vals = _expected(oracle, rec, R, rctx, q, flags, ctxcut, afcut, anicut, vo, ntop)
            assert rows["ref"].tolist() == want, i
            r = rows["ref"].astype(np.int64)
            assert np.all(rows["query"] == 0)
            assert np.array_equal(rows["num_ctx"], ctx[r, 0]) and np.array_equal(rows["diff_obj"], obj[r, 0])
            assert np.array_equal(rows["num_mut2_ctx"], ctx[r, 1])
            assert np.array_equal(rows["diff_obj_section"], obj[r, 1])
            assert np.all(np.abs(rows["value"] - vals[r]) <= 1e-12 * np.maximum(np.abs(vals[r]), 1e-300))
            assert st.total_ms >= st.device_ms >= 0.0 and 1 <= st.batch_queries <= 8
            n_rows += len(rows)
        assert n_rows > (len(reqs) if ntop == 0 else len(reqs) // 2)
        kinds = [len(got[i][0]) for i in range(len(reqs)) if i % 5 >= 3]
        assert sum(kinds) > 0  # read-set requests find their clades too
    finally:
        hip.close_index(h)


def test_serve_results_refuses_bad_options(hip, c5):
    rk, ri, reqs = c5
    R = len(ri) - 1
    rec = kt.build_index_np(rk, ri, OBJ_BITS)
    rctx = np.diff(ri).astype(np.uint32)
    vo = kt.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, OBJ_BITS, 44, 0.5, 1.0)
    h = hip.open_index(rec, R, OBJ_BITS)
    try:
        with pytest.raises(RuntimeError, match="ctxcut"):
            hip.serve_open(h, 4 << 20, select_opts=kt.SelectOpts(0, 0.2, 0.9, 0, vo), ref_ctx=rctx)
        with pytest.raises(RuntimeError, match="unique_rank"):
            hip.serve_open(h, 4 << 20, select_opts=kt.SelectOpts(3, 0.2, 0.9, 1, vo), ref_ctx=rctx)
        srv = hip.serve_open(h, 4 << 20, select_opts=kt.SelectOpts(3, 0.2, 0.9, 0, vo), ref_ctx=rctx)
        t = hip.serve_submit(srv, reqs[0])
        with pytest.raises(RuntimeError, match="serve_wait_rows"):
            hip.serve_wait(srv, t, R)
        rows, _ = hip.serve_wait_rows(srv, t)
        hip.serve_close(srv)
    finally:
        hip.close_index(h)
