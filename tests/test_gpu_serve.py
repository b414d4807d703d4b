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
