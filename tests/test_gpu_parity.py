"""GPU parity: the HIP path (through the C ABI, libkssdhip.so) against the CPU
oracle on the same seeded inputs.  Counts must be bit-exact."""
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


def _fixture(seed, n_ref=9, n_qry=7, pool_n=300, mean=60, conflict_p=0.2, obj_bits=OBJ_BITS):
    rng = np.random.default_rng(seed)
    pool = rng.integers(1, 1 << (64 - obj_bits - 1), size=pool_n, dtype=np.uint64)
    refs = kt.random_samples(rng, n_ref, pool_n, mean, obj_bits=obj_bits, shared_pool=pool,
                             conflict_p=conflict_p)
    qrys = kt.random_samples(rng, n_qry, pool_n, mean, obj_bits=obj_bits, shared_pool=pool,
                             conflict_p=conflict_p)
    return refs, qrys


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("flags", [0, 2, 4, 6])
def test_count_block_matches_oracle(hip, oracle, seed, flags):
    refs, qrys = _fixture(seed)
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    qkeys, qidx = _block(qrys)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        got = hip.count_block(h, len(refs), qkeys, qidx, flags=flags)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, qkeys, qidx, flags=flags)
    names = ("ctx", "obj", "unique_ctx", "unique_obj")
    for name, g, w in zip(names, got, want):
        assert np.array_equal(g, w), (seed, flags, name)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("ignore", [0, 4])
def test_lower_triangle_blocks(hip, oracle, seed, ignore):
    refs, _ = _fixture(1000 + seed, n_ref=23)
    keys, idx = _block(refs)
    rec = kt.build_index_np(keys, idx, OBJ_BITS)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        # blocks of 5 with absolute offsets, like command_ani.c:1694-1701
        for off in range(0, len(refs), 5):
            end = min(off + 5, len(refs))
            got = hip.count_block(h, len(refs), keys[idx[off]:idx[end]], idx[off:end + 1],
                                  qry_gid_offset=off, flags=1 | ignore)
            want = oracle.count_block(rec, len(refs), OBJ_BITS, keys[idx[off]:idx[end]],
                                      idx[off:end + 1], qry_gid_offset=off, flags=1 | ignore)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), off
    finally:
        hip.close_index(h)


@pytest.mark.parametrize("obj_bits", [20, 24, 28])
def test_object_widths(hip, oracle, obj_bits):
    # NUM_CODENS 11/10/9 with klen 32 give 20/24/28 object bits
    # (tests/test_unique_context.py:257-266 in the reference)
    refs, qrys = _fixture(77, obj_bits=obj_bits)
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, obj_bits)
    qkeys, qidx = _block(qrys)
    h = hip.open_index(rec, len(refs), obj_bits)
    try:
        got = hip.count_block(h, len(refs), qkeys, qidx, flags=2)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), obj_bits, qkeys, qidx, flags=2)
    assert all(np.array_equal(g, w) for g, w in zip(got, want))
    for qi, q in enumerate(qrys[:3]):
        for g, r in enumerate(refs[:3]):
            assert hip.pair(q, r, obj_bits) == oracle.pair(q, r, obj_bits)


def test_empty_and_ragged(hip, oracle):
    refs, qrys = _fixture(5, n_ref=6, n_qry=6)
    refs[2] = np.zeros(0, np.uint64)          # empty reference
    qrys[0] = np.zeros(0, np.uint64)          # empty first query
    qrys[3] = np.zeros(0, np.uint64)          # empty middle query
    qrys[5] = qrys[5][:1]                     # single-entry query
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    qkeys, qidx = _block(qrys)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        got = hip.count_block(h, len(refs), qkeys, qidx, flags=2)
        assert hip.count_block(h, len(refs), qkeys[:0], qidx[:1])[0].shape == (0, len(refs), 2)
        # all-empty block
        got0 = hip.count_block(h, len(refs), np.zeros(0, np.uint64), np.zeros(3, np.uint64))
        assert not got0[0].any() and not got0[1].any()
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, qkeys, qidx, flags=2)
    assert all(np.array_equal(g, w) for g, w in zip(got, want))
    assert hip.pair(np.zeros(0, np.uint64), refs[0], OBJ_BITS) == (0, 0, 0, 0)
    assert hip.pair(refs[0], np.zeros(0, np.uint64), OBJ_BITS) == (0, 0, 0, 0)


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("ignore", [False, True])
def test_pair_matches_oracle(hip, oracle, seed, ignore):
    refs, qrys = _fixture(300 + seed, n_ref=3, n_qry=3, mean=2000, pool_n=5000, conflict_p=0.3)
    for a in qrys:
        for b in refs:
            assert hip.pair(a, b, OBJ_BITS, ignore) == oracle.pair(a, b, OBJ_BITS, ignore)
    assert hip.pair(refs[0], refs[0], OBJ_BITS)[1:] == (0, 0, 0)


def test_find_first_matches_oracle(hip, oracle):
    refs, qrys = _fixture(9, n_ref=15, mean=500, pool_n=4000)
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    q = np.unique(np.concatenate(qrys))
    big = np.uint64(((1 << 43) - 5) << OBJ_BITS)
    q = np.unique(np.concatenate([q, [big, np.uint64(0)]]))
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        got = hip.find_first(h, q)
    finally:
        hip.close_index(h)
    assert np.array_equal(got, oracle.find_first(rec, q, OBJ_BITS))


def test_larger_random_block(hip, oracle):
    # a few thousand queries keys against a ~100k-record index
    rng = np.random.default_rng(4242)
    pool = rng.integers(1, 1 << 43, size=60000, dtype=np.uint64)
    refs = kt.random_samples(rng, 40, 60000, 2500, shared_pool=pool, conflict_p=0.05)
    qrys = kt.random_samples(rng, 64, 60000, 2500, shared_pool=pool, conflict_p=0.05)
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    qkeys, qidx = _block(qrys)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        got = hip.count_block(h, len(refs), qkeys, qidx, flags=2)
        again = hip.count_block(h, len(refs), qkeys, qidx, flags=2)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, qkeys, qidx, flags=2)
    assert all(np.array_equal(g, w) for g, w in zip(got, want))
    # run-to-run determinism
    assert all(np.array_equal(g, w) for g, w in zip(got, again))


def test_staged_host_copies(hip, oracle):
    """Host buffers over 8 MiB are copied through pinned chunks of 2 MiB by
    four host threads, unevenly shared here (7 and 5 chunks, a 11 KiB tail):
    a 12.3 MiB index upload, 8.8 MiB of query keys up and four 8.01 MiB
    counter arrays down (unique counters on), and the index builder's
    8.2 MiB of keys up and 12.3 MiB of records down."""
    rng = np.random.default_rng(777)
    pool = rng.integers(1, 1 << 43, size=200000, dtype=np.uint64)

    def samples(n, mean):  # contexts drawn from the shared pool (repeats: conflicts), random objects
        out = []
        for _ in range(n):
            m = int(rng.integers(mean // 2, mean * 3 // 2))
            ctx = pool[rng.integers(0, len(pool), size=m)]
            out.append(np.unique((ctx << np.uint64(OBJ_BITS)) | rng.integers(0, 1 << OBJ_BITS, size=m, dtype=np.uint64)))
        return out

    refs, qrys = samples(300, 3600), samples(3500, 330)
    rkeys, ridx = _block(refs)
    qkeys, qidx = _block(qrys)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    assert min(rec.nbytes, qkeys.nbytes, rkeys.nbytes, len(refs) * len(qrys) * 8) > 8 << 20
    built, _ = hip.build_index(rkeys, ridx)
    assert built.tobytes() == rec.tobytes()
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        got = hip.count_block(h, len(refs), qkeys, qidx, flags=6)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, qkeys, qidx, flags=6)
    assert len(got) == len(want) == 4
    assert all(np.array_equal(g, w) for g, w in zip(got, want))


def test_index_open_rejects_bad_index(hip):
    refs, _ = _fixture(3, n_ref=4)
    keys, idx = _block(refs)
    rec = kt.build_index_np(keys, idx, OBJ_BITS)
    with pytest.raises(RuntimeError, match="not sorted"):
        hip.open_index(rec[::-1].copy(), 4, OBJ_BITS)
    with pytest.raises(RuntimeError, match="gid"):
        hip.open_index(rec, 2, OBJ_BITS)


# (mode, chunk_records, block_threads, group_partitions)
TILINGS = [(1, 4096, 256, 1), (2, 64, 256, 1), (2, 128, 64, 1), (2, 4096, 256, 1), (2, 12288, 1024, 1),
           (2, 64, 256, 4), (2, 1024, 256, 8), (2, 6144, 512, 2), (2, 128, 64, 64)]


@pytest.mark.parametrize("mode,chunk,block,group", TILINGS)
@pytest.mark.parametrize("flags", [0, 2, 4, 6])
def test_tilings_match_oracle(hip, oracle, mode, chunk, block, group, flags):
    """Lookup and partitioned paths, small partitions (many boundaries, runs
    longer than a partition), the largest LDS stage, and partition groups
    whose LDS cell table holds every touched cell or overflows to the global
    counters (64 partitions of 128 records: about 4000 cells per group)."""
    rng = np.random.default_rng(chunk + flags + mode)
    pool = rng.integers(1, 1 << 43, size=3000, dtype=np.uint64)
    refs = kt.random_samples(rng, 60, 3000, 300, shared_pool=pool, conflict_p=0.15)
    # one context carried by every reference with two objects: a run of 120
    # records, longer than the 64/128-record partitions plus their slack
    hot = np.uint64(int(pool[7]) << 20)
    refs = [np.unique(np.concatenate([r, [hot | np.uint64(5), hot | np.uint64(9)]])) for r in refs]
    qrys = kt.random_samples(rng, 33, 3000, 400, shared_pool=pool, conflict_p=0.15)
    qrys[4] = np.zeros(0, np.uint64)
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    qkeys, qidx = _block(qrys)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        hip.set_tiling(h, mode=mode, chunk_records=chunk, block_threads=block, group_partitions=group)
        got = hip.count_block(h, len(refs), qkeys, qidx, flags=flags)
        low = hip.count_block(h, len(refs), rkeys[ridx[10]:], ridx[10:], qry_gid_offset=10,
                              flags=1 | (flags & 4))
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, qkeys, qidx, flags=flags)
    for name, g, w in zip(("ctx", "obj", "uctx", "uobj"), got, want):
        assert np.array_equal(g, w), (mode, chunk, group, flags, name)
    wl = oracle.count_block(rec, len(refs), OBJ_BITS, rkeys[ridx[10]:], ridx[10:], qry_gid_offset=10,
                            flags=1 | (flags & 4))
    assert np.array_equal(low[0], wl[0]) and np.array_equal(low[1], wl[1])


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("group", [1, 4])
def test_split_table_both_sides(hip, oracle, dense, group):
    """The split table from the partitions' side (few keys per table entry:
    queries of 0-3 keys against 64-record partitions) and from the keys' side
    (many keys per entry), with empty queries, single-key queries, keys below
    the first and above the last indexed context, and tile-boundary query
    ends (64-key tiles, 2048-key wave runs)."""
    rng = np.random.default_rng(91 + dense + 10 * group)
    pool = rng.integers(1 << 8, 1 << 42, size=4000, dtype=np.uint64)
    refs = kt.random_samples(rng, 50, 4000, 400, shared_pool=pool, conflict_p=0.1)
    if dense:
        sizes = [0, 1, 63, 64, 65, 0, 2047, 2048, 2049, 700, 0, 1, 3000]
    else:
        sizes = [0, 1, 3, 0, 2, 1, 0, 3, 1, 2, 0, 0, 1]
    qrys = []
    for n in sizes:
        ctx = np.unique(np.concatenate([rng.choice(pool, size=n), [np.uint64(1), np.uint64((1 << 44) - 1)]]))
        ctx = rng.permutation(ctx)[:n]
        qrys.append(np.sort((ctx << np.uint64(20)) | rng.integers(0, 1 << 20, size=len(ctx), dtype=np.uint64)))
    rkeys, ridx = _block(refs)
    rec = kt.build_index_np(rkeys, ridx, OBJ_BITS)
    qkeys, qidx = _block(qrys)
    h = hip.open_index(rec, len(refs), OBJ_BITS)
    try:
        hip.set_tiling(h, mode=2, chunk_records=64, block_threads=256, group_partitions=group)
        got = hip.count_block(h, len(refs), qkeys, qidx, flags=2)
    finally:
        hip.close_index(h)
    want = oracle.count_block(rec, len(refs), OBJ_BITS, qkeys, qidx, flags=2)
    for name, g, w in zip(("ctx", "obj", "uctx", "uobj"), got, want):
        assert np.array_equal(g, w), (dense, group, name)


def test_tiling_validation(hip):
    refs, _ = _fixture(3, n_ref=4)
    keys, idx = _block(refs)
    h = hip.open_index(kt.build_index_np(keys, idx, OBJ_BITS), 4, OBJ_BITS)
    try:
        t = hip.get_tiling(h)
        assert (t.mode, t.chunk_records, t.block_threads, t.group_partitions) == (0, 4096, 256, 1)
        for bad in (dict(mode=4), dict(mode=-1), dict(chunk_records=10), dict(chunk_records=20000),
                    dict(block_threads=100), dict(group_partitions=0), dict(group_partitions=65),
                    dict(chunk_records=8192, group_partitions=2), dict(query_threads=32),
                    dict(query_threads=576), dict(query_threads=1024), dict(query_slots=300), dict(query_slots=128),
                    dict(query_slots=16384)):
            with pytest.raises(RuntimeError):
                hip.set_tiling(h, **bad)
    finally:
        hip.close_index(h)


@pytest.mark.parametrize("obj_bits", [20, 32])
@pytest.mark.parametrize("mode", [1, 3])
def test_device_popcount_exhaustive(hip, obj_bits, mode):
    """dna_popcount on the device over every 20-bit object difference (and
    2^20 random 32-bit ones): 2^20 one-key queries against one reference
    holding each query's context with object 0.  Per query the pair has one
    shared context, and the minimum mismatch is the lane count of the
    query's object, capped at 12 (NUM_CODENS + 1, command_ani.h:535)."""
    n = 1 << 20
    rng = np.random.default_rng(obj_bits)
    objs = (np.arange(n, dtype=np.uint64) if obj_bits == 20
            else rng.integers(0, 1 << 32, size=n, dtype=np.uint64))
    ctxs = np.arange(1, n + 1, dtype=np.uint64)
    ref_keys = ctxs << np.uint64(obj_bits)
    ref_idx = np.array([0, n], np.uint64)
    qry_keys = (ctxs << np.uint64(obj_bits)) | objs
    qry_idx = np.arange(n + 1, dtype=np.uint64)
    rec = kt.build_index_np(ref_keys, ref_idx, obj_bits)
    lanes = np.zeros(n, np.uint32)
    for bit in range(0, 32, 2):
        lanes += ((objs >> np.uint64(bit)) & np.uint64(3)) != 0
    m = np.minimum(lanes, 12)
    h = hip.open_index(rec, 1, obj_bits)
    try:
        hip.set_tiling(h, mode=mode)
        ctx, obj, _, _ = hip.count_block(h, 1, qry_keys, qry_idx)
    finally:
        hip.close_index(h)
    assert np.all(ctx[:, 0, 0] == 1)
    assert np.array_equal(ctx[:, 0, 1], (m > 1).astype(np.uint32))
    assert np.array_equal(obj[:, 0, 0], (m > 0).astype(np.uint32))
    assert np.array_equal(obj[:, 0, 1], m)
