"""GPU index builder (kssdhip_build_index): the sorted records must be the
bytes of the reference's sortedcomb_ctxgid64obj32 and the unique-marker
bitset the bytes of its .unique_bits, on the reference's own sketches
(tests/golden) and on synthetic sketches checked against an independent
numpy build (stable argsort by (ctxgid, obj); a key's bit is set iff the
key occurs once across all samples, which is write_unique_marker_bits_from_sorted's
rule, command_sketch.c:855-914, since a sample holds each key once).
Context groups cover the three paths of the builder: up to 64 records
(thread per record), up to 4096 (LDS sort per group) and beyond (radix
sort in HBM)."""
from pathlib import Path

import numpy as np
import pytest

import kssdtest as kt

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden"
OBJ_BITS = 20


def _want(keys, index, obj_bits=OBJ_BITS):
    rec = kt.build_index_np(keys, index, obj_bits)
    orig = (rec["ctxgid"] >> np.uint64(20) << np.uint64(obj_bits)) | rec["obj"].astype(np.uint64)
    _, inv, cnt = np.unique(orig, return_inverse=True, return_counts=True)
    bits = np.packbits(cnt[inv] == 1, bitorder="little").tobytes()
    return rec, bits


def _block(samples):
    keys = np.concatenate(samples) if samples else np.zeros(0, np.uint64)
    idx = np.zeros(len(samples) + 1, np.uint64)
    idx[1:] = np.cumsum([len(s) for s in samples])
    return keys, idx


GOLDEN_DIRS = sorted({p.parent for p in GOLDEN.rglob("sortedcomb_ctxgid64obj32")})


@pytest.mark.parametrize("d", GOLDEN_DIRS, ids=lambda p: str(p.relative_to(GOLDEN)))
def test_reference_index_files(hip, d):
    sk = kt.read_sketch_dir(d)
    rec, bits = hip.build_index(sk.keys, sk.index, sk.obj_bits)
    assert rec.tobytes() == (d / "sortedcomb_ctxgid64obj32").read_bytes()
    ub = d / "sortedcomb_ctxgid64obj32.unique_bits"
    if ub.exists():
        assert bits == ub.read_bytes()
    assert bits == _want(sk.keys, sk.index, sk.obj_bits)[1]


@pytest.mark.parametrize("seed,n_samples,n_ctx,mean", [(1, 40, 3000, 400), (2, 300, 20000, 2000),
                                                       (3, 7, 50, 30)])
def test_random_sketches(hip, seed, n_samples, n_ctx, mean):
    rng = np.random.default_rng(seed)
    pool = rng.integers(1, 1 << 40, size=n_ctx, dtype=np.uint64)
    samples = kt.random_samples(rng, n_samples, n_ctx, mean, shared_pool=pool, conflict_p=0.2)
    samples[min(3, n_samples - 1)] = np.zeros(0, np.uint64)  # an empty sample
    keys, idx = _block(samples)
    rec, bits = hip.build_index(keys, idx)
    want_rec, want_bits = _want(keys, idx)
    assert rec.tobytes() == want_rec.tobytes()
    assert bits == want_bits


@pytest.mark.parametrize("group", [65, 700, 4096, 4097, 9000])
def test_large_context_groups(hip, group):
    """`group` samples share context 12345 (objects from a small pool, so
    some repeat and some are unique), plus their own random keys."""
    rng = np.random.default_rng(group)
    samples = []
    for g in range(group):
        own = rng.integers(1 << 30, 1 << 44, size=int(rng.integers(0, 6)), dtype=np.uint64) << np.uint64(20)
        own |= rng.integers(0, 1 << 20, size=len(own), dtype=np.uint64)
        shared = (np.uint64(12345) << np.uint64(20)) | np.uint64(int(rng.integers(0, group // 2 + 3)))
        samples.append(np.unique(np.concatenate([own, [shared]])))
    keys, idx = _block(samples)
    rec, bits = hip.build_index(keys, idx)
    want_rec, want_bits = _want(keys, idx)
    assert rec.tobytes() == want_rec.tobytes()
    assert bits == want_bits
    b = np.unpackbits(np.frombuffer(bits, np.uint8), bitorder="little")[:len(rec)]
    in_group = (rec["ctxgid"] >> np.uint64(20)) == 12345
    assert b[in_group].any() and not b[in_group].all()


def test_conflict_sample_and_tiles(hip):
    """One sample with many objects per context (a kept-conflict sketch),
    over several radix-sort tiles, and contexts spanning all 44 bits."""
    rng = np.random.default_rng(9)
    ctx = rng.integers(0, 1 << 44, size=3000, dtype=np.uint64)
    keys = np.unique(np.concatenate([(np.repeat(ctx, 5) << np.uint64(20)) |
                                     rng.integers(0, 1 << 20, size=15000, dtype=np.uint64),
                                     rng.integers(0, 2 ** 63, size=20000, dtype=np.uint64) * np.uint64(2)]))
    samples = [keys[: len(keys) // 2], keys, keys[::3]]
    k, idx = _block(samples)
    rec, bits = hip.build_index(k, idx)
    want_rec, want_bits = _want(k, idx)
    assert rec.tobytes() == want_rec.tobytes()
    assert bits == want_bits


def test_empty_and_errors(hip):
    rec, bits = hip.build_index(np.zeros(0, np.uint64), np.zeros(1, np.uint64))
    assert len(rec) == 0 and bits == b""
    rec, bits = hip.build_index(np.zeros(0, np.uint64), np.zeros(4, np.uint64))
    assert len(rec) == 0
    with pytest.raises(RuntimeError, match="ascending"):
        hip.build_index(np.array([5 << 20, 3 << 20], np.uint64), np.array([0, 2], np.uint64))
    with pytest.raises(RuntimeError, match="44 bits"):
        hip.build_index(np.array([1 << 62], np.uint64), np.array([0, 1], np.uint64), obj_bits=16)
    with pytest.raises(RuntimeError, match="start at 0"):
        hip.build_index(np.array([1, 2], np.uint64), np.array([1, 2], np.uint64))
