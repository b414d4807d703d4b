"""Failure handling of the kssd3a-hip front end, after the reference's own
checks: a sorted index whose size does not match the sketch
(command_ani.c:1565-1566; tests/test_unique_context.py:295-301), a query
sketch made with another hash (command_ani.c:533), an index demanded by
KSSD3A_FORCE_REF_INDEX but absent, and an output that cannot be written
(tests/test_unique_context.py:204-206).  The input checks run before the
first GPU call, so those tests need no GPU; each also asserts the message,
so a run that got past the check and then failed for want of a GPU would
not pass.  The /dev/full case writes results, so it runs on the GPU."""
import shutil
import struct
from pathlib import Path

import pytest

import kssdtest as kt
from test_cli import run_cli

C2 = kt.GOLDEN / "scale" / "c2"


@pytest.fixture
def sketches(tmp_path):
    ref = shutil.copytree(C2 / "ref", tmp_path / "ref")
    qry = shutil.copytree(C2 / "qry", tmp_path / "qry")
    return Path(ref), Path(qry), tmp_path


def _fails(p, msg):
    assert p.returncode != 0, p.stdout[-500:]
    assert msg in p.stderr, p.stderr
    assert p.stdout == ""


@pytest.mark.parametrize("cut", ["one_record", "minus_one_record", "partial_record", "empty"])
def test_missized_index_is_refused_before_output(sketches, cut):
    ref, qry, tmp = sketches
    idx = ref / "sortedcomb_ctxgid64obj32"
    data = idx.read_bytes()
    assert len(data) % 12 == 0 and len(data) > 24
    new = {"one_record": struct.pack("<QI", 1 << 20, 0),  # the reference's own case
           "minus_one_record": data[:-12],
           "partial_record": data[:-5],
           "empty": b""}[cut]
    idx.write_bytes(new)
    out = tmp / "report.tsv"
    p = run_cli(["ani", "-r", ref, "-q", qry, "-o", out])
    _fails(p, "size does not match the sketch")
    assert not out.exists()


@pytest.mark.parametrize("case", ["c2", "p2"])
def test_hash_id_mismatch_is_refused(tmp_path, case):
    """c2: as many queries as references (index path); p2: 12 queries
    against 200 references (the multi-query stream path)."""
    ref = shutil.copytree(kt.GOLDEN / "scale" / case / "ref", tmp_path / "ref")
    qry = shutil.copytree(kt.GOLDEN / "scale" / case / "qry", tmp_path / "qry")
    stat = Path(qry) / "lcofiles.stat"
    b = bytearray(stat.read_bytes())
    (hid,) = struct.unpack_from("<I", b, 0)  # dim_sketch_stat_t.hash_id comes first
    struct.pack_into("<I", b, 0, hid ^ 1)
    stat.write_bytes(bytes(b))
    out = tmp_path / "report.tsv"
    p = run_cli(["ani", "-r", ref, "-q", qry, "-o", out])
    _fails(p, "hash_id")
    assert not out.exists()


def test_forced_index_must_exist(sketches):
    ref, qry, tmp = sketches
    for f in ("sortedcomb_ctxgid64obj32", "sortedcomb_ctxgid64obj32.unique_bits"):
        (ref / f).unlink(missing_ok=True)
    p = run_cli(["ani", "-r", ref, "-q", qry], env={"KSSD3A_FORCE_REF_INDEX": "1"})
    _fails(p, "requires reference index")


def test_truncated_sketch_offsets_are_refused(sketches):
    ref, qry, tmp = sketches
    ix = qry / "comblco.index"
    ix.write_bytes(ix.read_bytes()[:-8])
    p = run_cli(["ani", "-r", ref, "-q", qry])
    _fails(p, "comblco.index")


@pytest.mark.gpu
@pytest.mark.skipif(not Path("/dev/full").exists(), reason="requires /dev/full")
@pytest.mark.parametrize("extra", [[], ["--unique-ctx", "--qraw"]])
def test_write_failure_returns_nonzero(sketches, extra):
    ref, qry, tmp = sketches
    args = ["ani", "-r", ref, "-q", qry, "-n0", "-f0", "-t1"]
    if extra:
        args = ["ani", "-r", ref, "--qraw", qry, "--unique-ctx", "-n0", "-f0", "-t1"]
    p = run_cli([*args, "-o", "/dev/full"])
    assert p.returncode != 0
    assert "Cannot write" in p.stderr or "No space" in p.stderr, p.stderr
