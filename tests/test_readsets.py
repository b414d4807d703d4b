"""Read-set sketch modes and --estimate-coverage against the reference.

tests/golden/readsets holds FASTQ (gzip and plain) and FASTA inputs, the
SHA-256 of every file the reference's `sketch` wrote for them in the counted
modes (-A abundance with its lcofiles.qc read-QC record, --asone,
-n/--LstKmerOcrs, --readsQC, --npercentile/--ncap, --sample-name; the
stat header and names for lcofiles.stat), and the reference's
`ani --qraw ... --estimate-coverage` detail output on those sketches
(make_golden.py, make_readsets).

CPU: `kssd3a-hip sketch` rebuilds every directory; each file must match the
reference's digest.  The read-QC range of lcofiles.qc is also checked
against properties that fix it independently of the reference: it contains
the histogram's mode, -n/--readsQC drop exactly the keys outside the range.
GPU: the coverage TSVs, byte for byte, on the rebuilt sketches (index path,
multi- and single-query stream paths, a reference without unique-marker
bits, a query without comblco.a, -N 1).
"""
import hashlib
import json
import shutil
import subprocess

import numpy as np
import pytest

import kssdtest as kt
from test_cli import BIN, run_cli

RS = kt.GOLDEN / "readsets"
DIGESTS = json.loads((RS / "sketch_digests.json").read_text())
SKETCHES = {  # the reference's `sketch` arguments (make_golden.py READSET_SKETCHES)
    "rs_A": ["-A", "-f", "2", "a_1.fq.gz", "a_2.fq.gz", "b.fq.gz", "c.fq"],
    "rs_plain": ["-f", "2", "a_1.fq.gz", "b.fq.gz"],
    "rs_asone": ["--asone", "--conflict", "-A", "-f", "2", "a_1.fq.gz", "a_2.fq.gz"],
    "rs_n2": ["-n", "2", "-f", "2", "a_1.fq.gz", "b.fq.gz", "c.fq"],
    "rs_qc": ["--readsQC", "-A", "-f", "2", "a_1.fq.gz", "b.fq.gz"],
    "rs_pct": ["--npercentile", "0.1", "--ncap", "3", "-A", "-f", "2", "a_1.fq.gz", "b.fq.gz"],
    "rs_named": ["--asone", "-A", "--sample-name", "sampleA", "-f", "2", "a_1.fq.gz", "a_2.fq.gz", "ga.fa"],
    "asm_asone": ["--asone", "-f", "2", "ga.fa", "gb.fa"],
    "refs": ["-f", "2", "ga.fa", "gb.fa", "gc.fa"],
}


def sketch(out, args):
    p = subprocess.run([str(BIN), "sketch", "-p", "4", "-o", str(out), *args], capture_output=True,
                       text=True, cwd=RS, timeout=120)
    assert p.returncode == 0, p.stderr
    return p


def check_dir(out, name):
    files = {k.split("/", 1)[1]: v for k, v in DIGESTS.items() if k.split("/", 1)[0] == name}
    produced = {f.name for f in out.iterdir()}
    assert produced == set(files), (name, produced ^ set(files))
    for f, want in files.items():
        data = (out / f).read_bytes()
        if f == "lcofiles.stat":
            assert data[:32].hex() == want["header"], (name, f)
            assert kt.read_sketch_dir(out).names == want["names"], (name, f)
        else:
            assert hashlib.sha256(data).hexdigest() == want, (name, f)


@pytest.mark.parametrize("name", SKETCHES)
def test_sketch_modes_match_reference(tmp_path, name):
    out = tmp_path / name
    sketch(out, SKETCHES[name])
    if name == "refs":  # the reference's digests are of the indexed directory
        p = run_cli(["sketch", "-i", str(out), "--host-index"])
        assert p.returncode == 0, p.stderr
    check_dir(out, name)


def test_sketch_thread_count_invariant(tmp_path):
    """-p 1 and -p 4 write the same files (the reference's per-file and
    in-file parallel modes agree on well-formed FASTQ)."""
    a, b = tmp_path / "p1", tmp_path / "p4"
    for out, p in ((a, "1"), (b, "4")):
        r = subprocess.run([str(BIN), "sketch", "-p", p, "-o", str(out), *SKETCHES["rs_A"]],
                           capture_output=True, text=True, cwd=RS, timeout=120)
        assert r.returncode == 0, r.stderr
    for f in a.iterdir():
        assert f.read_bytes() == (b / f.name).read_bytes(), f.name


def _qc(out):
    return np.fromfile(out / "lcofiles.qc", dtype=np.uint32).reshape(-1, 4)


def test_read_qc_range_properties(tmp_path):
    """The recorded read-QC range of each sample holds the count histogram's
    chosen mode; --readsQC keeps exactly the keys whose count lies inside
    [max(2, lower), upper] before conflicts are dropped; -n 2 keeps exactly
    the keys seen at least twice."""
    ab, qc = tmp_path / "ab", tmp_path / "qc"
    sketch(ab, ["-A", "--conflict", "-f", "2", "a_1.fq.gz", "b.fq.gz"])
    sketch(qc, ["--readsQC", "-A", "--conflict", "-f", "2", "a_1.fq.gz", "b.fq.gz"])
    ka, kq = kt.read_sketch_dir(ab), kt.read_sketch_dir(qc)
    ca = np.fromfile(ab / "comblco.a", dtype=np.uint32)
    cq = np.fromfile(qc / "comblco.a", dtype=np.uint32)
    for i, (flags, lo, hi, mode) in enumerate(_qc(ab)):
        assert flags == 1 and 2 <= lo <= mode <= hi
        a0, a1 = ka.index[i], ka.index[i + 1]
        keep = (ca[a0:a1] >= max(2, lo)) & (ca[a0:a1] <= hi)
        q0, q1 = kq.index[i], kq.index[i + 1]
        assert np.array_equal(ka.keys[a0:a1][keep], kq.keys[q0:q1])
        assert np.array_equal(ca[a0:a1][keep], cq[q0:q1])
    assert (_qc(qc)[:, 0] == 3).all()  # valid and applied
    n2 = tmp_path / "n2"
    sketch(n2, ["-n", "2", "--conflict", "-f", "2", "a_1.fq.gz", "b.fq.gz"])
    kn = kt.read_sketch_dir(n2)
    for i in range(2):
        a0, a1 = ka.index[i], ka.index[i + 1]
        assert np.array_equal(ka.keys[a0:a1][ca[a0:a1] >= 2], kn.keys[kn.index[i]:kn.index[i + 1]])


def test_counts_are_occurrences(tmp_path):
    """comblco.a counts every k-mer window of the inputs: an --asone sketch
    of a file twice over doubles every count of the file alone."""
    one, two = tmp_path / "one", tmp_path / "two"
    sketch(one, ["-A", "--conflict", "-f", "2", "c.fq"])
    sketch(two, ["--asone", "-A", "--conflict", "-f", "2", "c.fq", "c.fq"])
    assert (one / "comblco").read_bytes() == (two / "comblco").read_bytes()
    c1 = np.fromfile(one / "comblco.a", dtype=np.uint32)
    c2 = np.fromfile(two / "comblco.a", dtype=np.uint32)
    assert np.array_equal(c2, 2 * c1)


@pytest.mark.parametrize("args,msg", [
    (["--sample-name", "x", "-o", "o", "a.fq"], "--sample-name requires"),
    (["--npercentile", "1.5", "-o", "o", "a.fq"], "--npercentile"),
    (["-n", "0", "-o", "o", "a.fq"], "-n/--LstKmerOcrs"),
])
def test_sketch_option_errors(args, msg):
    p = run_cli(["sketch", *args])
    assert p.returncode == 64 and msg in p.stderr, p.stderr


def test_estimate_coverage_option_errors():
    p = run_cli(["ani", "-r", "x", "-q", "y", "--estimate-coverage"])
    assert p.returncode == 64 and "--qraw" in p.stderr, p.stderr
    p = run_cli(["ani", "-r", "x", "--qraw", "y", "-m1", "--estimate-coverage"])
    assert p.returncode == 64 and "detail output" in p.stderr, p.stderr


# ------------------------------------------------------------------ GPU
COVERAGE_CASES = {  # make_golden.py COVERAGE_CASES
    "cov_index": (["-r", "refs", "--qraw", "rs_A", "--estimate-coverage", "-n", "0.5"],
                  {"KSSD3A_FORCE_REF_INDEX": "1"}),
    "cov_stream": (["-r", "refs", "--qraw", "rs_A", "--estimate-coverage", "-n", "0.5"], {}),
    "cov_single": (["-r", "refs", "--qraw", "rs_asone", "--estimate-coverage", "-n", "0.5"], {}),
    "cov_noidx": (["-r", "refs_noidx", "--qraw", "rs_A", "--estimate-coverage", "-n", "0.5"], {}),
    "cov_noab": (["-r", "refs", "--qraw", "rs_plain", "--estimate-coverage", "-n", "0.5"], {}),
    "cov_unique": (["-r", "refs", "--qraw", "rs_A", "--estimate-coverage", "--unique-ctx", "-n", "0.5"], {}),
    "cov_top1": (["-r", "refs", "--qraw", "rs_A", "--estimate-coverage", "-n", "0.5", "-N", "1"],
                 {"KSSD3A_FORCE_REF_INDEX": "1"}),
    # --asone --conflict -A read sets queried with --qraw --unique-ctx
    # (the C4 combination), on the stream path and the index path
    "asone_unique": (["-r", "refs", "--qraw", "rs_asone", "--unique-ctx", "-n0", "-f0", "-t1"], {}),
    "asone_unique_index": (["-r", "refs", "--qraw", "rs_asone", "--unique-ctx", "-n0", "-f0", "-t1"],
                           {"KSSD3A_FORCE_REF_INDEX": "1"}),
    "A_unique_index": (["-r", "refs", "--qraw", "rs_A", "--unique-ctx", "-n0", "-f0", "-t1"],
                       {"KSSD3A_FORCE_REF_INDEX": "1"}),
}


@pytest.fixture(scope="module")
def coverage_dirs(tmp_path_factory):
    """The sketches the coverage cases read, rebuilt and checked against the
    reference's digests; refs indexed with `sketch -i` (GPU builder)."""
    root = tmp_path_factory.mktemp("cov")
    for name in ("refs", "rs_A", "rs_asone", "rs_plain"):
        sketch(root / name, SKETCHES[name])
    shutil.copytree(root / "refs", root / "refs_noidx")
    p = run_cli(["sketch", "-i", str(root / "refs")])
    assert p.returncode == 0, p.stderr
    for name in ("refs", "refs_noidx", "rs_A", "rs_asone", "rs_plain"):
        check_dir(root / name, name)
    return root


@pytest.mark.gpu
@pytest.mark.parametrize("case", COVERAGE_CASES)
def test_estimate_coverage_matches_reference(coverage_dirs, case):
    args, env = COVERAGE_CASES[case]
    args = [str(coverage_dirs / a) if a in ("refs", "refs_noidx", "rs_A", "rs_asone", "rs_plain") else a
            for a in args]
    p = run_cli(["ani", *args], env=env)
    assert p.returncode == 0, p.stderr
    assert p.stdout == (RS / f"{case}.tsv").read_text()
