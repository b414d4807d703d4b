"""The kssd3a-hip front end against the reference binary's outputs.

CPU tests: argument errors (exit status 64 like the reference's argp
errors), and `sketch -i`, whose index and unique-marker bits must be
byte-identical to the reference's files.  GPU tests: every golden TSV under
tests/golden/ regenerated by `kssd3a-hip ani` with the same arguments and
environment that produced it from the reference (make_golden.py), compared
byte for byte.  Each case runs on a private copy of its sketch directories
because some paths write an index next to the sketch, as the reference does.
"""
import os
import shutil
import subprocess

import pytest

import kssdtest as kt

BIN = kt.PKG / "kssd3a-hip"
GEN = kt.GOLDEN / "genomes"
COUNT_CASES = sorted((kt.GOLDEN / "counts").glob("case_*"))
INDEX_FILES = ("sortedcomb_ctxgid64obj32", "sortedcomb_ctxgid64obj32.unique_bits")


def run_cli(args, env=None, timeout=120):
    if not BIN.exists():
        raise RuntimeError(f"{BIN} missing: build with `make -C kssd3a-public_amd`")
    e = dict(os.environ)
    for k in ("KSSD3A_FORCE_REF_INDEX", "KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD"):
        e.pop(k, None)
    e.update(env or {})
    return subprocess.run([str(BIN), *map(str, args)], capture_output=True, text=True, env=e,
                          timeout=timeout)


def copy_dir(src, dst, with_index=True):
    shutil.copytree(src, dst)
    if not with_index:
        for f in INDEX_FILES:
            (dst / f).unlink(missing_ok=True)
    return dst


# ------------------------------------------------------------------ CPU
@pytest.mark.parametrize("args,msg", [
    (["ani", "-r", "x"], "missing query"),
    (["ani", "-q", "x"], "-r/--ref is required for detail output"),
    (["ani", "-r", "x", "-q", "y", "-s", "0"], "not 0"),
    (["ani", "-r", "x", "-q", "y", "-s", "12"], "-s/--slmetrics"),
    (["ani", "-r", "x", "-q", "y", "-f", "80"], "fraction"),
    (["ani", "-r", "x", "-q", "y", "-m1", "--max-pdist", "0.1"], "detail output"),
    (["ani", "-r", "x", "-q", "y", "--unique-ctx"], "unique-context options"),
    (["ani", "-r", "x", "--qraw", "y", "--unique-ctx", "-s9"], "--unique-ctx uses"),
    (["ani", "--reflist", "a", "--qrylist", "b"], "not offered"),
    (["ani", "-r", "x", "-q", "y", "--metric", "bogus"], "unknown --metric"),
    (["ani", "-r", "x", "-q", "y", "--format", "wide"], "--format must be"),
    (["sketch", "--splitmfa", "-o", "out", "a.fa"], "not offered"),
    (["sketch", "-o", "out"], "no input files"),
    (["sketch"], "give -o OUTDIR"),
])
def test_usage_errors_exit_64(args, msg):
    p = run_cli(args)
    assert p.returncode == 64, p.stderr
    if msg:
        assert msg in p.stderr


@pytest.mark.parametrize("sub,args", [("f8", []), ("f4", ["-f", "4"]), ("f8c", ["--conflict"])])
def test_sketch_matches_reference(tmp_path, sub, args):
    """`sketch -o` on the reference's FASTA inputs: comblco, comblco.index,
    lcofiles.infilemeta and the stat header byte-identical, names equal."""
    sk = kt.GOLDEN / "sketch"
    out = tmp_path / sub
    p = subprocess.run([str(BIN), "sketch", *args, "-o", str(out), "seq0.fa", "seq1.fa", "seq2.fa"],
                       capture_output=True, text=True, cwd=sk, timeout=120)
    assert p.returncode == 0, p.stderr
    for f in ("comblco", "comblco.index", "lcofiles.infilemeta"):
        assert (out / f).read_bytes() == (sk / sub / f).read_bytes(), f
    assert (out / "lcofiles.stat").read_bytes()[:32] == (sk / sub / "lcofiles.stat").read_bytes()[:32]
    assert kt.read_sketch_dir(out).names == kt.read_sketch_dir(sk / sub).names


def test_sketch_long_lines_and_gzip(tmp_path):
    """Lines longer than the reader's buffer (a one-line genome, a long
    header) and gzip input sketch like the wrapped plain file."""
    import gzip
    import numpy as np
    rng = np.random.default_rng(4)
    seq = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 300_000)].tobytes()
    wrapped = b">a\n" + b"\n".join(seq[i:i + 80] for i in range(0, len(seq), 80)) + b"\n"
    (tmp_path / "w.fa").write_bytes(wrapped)
    (tmp_path / "one.fa").write_bytes(b">a " + b"x" * 100_000 + b"\n" + seq + b"\n")
    (tmp_path / "z.fa.gz").write_bytes(gzip.compress(wrapped))
    p = run_cli(["sketch", "-o", tmp_path / "out", tmp_path / "w.fa", tmp_path / "one.fa", tmp_path / "z.fa.gz"])
    assert p.returncode == 0, p.stderr
    d = kt.read_sketch_dir(tmp_path / "out")
    assert len(d.sample(0)) > 500
    assert np.array_equal(d.sample(0), d.sample(1)) and np.array_equal(d.sample(0), d.sample(2))
    meta = (tmp_path / "out" / "lcofiles.infilemeta").read_bytes()
    assert meta[0:32][:16] == meta[32:64][:16]  # same record lengths summary


def test_sketch_plain_pieces_match_line_reader(tmp_path):
    """Plain files are mapped and sketched in pieces of about 1 MB by several
    threads (FASTA records cut with a 31-base overlap, FASTQ files in
    line-aligned chunks); gzip files go through the line reader.  The same
    content both ways must give the same keys and counts: a 3 Mbp two-record
    FASTA with 61-base lines, CRLF line ends, lowercase and N runs around the
    piece boundaries, sequence before the first header; a 2.5 MB FASTQ; and
    a 12 MB FASTQ with leading empty lines, CRLF records and a 1.5 Mbp read."""
    import gzip
    import numpy as np
    rng = np.random.default_rng(11)
    letters = np.frombuffer(b"ACGTacgt", np.uint8)
    seq = bytearray(letters[rng.integers(0, 8, 3_000_000)].tobytes())
    for at in (1 << 20, 2 << 20, (1 << 20) + 45, 2_000_007):
        seq[at - 3:at + 2] = b"NNNNN"
    lines = [bytes(seq[i:i + 61]) for i in range(0, len(seq), 61)]
    body = b"".join(ln + (b"\r\n" if k % 7 == 0 else b"\n") for k, ln in enumerate(lines))
    fa = b"ACGTTGCAAGGCTTACGATCGATCGGATCCATGCA\n>r1 first\n" + body[:1_700_000] + b"\n>r2\n" + body[1_700_000:]
    reads = []
    for r in range(8000):
        s0 = int(rng.integers(0, len(seq) - 150))
        reads.append(b"@q%d\n%s\n+\n%s\n" % (r, bytes(seq[s0:s0 + 150]).upper(), b"I" * 150))
    fq = b"".join(reads)
    # over 8 MB: cut into line-aligned chunks by all threads; leading empty
    # lines, CRLF records, a 1.5 Mbp read spanning chunk cuts
    big = [b"\n\r\n"]
    for r in range(30000):
        s0 = int(rng.integers(0, len(seq) - 150))
        end = b"\r\n" if r % 5 == 0 else b"\n"
        big.append(b"@b%d%s%s%s+%s%s%s" % (r, end, bytes(seq[s0:s0 + 150]).upper(), end, end, b"I" * 150, end))
        if r == 3000:
            big.append(b"@long\n%s\n+\n%s\n" % (bytes(seq[:1_500_000]).upper(), b"I" * 1_500_000))
    fqb = b"".join(big)
    assert len(fqb) > 8 << 20
    for name, data in (("g.fa", fa), ("r.fq", fq), ("b.fq", fqb)):
        (tmp_path / name).write_bytes(data)
        (tmp_path / (name + ".gz")).write_bytes(gzip.compress(data, 1))
    for args in (["-f", "4", "g.fa", "g.fa.gz"], ["-f", "2", "-A", "--conflict", "r.fq", "r.fq.gz"],
                 ["-f", "2", "-A", "--conflict", "b.fq", "b.fq.gz"]):
        out = tmp_path / ("o" + args[2])
        p = subprocess.run([str(BIN), "sketch", "-p", "8", "-o", str(out), *args], capture_output=True, text=True,
                           cwd=tmp_path, timeout=120)
        assert p.returncode == 0, p.stderr
        d = kt.read_sketch_dir(out)
        assert len(d.sample(0)) > 1000
        assert np.array_equal(d.sample(0), d.sample(1)), args
        if "-A" in args:
            a = np.fromfile(out / "comblco.a", dtype=np.uint32)
            n0 = int(d.index[1])
            assert np.array_equal(a[:n0], a[n0:])


# sketch -i builds the index on the GPU by default; --host-index on the host
BUILDERS = [pytest.param([], marks=pytest.mark.gpu, id="gpu"), pytest.param(["--host-index"], id="host")]


@pytest.mark.parametrize("builder", BUILDERS)
def test_sketch_index_matches_reference(tmp_path, builder):
    d = copy_dir(GEN / "ref", tmp_path / "ref", with_index=False)
    p = run_cli(["sketch", "-i", d, "-p", "2", *builder])
    assert p.returncode == 0, p.stderr
    for f in INDEX_FILES:
        assert (d / f).read_bytes() == (GEN / "ref" / f).read_bytes(), f


@pytest.mark.parametrize("builder", BUILDERS)
def test_sketch_index_conflict_sketch(tmp_path, builder):
    """Conflict-keeping sketch (object conflicts inside contexts): the index
    and unique bits against the numpy restatement of the format."""
    import numpy as np
    case = COUNT_CASES[-1]
    d = copy_dir(case / "ref", tmp_path / "ref", with_index=False)
    p = run_cli(["sketch", "-i", d, *builder])
    assert p.returncode == 0, p.stderr
    sk = kt.read_sketch_dir(d)
    want = kt.build_index_np(sk.keys, sk.index, sk.obj_bits)
    assert (d / INDEX_FILES[0]).read_bytes() == want.tobytes()
    ctx = want["ctxgid"] >> np.uint64(20)
    bits = np.zeros(len(want), bool)
    for i in range(len(want)):
        grp = want["obj"][ctx == ctx[i]]
        bits[i] = int((grp == want["obj"][i]).sum()) == 1
    assert (d / INDEX_FILES[1]).read_bytes() == np.packbits(bits, bitorder="little").tobytes()


# ------------------------------------------------------------------ GPU
def genome_dirs(tmp_path, ref_index):
    return {
        "R": copy_dir(GEN / "ref", tmp_path / "ref", with_index=ref_index),
        "Q": copy_dir(GEN / "qry", tmp_path / "qry"),
        "S": copy_dir(GEN / "all", tmp_path / "all", with_index=False),
        "Q1": copy_dir(GEN / "qry1", tmp_path / "qry1"),
        "Q3": copy_dir(GEN / "qry3", tmp_path / "qry3"),
    }


T0 = ["-n0", "-f0", "-t1"]
GENOME_CASES = [
    # (golden, reference index present, args, env)
    ("detail_best_hash.tsv", False, ["-r", "R", "-q", "Q"], {}),
    ("detail_all_s3_hash.tsv", False, ["-r", "R", "-q", "Q", "-s3", *T0], {}),
    ("matrix_ani_rq.tsv", False, ["-r", "R", "-q", "Q", "--format", "matrix", "--values", "ani", "-s3"], {}),
    ("self_matrix_direct_s3.tsv", False, ["-q", "S", "--format", "matrix", "--values", "ani", "-s3"], {}),
    ("self_matrix_direct_s9.tsv", False, ["-q", "S", "--format", "matrix", "-s9"], {}),
    ("self_matrix_direct_s9.tsv", False, ["-r", "S", "-q", "S", "--format", "matrix", "-s9"], {}),
    ("self_matrix_indexed_s3.tsv", False, ["-q", "S", "--format", "matrix", "--values", "ani", "-s3"],
     {"KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD": "0"}),
    ("self_triangle_indexed_s4.tsv", False, ["-q", "S", "--format", "triangle", "-s4"],
     {"KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD": "0"}),
    ("detail_best_index.tsv", True, ["-r", "R", "-q", "Q"], {}),
    ("detail_all_s3_index.tsv", True, ["-r", "R", "-q", "Q", "-s3", *T0], {}),
    ("detail_all_s9_index.tsv", True, ["-r", "R", "-q", "Q", "-s9", *T0], {}),
    ("detail_qraw_index.tsv", True, ["-r", "R", "--qraw", "Q", *T0], {}),
]
GENOME_CASES += [(f"detail_all_s{s}_index.tsv", True, ["-r", "R", "-q", "Q", f"-s{s}", *T0], {})
                 for s in (2, 5, 6, 7, 8)]
GENOME_CASES += [(f"matrix_rq_s{s}.tsv", True, ["-r", "R", "-q", "Q", "--format", "matrix", f"-s{s}"], {})
                 for s in (5, 6, 7, 8, -9)]
for tag in ("1", "3"):
    GENOME_CASES += [
        (f"stream_qry{tag}_best.tsv", True, ["-r", "R", "-q", f"Q{tag}"], {}),
        (f"stream_qry{tag}_all_s3.tsv", False, ["-r", "R", "-q", f"Q{tag}", "-s3", *T0], {}),
        (f"stream_qry{tag}_all_s9_top2.tsv", True, ["-r", "R", "-q", f"Q{tag}", "-s9", *T0, "-N2"], {}),
        (f"stream_qry{tag}_qraw.tsv", True, ["-r", "R", "--qraw", f"Q{tag}", "-n0", "-t1"], {}),
        (f"stream_qry{tag}_maxpdist.tsv", True,
         ["-r", "R", "-q", f"Q{tag}", "-s4", *T0, "--max-pdist", "0.06"], {}),
        (f"stream_qry{tag}_maxsection.tsv", True,
         ["-r", "R", "-q", f"Q{tag}", "-s4", *T0, "--max-diff-section", "70"], {}),
    ]


@pytest.mark.gpu
@pytest.mark.parametrize("golden,ref_index,args,env", GENOME_CASES,
                         ids=[f"{c[0]}:{' '.join(c[2])}" for c in GENOME_CASES])
def test_cli_genomes_match_reference(tmp_path, golden, ref_index, args, env):
    dirs = genome_dirs(tmp_path, ref_index)
    argv = [str(dirs.get(a, a)) for a in args]
    out = tmp_path / "out.tsv"
    p = run_cli(["ani", *argv, "-o", out], env=env)
    assert p.returncode == 0, p.stderr
    assert out.read_text() == (GEN / golden).read_text()


SCALE = kt.GOLDEN / "scale"


def _golden_text(stem):
    import gzip
    f = SCALE / f"{stem}.tsv"
    return f.read_text() if f.exists() else gzip.decompress((SCALE / f"{stem}.tsv.gz").read_bytes()).decode()


SCALE_CASES = __import__("json").loads((SCALE / "cases.json").read_text())


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SCALE_CASES))
def test_cli_scale_cases_match_reference(tmp_path, name):
    """Down-scaled C2 and C3 workloads, the heap top-N with ties and the
    block boundary with empty queries, each byte-identical to the reference's
    report (tests/golden/make_golden.py, make_scale)."""
    ref, qflag, qry, args, env = SCALE_CASES[name]
    top = ref.split("/")[0]
    shutil.copytree(SCALE / top, tmp_path / top)
    e = dict(os.environ)
    e.pop("KSSD3A_FORCE_REF_INDEX", None)
    e.update(env)
    p = subprocess.run([str(BIN), "ani", "-r", ref, qflag, qry, *args], capture_output=True, text=True,
                       cwd=tmp_path, env=e, timeout=300)
    assert p.returncode == 0, p.stderr
    assert p.stdout == _golden_text(name)


@pytest.mark.gpu
def test_cli_stdout_and_auto_index(tmp_path):
    """Without -o the rows go to stdout; the self full matrix above the
    direct threshold writes the index next to the sketch first."""
    dirs = genome_dirs(tmp_path, False)
    p = run_cli(["ani", "-q", dirs["S"], "--format", "matrix", "--values", "ani", "-s3"],
                env={"KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD": "0"})
    assert p.returncode == 0, p.stderr
    assert p.stdout == (GEN / "self_matrix_indexed_s3.tsv").read_text()
    assert (dirs["S"] / INDEX_FILES[0]).read_bytes() == (GEN / "all" / INDEX_FILES[0]).read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("case", COUNT_CASES, ids=lambda p: p.name)
def test_cli_conflict_fixtures_match_reference(tmp_path, case):
    ref = copy_dir(case / "ref", tmp_path / "ref")
    qry = copy_dir(case / "qry", tmp_path / "qry")
    base = ["ani", "-r", ref, "--qraw", qry, "-p2", "-n0", "-f0", "-t1"]
    force = {"KSSD3A_FORCE_REF_INDEX": "1"}
    for extra, env, golden in (([], force, "detail_qraw.tsv"),
                               (["--unique-ctx"], {}, "detail_qraw_unique.tsv"),
                               (["--ignoreconflict"], force, "detail_qraw_ignoreconflict.tsv")):
        p = run_cli([*base, *extra], env=env)
        assert p.returncode == 0, p.stderr
        assert p.stdout == (case / golden).read_text(), golden


@pytest.mark.parametrize("builder", BUILDERS)
def test_sketch_index_unique_bits_large_groups(tmp_path, builder):
    """Context groups larger than the small-group scan (many samples share
    one context, some with the same object)."""
    import numpy as np
    rng = np.random.default_rng(5)
    samples = []
    for i in range(60):
        ctxs = np.array([7, 9, 11 + (i % 3)], np.uint64)
        objs = np.array([rng.integers(0, 40), i, rng.integers(0, 4)], np.uint64)
        samples.append(np.sort((ctxs << np.uint64(20)) | objs))
    kt.write_sketch_dir(tmp_path / "big", samples)
    p = run_cli(["sketch", "-i", tmp_path / "big", *builder])
    assert p.returncode == 0, p.stderr
    sk = kt.read_sketch_dir(tmp_path / "big")
    want = kt.build_index_np(sk.keys, sk.index, sk.obj_bits)
    assert (tmp_path / "big" / INDEX_FILES[0]).read_bytes() == want.tobytes()
    ctx = want["ctxgid"] >> np.uint64(20)
    bits = np.array([int((want["obj"][ctx == ctx[i]] == want["obj"][i]).sum()) == 1
                     for i in range(len(want))])
    assert bits.any() and not bits.all()
    assert (tmp_path / "big" / INDEX_FILES[1]).read_bytes() == np.packbits(bits, bitorder="little").tobytes()


MATRIX_CASES = [("matrix_triangle_default.tsv", ["S"]), ("matrix_full_default.tsv", ["--format", "full", "S"]),
                ("matrix_rq_default.tsv", ["-r", "R", "-q", "Q"]),
                ("matrix_triangle_diag.tsv", ["--diagonal-value", "0.25", "-e", "2", "S"]),
                ("matrix_full_diag.tsv", ["--format", "full", "--diagonal-value", "0.25", "-m", "aaf", "S"]),
                ("matrix_rq_e2.tsv", ["-r", "R", "-q", "Q", "-e", "2", "-m", "p_dist"]),
                ("matrix_self_rq.tsv", ["-r", "S", "-q", "S", "--diagonal-value", "0.5"])]
for _m in ("ctx-moe", "p_dist", "mash", "aaf"):
    MATRIX_CASES += [(f"matrix_triangle_{_m}.tsv", ["-m", _m, "S"]),
                     (f"matrix_full_{_m}.tsv", ["--format", "full", "-m", _m, "S"]),
                     (f"matrix_rq_{_m}.tsv", ["-r", "R", "-q", "Q", "-m", _m])]


@pytest.mark.gpu
@pytest.mark.parametrize("golden,args", MATRIX_CASES, ids=[c[0] for c in MATRIX_CASES])
def test_cli_matrix_command_matches_reference(tmp_path, golden, args):
    dirs = genome_dirs(tmp_path, True)
    # the reference ran these with the index built next to "all" (the indexed ani runs came first)
    shutil.copy(GEN / "all" / INDEX_FILES[0], dirs["S"] / INDEX_FILES[0])
    argv = [str(dirs.get(x, x)) for x in args]
    p = run_cli(["matrix", *argv])
    assert p.returncode == 0, p.stderr
    assert p.stdout == (GEN / golden).read_text()


EDGE = ["--format", "edges"]
EDGE_CASES = [  # (golden, args, the sketch keeps its index)
    ("edges_default.tsv", [*EDGE, "--cut", "0.06", "S"], True),
    ("edges_mash.tsv", [*EDGE, "--cut", "0.2", "-m", "mash", "S"], True),
    ("edges_and.tsv", [*EDGE, "--cut", "0.06", "-m", "ctx-naive&aaf", "S"], True),
    ("edges_or.tsv", [*EDGE, "--cut", "0.05", "-m", "ctx-moe|mash", "S"], True),
    ("edges_cuts.tsv", [*EDGE, "--cut", "0.06", "--ctxcut", "150", "--max-afcut", "0.35", "S"], True),
    ("edges_maxfreq.tsv", [*EDGE, "--cut", "0.06", "--index-max-ctx-freq", "3", "S"], True),
    ("edges_votes.tsv", [*EDGE, "--cut", "0.07", "--index-min-votes", "150", "-m", "p_dist", "S"], True),
    ("edges_step.tsv", [*EDGE, "--cut", "0.06", "--index-sample-step", "3", "--index-max-ctx-freq", "5", "S"], True),
    ("edges_noindex.tsv", ["--sparse", "--cut", "0.06", "S"], False),
    ("edges_noindex_aaf.tsv", ["--sparse", "--cut", "0.3", "-m", "aaf|p_dist", "S"], False),
    ("edges_rq.tsv", [*EDGE, "--cut", "0.06", "-r", "R", "-q", "Q"], False),
    ("edges_rq_same.tsv", [*EDGE, "--cut", "0.05", "-r", "S", "-q", "S"], True),
]


@pytest.mark.gpu
@pytest.mark.parametrize("golden,args,indexed", EDGE_CASES, ids=[c[0] for c in EDGE_CASES])
def test_cli_matrix_edges_match_reference(tmp_path, golden, args, indexed):
    """Sparse edge reports (the reference ran them with -p 1): the
    vote-nominated indexed scan on the indexed sketch, every pair without
    the index, and -r/-q."""
    dirs = genome_dirs(tmp_path, True)
    if indexed:
        for f in INDEX_FILES:
            shutil.copy(GEN / "all" / f, dirs["S"] / f)
    p = run_cli(["matrix", "-p", "4", *[str(dirs.get(x, x)) for x in args]])
    assert p.returncode == 0, p.stderr
    assert p.stdout == (GEN / golden).read_text()


CLUSTER_CASES = [  # (golden stem, args, the sketch keeps its index and metadata)
    ("clusters_default", ["--cut", "0.05"], True),
    ("clusters_wide", ["--cut", "0.9", "-m", "mash|aaf"], True),
    ("clusters_cuts", ["--cut", "0.06", "--ctxcut", "200", "--index-max-ctx-freq", "4"], True),
    ("clusters_noindex", ["--cut", "0.05"], False),
    ("clusters_noindex_wide", ["--cut", "0.9", "-m", "ctx-moe&p_dist"], False),
]


@pytest.mark.gpu
@pytest.mark.parametrize("stem,args,indexed", CLUSTER_CASES, ids=[c[0] for c in CLUSTER_CASES])
def test_cli_matrix_clusters_match_reference(tmp_path, stem, args, indexed):
    """Cluster reports and their --edge-out edges: components of the
    indexed scan's edges (representatives by input metadata) and of every
    pair's (no index, no metadata: by sketch size)."""
    dirs = genome_dirs(tmp_path, True)
    if indexed:
        for f in INDEX_FILES:
            shutil.copy(GEN / "all" / f, dirs["S"] / f)
    else:
        (dirs["S"] / "lcofiles.infilemeta").unlink(missing_ok=True)
    edges = tmp_path / "edges.tsv"
    p = run_cli(["matrix", "--format", "clusters", "--edge-out", edges, *args, dirs["S"]])
    assert p.returncode == 0, p.stderr
    assert p.stdout == (GEN / f"{stem}.tsv").read_text()
    assert edges.read_text() == (GEN / f"{stem}_edges.tsv").read_text()


DEDUP_CASES = {  # make_golden.py DEDUP_CASES: (sketch keeps its index and metadata, args)
    "dedup_default": (True, ["--cut", "0.05"]),
    "dedup_full": (True, ["--cut", "0.06", "--max-afcut", "0.1", "--dedup-strategy", "full-linkage",
                          "--keep-matrix-format", "phylip", "--keep-matrix-idmap", "@idmap"]),
    "dedup_greedy": (True, ["--cut", "0.06", "--max-afcut", "0.1"]),
    "dedup_wide": (True, ["--cut", "0.9", "--max-afcut", "0.3", "-m", "mash|aaf"]),
    "dedup_noindex": (False, ["--cut", "0.055", "--max-afcut", "0", "-m", "ctx-moe&p_dist",
                              "--diagonal-value", "0.5"]),
    "dedup_noindex_full": (False, ["--cut", "0.9", "--dedup-strategy", "clique", "--max-afcut", "0"]),
}
DEDUP_OUTPUTS = {"--keep-out": "keep", "--remove-out": "remove", "--keep-matrix-out": "keep_matrix",
                 "--edge-out": "edges"}


@pytest.mark.gpu
@pytest.mark.parametrize("stem", DEDUP_CASES)
def test_cli_matrix_dedup_plan_matches_reference(tmp_path, stem):
    """Dedup plans (greedy and full-linkage, over the indexed scan's edges
    and every pair's) with every side output: keep and remove lists, the
    kept samples' matrix (TSV; PHYLIP with its id map), --edge-out."""
    indexed, args = DEDUP_CASES[stem]
    dirs = genome_dirs(tmp_path, True)
    if indexed:
        for f in INDEX_FILES:
            shutil.copy(GEN / "all" / f, dirs["S"] / f)
    else:
        (dirs["S"] / "lcofiles.infilemeta").unlink(missing_ok=True)
    extra = []
    for opt, suffix in DEDUP_OUTPUTS.items():
        extra += [opt, tmp_path / f"{suffix}.out"]
    args = [str(tmp_path / "idmap.out") if a == "@idmap" else a for a in args]
    p = run_cli(["matrix", "--format", "dedup-plan", *extra, *args, dirs["S"]])
    assert p.returncode == 0, p.stderr
    assert p.stdout == (GEN / f"{stem}.tsv").read_text()
    for suffix in [*DEDUP_OUTPUTS.values(), "idmap"]:
        want = GEN / f"{stem}_{suffix}.txt"
        got = tmp_path / f"{suffix}.out"
        assert got.exists() == want.exists(), suffix
        if want.exists():
            assert got.read_text() == want.read_text(), suffix


@pytest.mark.gpu
def test_cli_matrix_edges_conflict_sketch(tmp_path):
    """A conflict sketch with an index still takes the every-pair path."""
    case = COUNT_CASES[-1]
    d = copy_dir(case / "ref", tmp_path / "ref")
    p = run_cli(["matrix", "--format", "edges", "--cut", "0.9", "-m", "p_dist", d])
    assert p.returncode == 0, p.stderr
    assert p.stdout == (case / "edges_conflict.tsv").read_text()


def test_matrix_usage_errors():
    for args, msg in ((["matrix", "-r", "x", "--format", "triangle", "y"], "only valid for one sketch"),
                      (["matrix", "-m", "aaf&mash", "x"], "combined --metric"),
                      (["matrix", "-m", "aaf&mash|p_dist", "--sparse", "--cut", "1", "x"], "-m/--metric must be"),
                      (["matrix", "-m", "bogus", "x"], "-m/--metric must be"),
                      (["matrix", "--sparse", "x"], "requires --cut"),
                      (["matrix", "--format", "dedup-plan", "-r", "x", "-q", "y", "--cut", "1"], "requires one sketch"),
                      (["matrix", "--format", "clusters", "--cut", "0.1", "-r", "x", "-q", "y"], "requires one sketch"),
                      (["matrix", "--keep-out", "k.txt", "x"], "require --format dedup-plan"),
                      (["matrix", "--dedup-strategy", "greedy", "x"], "requires --format dedup-plan"),
                      (["matrix", "--format", "dedup", "--cut", "1", "--dedup-strategy", "x", "y"], "--dedup-strategy must"),
                      (["matrix", "--format", "dedup", "--cut", "1", "--keep-matrix-format", "phylip", "y"],
                       "requires --keep-matrix-out"),
                      (["matrix", "--format", "dedup", "--cut", "1", "--keep-matrix-out", "m", "--keep-matrix-idmap",
                        "i", "y"], "requires --keep-matrix-format phylip"),
                      (["matrix", "--format", "dedup", "--cut", "1", "--keep-out", "a", "--remove-out", "a", "y"],
                       "must be different files"),
                      (["matrix", "--format", "dedup", "--cut", "1", "--keep-matrix-out", "a", "-o", "a", "y"],
                       "--keep-matrix-out must differ from --outfile"),
                      (["matrix", "--glist", "g", "x"], "not offered"),
                      (["matrix"], "missing sketch input")):
        p = run_cli(args)
        assert p.returncode == 64 and msg in p.stderr, (args, p.stderr)
