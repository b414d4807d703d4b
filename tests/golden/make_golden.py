#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference checker
binary (oracle/_ref/kssd3a_ref, compiled from /root/reference by
`make -C oracle ref`).  Run in the container that has /root/reference:

    python tests/golden/make_golden.py

Every fixture is data: input sketch directories (or small FASTA files) and
the reference's outputs on them.  Inputs are generated here from fixed seeds.

  counts/case_XX/      hand-built sketch dirs (test_unique_context.py style,
                       with object conflicts) + the reference's
                       --unique-ctx-report (every count column of every pair
                       with a shared context) and its detail output
  genomes/             8 references + 8 queries simulated from 2 ancestors,
                       sketched by the reference, plus its detail, matrix,
                       triangle and self-matrix outputs (direct and indexed)
  sketch/              small FASTA inputs and the reference's sketches of them
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "tests"))
import kssdtest as kt  # noqa: E402

BIN = ROOT / "oracle" / "_ref" / "kssd3a_ref"


def run(args, env=None, cwd=None):
    e = dict(os.environ, OMP_NUM_THREADS="2")
    e.pop("KSSD3A_FORCE_REF_INDEX", None)
    e.update(env or {})
    p = subprocess.run([str(BIN), *map(str, args)], capture_output=True, text=True, env=e,
                       cwd=cwd, timeout=600)
    if p.returncode != 0:
        raise RuntimeError(f"{args}: {p.stderr}")
    return p.stdout


# ------------------------------------------------------------------ counts
def count_fixture(rng, n_ref, n_qry, ref_entries, qry_entries, ctx_range, obj_range):
    def sample(n):
        pairs = {(int(rng.integers(1, ctx_range)), int(rng.integers(0, obj_range))) for _ in range(n)}
        return sorted(pairs)
    refs = [sample(ref_entries) for _ in range(n_ref)]
    qrys = [sample(qry_entries) for _ in range(n_qry)]
    return refs, qrys


def to_keys(sample, obj_bits=20):
    return np.array([(c << obj_bits) | o for c, o in sample], dtype=np.uint64)


def make_counts():
    out = HERE / "counts"
    shutil.rmtree(out, ignore_errors=True)
    cases = []
    # the reference's randomized oracle shape (test_unique_context.py:268-278)
    for seed in range(8):
        cases.append((seed, 4, 2, 25, 35, 40, 8))
    # larger, more conflicts and more sharing
    for seed in range(8, 12):
        cases.append((seed, 12, 6, 200, 240, 400, 64))
    # ragged: one empty reference, one empty query
    cases.append((12, 5, 4, 60, 60, 120, 16))
    for seed, nr, nq, re_, qe, cr, orr in cases:
        rng = np.random.default_rng(seed)
        refs, qrys = count_fixture(rng, nr, nq, re_, qe, cr, orr)
        if seed == 12:
            refs[1] = []
            qrys[2] = []
        d = out / f"case_{seed:02d}"
        kt.write_sketch_dir(d / "ref", [to_keys(s) for s in refs],
                            names=[f"ref_{i}" for i in range(nr)], indexed=True)
        kt.write_sketch_dir(d / "qry", [to_keys(s) for s in qrys],
                            names=[f"qry_{i}" for i in range(nq)])
        base = ["ani", "-r", d / "ref", "--qraw", d / "qry", "-p2", "-n0", "-f0", "-t1"]
        report = d / "unique_report.tsv"
        report.unlink(missing_ok=True)
        (d / "detail_qraw.tsv").write_text(run(base, env={"KSSD3A_FORCE_REF_INDEX": "1"}))
        run([*base, "--unique-ctx-report", report])
        (d / "detail_qraw_unique.tsv").write_text(run([*base, "--unique-ctx"]))
        (d / "detail_qraw_ignoreconflict.tsv").write_text(
            run([*base, "--ignoreconflict"], env={"KSSD3A_FORCE_REF_INDEX": "1"}))


# ------------------------------------------------------------------ genomes
def mutate(rng, seq, rate):
    s = seq.copy()
    hit = np.flatnonzero(rng.random(len(s)) < rate)
    s[hit] = (s[hit] + rng.integers(1, 4, size=len(hit))) % 4
    return s


def write_fasta(path, name, seq):
    letters = np.frombuffer(b"ACGT", dtype=np.uint8)[seq]
    body = letters.tobytes()
    with open(path, "wb") as fh:
        fh.write(f">{name}\n".encode())
        for i in range(0, len(body), 80):
            fh.write(body[i:i + 80] + b"\n")


def make_genomes():
    out = HERE / "genomes"
    shutil.rmtree(out, ignore_errors=True)
    tmp = out / "_fasta"
    tmp.mkdir(parents=True)
    rng = np.random.default_rng(20261203)
    refs, qrys = [], []
    for a in range(2):
        anc = rng.integers(0, 4, size=120_000, dtype=np.uint8)
        for d in range(4):
            r = mutate(rng, anc, rng.uniform(0.005, 0.05))
            p = tmp / f"ref{a}{d}.fa"
            write_fasta(p, f"ref{a}{d}", r)
            refs.append(p)
            q = mutate(rng, anc, rng.uniform(0.005, 0.05))
            p = tmp / f"qry{a}{d}.fa"
            write_fasta(p, f"qry{a}{d}", q)
            qrys.append(p)
    # names in lcofiles.stat are the paths as given: use relative names
    run(["sketch", "-o", "ref", *[p.relative_to(out) for p in refs]], cwd=out)
    run(["sketch", "-o", "qry", *[p.relative_to(out) for p in qrys]], cwd=out)
    run(["sketch", "-o", "all", *[p.relative_to(out) for p in refs + qrys]], cwd=out)
    R, Q, S = out / "ref", out / "qry", out / "all"
    w = lambda name, text: (out / name).write_text(text)  # noqa: E731
    # hash path (no index yet; ref_n <= 100)
    w("detail_best_hash.tsv", run(["ani", "-r", R, "-q", Q]))
    w("detail_all_s3_hash.tsv", run(["ani", "-r", R, "-q", Q, "-s3", "-n0", "-f0", "-t1"]))
    w("matrix_ani_rq.tsv", run(["ani", "-r", R, "-q", Q, "--format", "matrix", "--values", "ani", "-s3"]))
    w("self_matrix_direct_s3.tsv", run(["ani", "-q", S, "--format", "matrix", "--values", "ani", "-s3"]))
    w("self_matrix_direct_s9.tsv", run(["ani", "-q", S, "--format", "matrix", "-s9"]))
    w("self_matrix_indexed_s3.tsv", run(["ani", "-q", S, "--format", "matrix", "--values", "ani", "-s3"],
                                        env={"KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD": "0"}))
    w("self_triangle_indexed_s4.tsv", run(["ani", "-q", S, "--format", "triangle", "-s4"],
                                          env={"KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD": "0"}))
    w("matrix_triangle_default.tsv", run(["matrix", S]))
    w("matrix_full_default.tsv", run(["matrix", "--format", "full", S]))
    w("matrix_rq_default.tsv", run(["matrix", "-r", R, "-q", Q]))
    # index path: build the inverted index, then the same detail requests
    run(["sketch", "-i", R])
    w("detail_best_index.tsv", run(["ani", "-r", R, "-q", Q]))
    w("detail_all_s3_index.tsv", run(["ani", "-r", R, "-q", Q, "-s3", "-n0", "-f0", "-t1"]))
    w("detail_all_s9_index.tsv", run(["ani", "-r", R, "-q", Q, "-s9", "-n0", "-f0", "-t1"]))
    w("detail_qraw_index.tsv", run(["ani", "-r", R, "--qraw", Q, "-n0", "-f0", "-t1"]))
    shutil.rmtree(tmp)


def subset_dir(src, dst, keep):
    """Sketch dir holding samples `keep` of src (same header fields)."""
    src, dst = Path(src), Path(dst)
    shutil.rmtree(dst, ignore_errors=True)
    dst.mkdir(parents=True)
    stat = (src / "lcofiles.stat").read_bytes()
    head = bytearray(stat[:32])
    head[28:32] = np.uint32(len(keep)).tobytes()
    names = [stat[32 + 256 * i:32 + 256 * (i + 1)] for i in keep]
    (dst / "lcofiles.stat").write_bytes(bytes(head) + b"".join(names))
    keys = np.fromfile(src / "comblco", dtype="<u8")
    index = np.fromfile(src / "comblco.index", dtype="<u8")
    parts = [keys[index[i]:index[i + 1]] for i in keep]
    np.concatenate(parts).astype("<u8").tofile(dst / "comblco")
    np.concatenate([[0], np.cumsum([len(x) for x in parts])]).astype("<u8").tofile(dst / "comblco.index")
    for name, width in (("lcofiles.infilemeta", 32), ("lcofiles.anno", 256)):
        if (src / name).exists():
            raw = (src / name).read_bytes()
            (dst / name).write_bytes(b"".join(raw[width * i:width * (i + 1)] for i in keep))


def make_stream():
    """Few-query requests against the 8-reference set: one query takes the
    single-query streaming path, three take the multi-query path
    (command_ani_wrapper.c:1168-1191)."""
    out = HERE / "genomes"
    R = out / "ref"
    subset_dir(out / "qry", out / "qry1", [5])
    subset_dir(out / "qry", out / "qry3", [0, 3, 6])
    w = lambda name, text: (out / name).write_text(text)  # noqa: E731
    for tag in ("qry1", "qry3"):
        Q = out / tag
        w(f"stream_{tag}_best.tsv", run(["ani", "-r", R, "-q", Q]))
        w(f"stream_{tag}_all_s3.tsv", run(["ani", "-r", R, "-q", Q, "-s3", "-n0", "-f0", "-t1"]))
        w(f"stream_{tag}_all_s9_top2.tsv", run(["ani", "-r", R, "-q", Q, "-s9", "-n0", "-f0", "-t1", "-N2"]))
        w(f"stream_{tag}_qraw.tsv", run(["ani", "-r", R, "--qraw", Q, "-n0", "-t1"]))
        w(f"stream_{tag}_maxpdist.tsv", run(["ani", "-r", R, "-q", Q, "-s4", "-n0", "-f0", "-t1",
                                             "--max-pdist", "0.06"]))
        w(f"stream_{tag}_maxsection.tsv", run(["ani", "-r", R, "-q", Q, "-s4", "-n0", "-f0", "-t1",
                                               "--max-diff-section", "70"]))


def make_metrics():
    """Every selectable metric on the index path (detail) and the direct
    rectangular matrix path, so each distance formula is pinned."""
    out = HERE / "genomes"
    R, Q = out / "ref", out / "qry"
    w = lambda name, text: (out / name).write_text(text)  # noqa: E731
    for s in (2, 5, 6, 7, 8):
        w(f"detail_all_s{s}_index.tsv", run(["ani", "-r", R, "-q", Q, f"-s{s}", "-n0", "-f0", "-t1"]))
    for s in (5, 6, 7, 8, -9):
        w(f"matrix_rq_s{s}.tsv", run(["ani", "-r", R, "-q", Q, "--format", "matrix", f"-s{s}"],
                                    env={"KSSD3A_ANI_MATRIX_DIRECT_THRESHOLD": "1000"}))


def make_matrix_cmd():
    """The `matrix` command on the genome sketches: one-sketch triangle and
    full matrices for every single metric, the rectangular report, the
    exception and diagonal options."""
    out = HERE / "genomes"
    R, Q, S = out / "ref", out / "qry", out / "all"
    w = lambda name, text: (out / name).write_text(text)  # noqa: E731
    for m in ("ctx-moe", "p_dist", "mash", "aaf"):
        w(f"matrix_triangle_{m}.tsv", run(["matrix", "-m", m, S]))
        w(f"matrix_full_{m}.tsv", run(["matrix", "--format", "full", "-m", m, S]))
        w(f"matrix_rq_{m}.tsv", run(["matrix", "-r", R, "-q", Q, "-m", m]))
    w("matrix_triangle_diag.tsv", run(["matrix", "--diagonal-value", "0.25", "-e", "2", S]))
    w("matrix_full_diag.tsv", run(["matrix", "--format", "full", "--diagonal-value", "0.25", "-m", "aaf", S]))
    w("matrix_rq_e2.tsv", run(["matrix", "-r", R, "-q", Q, "-e", "2", "-m", "p_dist"]))
    w("matrix_self_rq.tsv", run(["matrix", "-r", S, "-q", S, "--diagonal-value", "0.5"]))


def make_matrix_edges():
    """`matrix --format edges` (sparse edge reports) with one thread: the
    vote-nominated indexed scan on the indexed genome sketch (metrics,
    expressions, cuts and the index vote options), every pair on the same
    sketch without its index and on a conflict sketch, and -r/-q."""
    out = HERE / "genomes"
    R, Q, S = out / "ref", out / "qry", out / "all"
    w = lambda name, text: (out / name).write_text(text)  # noqa: E731
    cases = {
        "edges_default.tsv": ["--cut", "0.06"],
        "edges_mash.tsv": ["--cut", "0.2", "-m", "mash"],
        "edges_and.tsv": ["--cut", "0.06", "-m", "ctx-naive&aaf"],
        "edges_or.tsv": ["--cut", "0.05", "-m", "ctx-moe|mash"],
        "edges_cuts.tsv": ["--cut", "0.06", "--ctxcut", "150", "--max-afcut", "0.35"],
        "edges_maxfreq.tsv": ["--cut", "0.06", "--index-max-ctx-freq", "3"],
        "edges_votes.tsv": ["--cut", "0.07", "--index-min-votes", "150", "-m", "p_dist"],
        "edges_step.tsv": ["--cut", "0.06", "--index-sample-step", "3", "--index-max-ctx-freq", "5"],
    }
    for name, args in cases.items():
        w(name, run(["matrix", "--format", "edges", "-p", "1", *args, S]))
    # the same sketch without its index: every pair
    tmp = HERE / "_noindex"
    shutil.rmtree(tmp, ignore_errors=True)
    shutil.copytree(S, tmp)
    for f in ("sortedcomb_ctxgid64obj32", "sortedcomb_ctxgid64obj32.unique_bits"):
        (tmp / f).unlink(missing_ok=True)
    w("edges_noindex.tsv", run(["matrix", "--sparse", "-p", "1", "--cut", "0.06", tmp]))
    w("edges_noindex_aaf.tsv", run(["matrix", "--sparse", "-p", "1", "--cut", "0.3", "-m", "aaf|p_dist", tmp]))
    shutil.rmtree(tmp)
    w("edges_rq.tsv", run(["matrix", "--format", "edges", "-p", "1", "--cut", "0.06", "-r", R, "-q", Q]))
    w("edges_rq_same.tsv", run(["matrix", "--format", "edges", "-p", "1", "--cut", "0.05", "-r", S, "-q", S]))
    # a conflict sketch (indexed, but conflicts rule the indexed scan out)
    c = HERE / "counts" / "case_12"
    (c / "edges_conflict.tsv").write_text(run(["matrix", "--format", "edges", "-p", "1", "--cut", "0.9",
                                               "-m", "p_dist", c / "ref"]))


def make_matrix_clusters():
    """`matrix --format clusters` (with one thread): components of the
    indexed scan's edges and of every pair's, their representatives (by the
    input metadata, or by sketch size without it) and --edge-out."""
    out = HERE / "genomes"
    S = out / "all"
    tmp = HERE / "_noindex"
    shutil.rmtree(tmp, ignore_errors=True)
    shutil.copytree(S, tmp)
    for f in ("sortedcomb_ctxgid64obj32", "sortedcomb_ctxgid64obj32.unique_bits", "lcofiles.infilemeta"):
        (tmp / f).unlink(missing_ok=True)
    for name, d, args in (("clusters_default", S, ["--cut", "0.05"]),
                          ("clusters_wide", S, ["--cut", "0.9", "-m", "mash|aaf"]),
                          ("clusters_cuts", S, ["--cut", "0.06", "--ctxcut", "200", "--index-max-ctx-freq", "4"]),
                          ("clusters_noindex", tmp, ["--cut", "0.05"]),
                          ("clusters_noindex_wide", tmp, ["--cut", "0.9", "-m", "ctx-moe&p_dist"])):
        e = tmp.parent / "_edges.tsv"
        (out / f"{name}.tsv").write_text(run(["matrix", "--format", "clusters", "-p", "1", "--edge-out", e,
                                              *args, d]))
        shutil.move(e, out / f"{name}_edges.tsv")
    shutil.rmtree(tmp)


DEDUP_CASES = {  # name: (sketch: "all" or "noindex", arguments), outputs beside the main report
    "dedup_default": ("all", ["--cut", "0.05"]),
    "dedup_full": ("all", ["--cut", "0.06", "--max-afcut", "0.1", "--dedup-strategy", "full-linkage",
                           "--keep-matrix-format", "phylip", "--keep-matrix-idmap", "@idmap"]),
    "dedup_greedy": ("all", ["--cut", "0.06", "--max-afcut", "0.1"]),
    "dedup_wide": ("all", ["--cut", "0.9", "--max-afcut", "0.3", "-m", "mash|aaf"]),
    "dedup_noindex": ("noindex", ["--cut", "0.055", "--max-afcut", "0", "-m", "ctx-moe&p_dist",
                                  "--diagonal-value", "0.5"]),
    "dedup_noindex_full": ("noindex", ["--cut", "0.9", "--dedup-strategy", "clique", "--max-afcut", "0"]),
}
DEDUP_OUTPUTS = {"--keep-out": "keep", "--remove-out": "remove", "--keep-matrix-out": "keep_matrix",
                 "--edge-out": "edges"}


def make_matrix_dedup():
    """`matrix --format dedup-plan` (one thread): greedy and full-linkage
    plans over the indexed scan's edges and every pair's, with --keep-out,
    --remove-out, --edge-out and the kept samples' matrix (TSV, or PHYLIP
    with its id map)."""
    out = HERE / "genomes"
    tmp = HERE / "_noindex"
    shutil.rmtree(tmp, ignore_errors=True)
    shutil.copytree(out / "all", tmp)
    for f in ("sortedcomb_ctxgid64obj32", "sortedcomb_ctxgid64obj32.unique_bits", "lcofiles.infilemeta"):
        (tmp / f).unlink(missing_ok=True)
    for name, (which, args) in DEDUP_CASES.items():
        d = out / "all" if which == "all" else tmp
        extra = []
        for opt, suffix in DEDUP_OUTPUTS.items():
            extra += [opt, HERE / f"_{suffix}.out"]
        args = [str(HERE / "_idmap.out") if a == "@idmap" else a for a in args]
        (out / f"{name}.tsv").write_text(run(["matrix", "--format", "dedup-plan", "-p", "1", *extra, *args, d]))
        for suffix in [*DEDUP_OUTPUTS.values(), "idmap"]:
            f = HERE / f"_{suffix}.out"
            if f.exists():
                shutil.move(f, out / f"{name}_{suffix}.txt")
    shutil.rmtree(tmp)


def make_sketch():
    """Small FASTA inputs (multi-record, lowercase, non-ACGT runs) and the
    reference's sketches of them, at the default -f8 and at -f4."""
    out = HERE / "sketch"
    shutil.rmtree(out, ignore_errors=True)
    out.mkdir(parents=True)
    rng = np.random.default_rng(77)
    files = []
    for i in range(3):
        seq = rng.integers(0, 4, size=40_000, dtype=np.uint8)
        letters = bytearray(np.frombuffer(b"ACGT", dtype=np.uint8)[seq].tobytes())
        if i == 1:  # lowercase stretch and an N run
            letters[1000:3000] = letters[1000:3000].lower()
            letters[5000:5040] = b"N" * 40
        text = b">rec_a\n" + bytes(letters[:25_000]) + b"\n"
        if i != 0:  # a second record restarts the k-mer window
            text += b">rec_b desc\n" + bytes(letters[25_000:]) + b"\nRYKM\n"
        p = out / f"seq{i}.fa"
        p.write_bytes(text)
        files.append(p.name)
    run(["sketch", "-o", "f8", *files], cwd=out)
    run(["sketch", "-f", "4", "-o", "f4", *files], cwd=out)
    run(["sketch", "--conflict", "-o", "f8c", *files], cwd=out)


def write_reads(path, rng, genome, n_reads, read_len=100, err=0.01):
    """Single-end reads of `genome` (both strands) with i.i.d. base errors,
    as 4-line FASTQ with a constant quality line ('@' never starts one)."""
    comp = np.array([3, 2, 1, 0], dtype=np.uint8)
    letters = np.frombuffer(b"ACGT", dtype=np.uint8)
    recs = []
    for r in range(n_reads):
        s = int(rng.integers(0, len(genome) - read_len))
        read = genome[s:s + read_len].copy()
        if rng.random() < 0.5:
            read = comp[read[::-1]]
        hit = rng.random(read_len) < err
        read[hit] = (read[hit] + rng.integers(1, 4, size=int(hit.sum()), dtype=np.uint8)) % 4
        recs.append(b"@r%d\n%s\n+\n%s\n" % (r, letters[read].tobytes(), b"I" * read_len))
    data = b"".join(recs)
    if str(path).endswith(".gz"):
        import gzip
        with gzip.GzipFile(path, "wb", mtime=0) as fh:
            fh.write(data)
    else:
        Path(path).write_bytes(data)


READSET_SKETCHES = {  # output dir: arguments of `sketch` (inputs under readsets/)
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
COVERAGE_CASES = {  # golden: (ani arguments, environment), sketches from READSET_SKETCHES
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


def make_readsets():
    """Read sets (FASTQ, gzip and plain) of two related 20 kbp genomes and
    the reference's sketches of them in the counted modes (-A, --asone, -n,
    --readsQC, --npercentile/--ncap, --sample-name), plus its
    --estimate-coverage detail output against the genomes' sketch."""
    out = HERE / "readsets"
    shutil.rmtree(out, ignore_errors=True)
    out.mkdir(parents=True)
    rng = np.random.default_rng(4242)
    ga = rng.integers(0, 4, size=20_000, dtype=np.uint8)
    gb = mutate(rng, ga, 0.02)
    gc = rng.integers(0, 4, size=15_000, dtype=np.uint8)
    for name, g in (("ga", ga), ("gb", gb), ("gc", gc)):
        write_fasta(out / f"{name}.fa", name, g)
    write_reads(out / "a_1.fq.gz", rng, ga, 1000)
    write_reads(out / "a_2.fq.gz", rng, ga, 1000)
    write_reads(out / "b.fq.gz", rng, gb, 3000)
    write_reads(out / "c.fq", rng, ga, 400)
    for d, args in READSET_SKETCHES.items():
        run(["sketch", "-o", d, *args], cwd=out)
    shutil.copytree(out / "refs", out / "refs_noidx")
    run(["sketch", "-i", "refs"], cwd=out)
    for name, (args, env) in COVERAGE_CASES.items():
        (out / f"{name}.tsv").write_text(run(["ani", *args], env=env, cwd=out))
    # the sketch directories are kept as SHA-256 digests of their files (the
    # tests rebuild them from the inputs above and compare)
    import hashlib
    import json
    digests = {}
    for d in [*READSET_SKETCHES, "refs_noidx"]:
        for f in sorted((out / d).iterdir()):
            if f.name == "lcofiles.stat":  # name slots may hold bytes past the terminator
                sk = kt.read_sketch_dir(out / d)
                digests[f"{d}/{f.name}"] = {"header": f.read_bytes()[:32].hex(), "names": sk.names}
            else:
                digests[f"{d}/{f.name}"] = hashlib.sha256(f.read_bytes()).hexdigest()
        shutil.rmtree(out / d)
    (out / "sketch_digests.json").write_text(json.dumps(digests, indent=1, sort_keys=True) + "\n")


def indel(rng, seq, frac):
    """Insertions and deletions of 1-50 bases (half each) whose lengths add
    up to about `frac` of the sequence."""
    s = seq.copy()
    budget = int(frac * len(s))
    while budget > 0:
        n = int(rng.integers(1, 51))
        at = int(rng.integers(0, max(1, len(s) - n)))
        if rng.random() < 0.5:
            s = np.concatenate([s[:at], rng.integers(0, 4, size=n, dtype=np.uint8), s[at:]])
        else:
            s = np.concatenate([s[:at], s[at + n:]])
        budget -= n
    return s


def sketch_fastas(out, name, seqs, prefix):
    """FASTA files of `seqs` under out/_fasta, sketched by the reference into
    out/name (names relative to out, as make_genomes does)."""
    tmp = out / "_fasta"
    tmp.mkdir(parents=True, exist_ok=True)
    paths = []
    for i, g in enumerate(seqs):
        p = tmp / f"{prefix}{i}.fa"
        write_fasta(p, f"{prefix}{i}", g)
        paths.append(p.relative_to(out))
    run(["sketch", "-o", name, *paths], cwd=out)


FORCE = {"KSSD3A_FORCE_REF_INDEX": "1"}
SCALE_CASES = {  # scale/<name>.tsv: (reference dir, query flag, query dir, arguments, environment)
    # C2 recipe (lognormal lengths, 0.1-10% substitutions + 1% indels), 50 x 50
    # (as many queries as references: the index path when the index exists)
    "c2_best_index": ("c2/ref", "-q", "c2/qry", [], {}),
    "c2_s3_index": ("c2/ref", "-q", "c2/qry", ["-s3", "-n0", "-f0", "-t1"], {}),
    "c2_best_noindex": ("c2/ref_noidx", "-q", "c2/qry", [], {}),
    # C3 recipe: 10 clades of 20 references (members U(1%, 15%) from the
    # ancestor) x 300 new clade members; the index path with its defaults,
    # looser cuts, all pairs, and unique contexts
    "c3_best": ("c3/ref", "-q", "c3/qry", [], {}),
    "c3_s3_loose": ("c3/ref", "-q", "c3/qry", ["-s3", "-n", "0.9", "-f", "0.1"], {}),
    "c3_s4_all": ("c3/ref", "-q", "c3/qry", ["-s4", "-n0", "-f0", "-t1"], {}),
    "c3_qraw_unique": ("c3/ref", "--qraw", "c3/qry", ["--unique-ctx", "-n0", "-f0", "-t1"], {}),
    # many tied ANIs: 200 references (40 variants x 5 copies) all sharing
    # contexts with each of 12 queries.  Fewer queries than references take
    # the multi-query stream path; with the index forced, -N 3 / 7 take the
    # heap branch of idani_select_topN_sorted (m > 128, N < m / 2), -N 120
    # and the full listing its stable sort
    "p2_top3": ("p2/ref", "-q", "p2/qry", ["-N", "3", "-n0", "-f0", "-t1"], {}),
    "p2_top3_s3": ("p2/ref", "-q", "p2/qry", ["-s3", "-N", "3", "-n0", "-f0", "-t1"], {}),
    "p2_index_top3": ("p2/ref", "-q", "p2/qry", ["-N", "3", "-n0", "-f0", "-t1"], FORCE),
    "p2_index_top3_s3": ("p2/ref", "-q", "p2/qry", ["-s3", "-N", "3", "-n0", "-f0", "-t1"], FORCE),
    "p2_index_top7_s4": ("p2/ref", "-q", "p2/qry", ["-s4", "-N", "7", "-n0", "-f0", "-t1"], FORCE),
    "p2_index_top120": ("p2/ref", "-q", "p2/qry", ["-N", "120", "-n0", "-f0", "-t1"], FORCE),
    "p2_index_all_s3": ("p2/ref", "-q", "p2/qry", ["-s3", "-n0", "-f0", "-t1"], FORCE),
    "p2_index_unique_top3": ("p2/ref", "--qraw", "p2/qry", ["--unique-ctx", "-N", "3", "-n0", "-f0", "-t1"], {}),
    # the reference's block boundary (tests/test_unique_context.py:216-223):
    # 4096 one-context references, 310 queries, every third one empty
    "boundary_unique": ("boundary/ref", "--qraw", "boundary/qry", ["--unique-ctx", "-n0", "-f0", "-t1", "-p4"], {}),
    "boundary_detail": ("boundary/ref", "-q", "boundary/qry", ["-n0", "-f0", "-t1"], {}),
    "boundary_index_detail": ("boundary/ref", "-q", "boundary/qry", ["-n0", "-f0", "-t1"], FORCE),
    # the align-fraction bound of the one-query stream path: references that
    # share exactly the needed number of contexts, the last of their walk
    # shared or not (command_ani.c:2154-2156); three queries take the
    # multi-query path, the forced index the index path, neither of which
    # gives a reference up early
    "afb_stream_one": ("afb/ref", "-q", "afb/qry1", ["-n0", "-t1"], {}),
    "afb_stream_one_s3": ("afb/ref", "-q", "afb/qry1", ["-s3", "-n0", "-t1"], {}),
    "afb_stream_one_ignoreconflict": ("afb/ref", "-q", "afb/qry1", ["-n0", "-t1", "--ignoreconflict"], {}),
    "afb_stream_multi": ("afb/ref", "-q", "afb/qry3", ["-n0", "-t1"], {}),
    "afb_index": ("afb/ref", "-q", "afb/qry3", ["-n0", "-t1"], FORCE),
}


def afb_sketches(out):
    """Query contexts 1000..1099; references of 80 contexts (need = 40 at
    afcut 0.5) sharing 40 of them with the last walked context shared or
    not, 41, all, and a 10-context one sharing its last 5; one reference
    whose last context holds two objects (walked or skipped under
    --ignoreconflict)."""
    q = [(c, 0) for c in range(1000, 1100)]
    refs = {
        "r_shared40_last_unshared": [(c, 0) for c in range(1000, 1040)] + [(c, 0) for c in range(2000, 2040)],
        "r_shared40_last_shared": [(c, 0) for c in range(10, 50)] + [(c, 0) for c in range(1050, 1090)],
        "r_shared41": [(c, 0) for c in range(10, 49)] + [(c, 0) for c in range(1050, 1091)],
        "r_shared40_diffobj": [(c, 0) for c in range(10, 50)] + [(c, 3 * (c % 2)) for c in range(1050, 1090)],
        "r_all": list(q),
        "r_small_last_shared": [(c, 0) for c in range(5, 10)] + [(c, 0) for c in range(1095, 1100)],
        "r_conflict_last": [(c, 0) for c in range(10, 49)] + [(c, 0) for c in range(1050, 1090)]
        + [(1090, 0), (1090, 5)],
        "r_none": [(c, 0) for c in range(3000, 3050)],
    }
    kt.write_sketch_dir(out / "afb" / "ref", [to_keys(v) for v in refs.values()], names=list(refs), conflict=True)
    kt.write_sketch_dir(out / "afb" / "qry1", [to_keys(q)], names=["q_all"], conflict=True)
    q2 = [(c, 0) for c in range(1000, 1060)]
    q3 = [(c, 1) for c in range(1040, 1100)]
    kt.write_sketch_dir(out / "afb" / "qry3", [to_keys(q), to_keys(q2), to_keys(q3)], names=["q_all", "q_low", "q_high"],
                        conflict=True)
    run(["sketch", "-i", "ref"], cwd=out / "afb")


def make_scale():
    """Down-scaled slices of the benchmark configurations and the ranking
    and block-boundary cases, sketched and answered by the reference."""
    out = HERE / "scale"
    shutil.rmtree(out, ignore_errors=True)
    out.mkdir(parents=True)
    rng = np.random.default_rng(20261204)
    # C2 slice
    lens = np.clip(rng.lognormal(np.log(1.5e5), 0.5, size=50), 5e4, 4.5e5).astype(int)
    refs = [rng.integers(0, 4, size=n, dtype=np.uint8) for n in lens]
    qrys = [indel(rng, mutate(rng, refs[int(rng.integers(0, 50))], rng.uniform(0.001, 0.10)), 0.01)
            for _ in range(50)]
    (out / "c2").mkdir()
    sketch_fastas(out / "c2", "ref", refs, "r")
    sketch_fastas(out / "c2", "qry", qrys, "q")
    shutil.copytree(out / "c2" / "ref", out / "c2" / "ref_noidx")
    run(["sketch", "-i", "ref"], cwd=out / "c2")
    # C3 slice
    anc = [rng.integers(0, 4, size=int(rng.uniform(4e4, 8e4)), dtype=np.uint8) for _ in range(10)]
    refs = [mutate(rng, anc[c], rng.uniform(0.01, 0.15)) for c in range(10) for _ in range(20)]
    qrys = [mutate(rng, anc[int(rng.integers(0, 10))], rng.uniform(0.01, 0.15)) for _ in range(300)]
    (out / "c3").mkdir()
    sketch_fastas(out / "c3", "ref", refs, "r")
    sketch_fastas(out / "c3", "qry", qrys, "q")
    run(["sketch", "-i", "ref"], cwd=out / "c3")
    # ties and the heap top-N
    a = rng.integers(0, 4, size=50_000, dtype=np.uint8)
    variants = [mutate(rng, a, [0.005, 0.01, 0.02][v % 3]) for v in range(40)]
    order = rng.permutation(200)
    refs = [variants[i % 40] for i in order]
    qrys = [mutate(rng, a, rng.uniform(0.005, 0.03)) for _ in range(12)]
    (out / "p2").mkdir()
    sketch_fastas(out / "p2", "ref", refs, "r")
    sketch_fastas(out / "p2", "qry", qrys, "q")
    run(["sketch", "-i", "ref"], cwd=out / "p2")
    # block boundary (hand-built sketches, as the reference's test builds them)
    kt.write_sketch_dir(out / "boundary" / "ref", [to_keys([(i + 1, 0)]) for i in range(4096)],
                        names=[f"ref_{i}" for i in range(4096)])
    kt.write_sketch_dir(out / "boundary" / "qry", [to_keys([(1, 0)] if i % 3 else []) for i in range(310)],
                        names=[f"qry_{i}" for i in range(310)])
    run(["sketch", "-i", "ref"], cwd=out / "boundary")
    afb_sketches(out)
    import gzip
    import json
    (out / "cases.json").write_text(json.dumps(SCALE_CASES, indent=1, sort_keys=True) + "\n")
    for name, (r, qflag, q, args, env) in SCALE_CASES.items():
        text = run(["ani", "-r", r, qflag, q, *args], env=env, cwd=out).encode()
        if len(text) > 100_000:  # large reports kept compressed
            with gzip.GzipFile(out / f"{name}.tsv.gz", "wb", mtime=0) as fh:
                fh.write(text)
        else:
            (out / f"{name}.tsv").write_bytes(text)
    for d in out.glob("*/_fasta"):
        shutil.rmtree(d)


def main():
    if not BIN.exists():
        raise SystemExit(f"{BIN} missing: run `make -C oracle ref` first")
    if len(sys.argv) > 1:  # named generators only, e.g. `make_golden.py matrix_edges`
        for name in sys.argv[1:]:
            globals()[f"make_{name}"]()
        return
    make_counts()
    make_genomes()
    make_stream()
    make_metrics()
    make_matrix_cmd()
    make_matrix_edges()
    make_matrix_clusters()
    make_matrix_dedup()
    make_sketch()
    make_readsets()
    make_scale()
    print("golden fixtures written under", HERE)


if __name__ == "__main__":
    main()
