"""CPU tests pinning the oracle to the reference binary's own outputs
(tests/golden/, written by tests/golden/make_golden.py from the reference
compiled out of /root/reference).  Counts must match exactly; FP64 values
must print identically with the reference's %lf / %f formatting."""
import csv
import io
from pathlib import Path

import numpy as np
import pytest

import kssdtest as kt

COUNT_CASES = sorted((kt.GOLDEN / "counts").glob("case_*"))
GEN = kt.GOLDEN / "genomes"


def read_tsv(path):
    return list(csv.DictReader(io.StringIO(Path(path).read_text()), delimiter="\t"))


@pytest.mark.parametrize("case", COUNT_CASES, ids=lambda p: p.name)
def test_counts_against_reference_unique_report(oracle, case):
    ref = kt.read_sketch_dir(case / "ref")
    qry = kt.read_sketch_dir(case / "qry")
    assert ref.records is not None
    ctx, obj, uctx, uobj = oracle.count_block(ref.records, ref.n, ref.obj_bits, qry.keys, qry.index,
                                              flags=kt.FLAGS["UNIQUE"], threads=2)
    rows = read_tsv(case / "unique_report.tsv")
    seen = set()
    for r in rows:
        qi, gi = qry.names.index(r["Qry"]), ref.names.index(r["Ref"])
        seen.add((qi, gi))
        got = dict(XnY_ctx=ctx[qi, gi, 0], N_diff_obj=obj[qi, gi, 0],
                   N_diff_obj_section=obj[qi, gi, 1], N_mut2_ctx=ctx[qi, gi, 1],
                   Unique_XnY_ctx=uctx[qi, gi, 0], Unique_N_diff_obj=uobj[qi, gi, 0],
                   Unique_N_diff_obj_section=uobj[qi, gi, 1], Unique_N_mut2_ctx=uctx[qi, gi, 1])
        for k, v in got.items():
            assert int(r[k]) == int(v), (case.name, r["Qry"], r["Ref"], k)
        # context totals (command_ani.c:383-401 and :351-370)
        assert int(r["Qry_ctx"]) == oracle.lib.orc_ctx_runs(
            kt._p(qry.sample(qi)), len(qry.sample(qi)), qry.obj_bits, 0)
    # the report lists exactly the cells with a shared context (command_ani.c:471)
    assert seen == {tuple(x) for x in np.argwhere(ctx[:, :, 0] > 0)}


@pytest.mark.parametrize("case", COUNT_CASES, ids=lambda p: p.name)
def test_ignoreconflict_counts_against_reference(oracle, case):
    ref = kt.read_sketch_dir(case / "ref")
    qry = kt.read_sketch_dir(case / "qry")
    ctx, obj, _, _ = oracle.count_block(ref.records, ref.n, ref.obj_bits, qry.keys, qry.index,
                                        flags=kt.FLAGS["IGNORECONFLICT"], threads=2)
    for r in read_tsv(case / "detail_qraw_ignoreconflict.tsv"):
        qi, gi = qry.names.index(r["Qry"]), ref.names.index(r["Ref"])
        assert (int(r["XnY_ctx"]), int(r["N_diff_obj"]), int(r["N_diff_obj_section"]),
                int(r["N_mut2_ctx"])) == (ctx[qi, gi, 0], obj[qi, gi, 0], obj[qi, gi, 1], ctx[qi, gi, 1])


def test_genome_counts_index_and_merge_join_agree_with_reference(oracle):
    ref = kt.read_sketch_dir(GEN / "ref")
    qry = kt.read_sketch_dir(GEN / "qry")
    # the reference wrote the inverted index itself (sketch -i)
    assert ref.records.tobytes() == kt.build_index_np(ref.keys, ref.index, ref.obj_bits).tobytes()
    ctx, obj, _, _ = oracle.count_block(ref.records, ref.n, ref.obj_bits, qry.keys, qry.index, threads=2)
    rows = read_tsv(GEN / "detail_all_s3_index.tsv")
    assert len(rows) == int((ctx[:, :, 0] >= 1).sum())
    for r in rows:
        qi, gi = qry.names.index(r["Qry"]), ref.names.index(r["Ref"])
        want = (int(r["XnY_ctx"]), int(r["N_diff_obj"]), int(r["N_diff_obj_section"]), int(r["N_mut2_ctx"]))
        assert want == (ctx[qi, gi, 0], obj[qi, gi, 0], obj[qi, gi, 1], ctx[qi, gi, 1])
        assert want == oracle.pair(qry.sample(qi), ref.sample(gi), ref.obj_bits)


def _fmt(x):
    return "%f" % x


def test_ctxmoe_and_af_values_against_reference(oracle):
    """-s3 detail rows: ANI = 1 - lm3ways distance; AFs and blastn AFs
    (command_ani.c:3489-3494, float-cast in the index path)."""
    ref = kt.read_sketch_dir(GEN / "ref")
    qry = kt.read_sketch_dir(GEN / "qry")
    rows = read_tsv(GEN / "detail_all_s3_index.tsv")
    for r in rows:
        qi, gi = qry.names.index(r["Qry"]), ref.names.index(r["Ref"])
        nq, nr = len(qry.sample(qi)), len(ref.sample(gi))
        f = kt.Features(int(r["XnY_ctx"]), 0, int(r["N_diff_obj_section"]), int(r["N_mut2_ctx"]),
                        int(r["N_diff_obj"]))
        dist = oracle.lib.orc_lm3ways_dist(f, 8)
        assert _fmt(dist) == r["Distance"]
        assert _fmt(1.0 - dist) == r["ANI"]
        afq = np.float32(np.float32(f.XnY_ctx) / np.float32(nq))
        afr = np.float32(np.float32(f.XnY_ctx) / np.float32(nr))
        assert _fmt(float(afq)) == r["Qry_align_fraction"]
        assert _fmt(float(afr)) == r["Ref_align_fraction"]
        f.X_ctx = nq
        bq = float(np.float32(oracle.lib.orc_af_anib(f, 8)))
        f.X_ctx = nr
        br = float(np.float32(oracle.lib.orc_af_anib(f, 8)))
        assert _fmt(min(max(bq, 0.0), 1.0)) == r["blastn_Qry_align_fraction"]
        assert _fmt(min(max(br, 0.0), 1.0)) == r["blastn_Ref_align_fraction"]


def test_pdist_values_against_reference(oracle):
    ref = kt.read_sketch_dir(GEN / "ref")
    for r in read_tsv(GEN / "detail_all_s9_index.tsv"):
        f = kt.Features(int(r["XnY_ctx"]), 0, int(r["N_diff_obj_section"]), int(r["N_mut2_ctx"]),
                        int(r["N_diff_obj"]))
        assert _fmt(oracle.lib.orc_p_dist(f, ref.obj_bits)) == r["Distance"]


def test_naive_values_against_reference(oracle):
    # --qraw forces Naive (command_ani.c:1263-1264)
    for case in COUNT_CASES:
        for r in read_tsv(case / "detail_qraw.tsv"):
            f = kt.Features(int(r["XnY_ctx"]), 0, int(r["N_diff_obj_section"]), int(r["N_mut2_ctx"]),
                            int(r["N_diff_obj"]))
            assert _fmt(oracle.lib.orc_naive_dist(f)) == r["Distance"], (case.name, r)
            assert r["Selected_metric"] == "Naive"


def test_self_matrix_paths_agree(oracle):
    """Direct (merge-join) and indexed (lower-triangle) self matrices are
    byte-identical in the reference (smoke_main.sh:480-492); the oracle's
    lower-triangle counts reproduce the -s3 ANI cells of both."""
    a = (GEN / "self_matrix_direct_s3.tsv").read_text()
    b = (GEN / "self_matrix_indexed_s3.tsv").read_text()
    assert a == b
    s = kt.read_sketch_dir(GEN / "all")
    rec = kt.build_index_np(s.keys, s.index, s.obj_bits)
    ctx, obj, _, _ = oracle.count_block(rec, s.n, s.obj_bits, s.keys, s.index, flags=kt.FLAGS["LOWER"],
                                        threads=2)
    lines = a.splitlines()
    assert lines[0].split("\t")[1:] == s.names
    for qi, line in enumerate(lines[1:]):
        cells = line.split("\t")
        assert cells[0] == s.names[qi]
        for gi in range(qi):
            f = kt.Features(int(ctx[qi, gi, 0]), 0, int(obj[qi, gi, 1]), int(ctx[qi, gi, 1]),
                            int(obj[qi, gi, 0]))
            nq, nr = len(s.sample(qi)), len(s.sample(gi))
            afq, afr = f.XnY_ctx / nq, f.XnY_ctx / nr
            if max(afq, afr) >= 0.5:
                ani = 1.0 - oracle.lib.orc_lm3ways_dist(f, 8)
            else:
                ani = 1.0 - 1.0  # af gate: moe distance is the exception value e=1
            scaled = int(ani * 1e6 + 0.5) if ani >= 0 else int(ani * 1e6 - 0.5)
            assert cells[1 + gi] == "%d.%06d" % divmod(scaled, 1000000), (qi, gi)


SCALE = kt.GOLDEN / "scale"


def _scale_text(stem):
    import gzip
    f = SCALE / f"{stem}.tsv"
    return f.read_text() if f.exists() else gzip.decompress((SCALE / f"{stem}.tsv.gz").read_bytes()).decode()


@pytest.mark.parametrize("stem,ntop", [("p2_index_top3", 3), ("p2_index_top3_s3", 3), ("p2_index_top7_s4", 7),
                                       ("p2_index_top120", 120), ("p2_index_all_s3", 0)])
def test_index_rank_against_reference_order(oracle, stem, ntop):
    """The oracle's restatement of the index path's gate and top-N
    (command_ani.h:577-621) gives, per query, exactly the references the
    reference printed, in its order: 200 references of 40 tied variants,
    so -N 3 / 7 take the heap branch (m > 128, N < m / 2) and -N 120 and
    the full listing the stable sort; with -n0 every ranked row prints."""
    ref = kt.read_sketch_dir(SCALE / "p2" / "ref")
    qry = kt.read_sketch_dir(SCALE / "p2" / "qry")
    ctx, obj, uctx, uobj = oracle.count_block(ref.records, ref.n, ref.obj_bits, qry.keys, qry.index, threads=4)
    ranked = oracle.index_rank(ctx, obj, uctx, uobj, np.diff(ref.index), np.diff(qry.index), ctxcut=1, afcut=0.0,
                               anicut=0.0, ntop=ntop)
    rows = list(csv.DictReader(io.StringIO(_scale_text(stem)), delimiter="\t"))
    for q, (_, _, printed) in enumerate(ranked):
        got = [ref.names[i] for i in printed]
        want = [r["Ref"] for r in rows if r["Qry"] == qry.names[q]]
        assert got == want, (stem, q)
    assert ntop == 0 or all(len(p[2]) == ntop for p in ranked)


@pytest.mark.parametrize("stem,metric,values_ani", [
    ("matrix_ani_rq", 3, 1), ("matrix_rq_s5", 5, 0), ("matrix_rq_s6", 6, 0), ("matrix_rq_s7", 7, 0),
    ("matrix_rq_s8", 8, 0), ("matrix_rq_s-9", 9, 1)])
def test_matrix_values_match_reference(oracle, stem, metric, values_ani):
    """orc_matrix_values (the FP64 matrix epilogue restated from
    command_ani.c:1321-1379, :1458-1496) on the oracle's counts, printed as
    the reference prints matrix cells ("%f"), equals the reference binary's
    `ani -r ref -q qry --format matrix` output byte for byte."""
    g = kt.GOLDEN / "genomes"
    ref, qry = kt.read_sketch_dir(g / "ref"), kt.read_sketch_dir(g / "qry")
    R, Q = ref.n, qry.n
    ctx, obj, _, _ = oracle.count_block(kt.build_index_np(ref.keys, ref.index, 20), R, 20, qry.keys, qry.index)
    qctx = np.diff(qry.index).astype(np.uint32)
    rctx = np.diff(ref.index).astype(np.uint32)
    vals = np.zeros(Q * R, np.float64)
    oracle.lib.orc_matrix_values(kt._p(ctx), kt._p(obj), Q, R, kt._p(qctx), kt._p(rctx), metric, values_ani, 0.5,
                                 1.0, ref.drfold, 20, 44, kt._p(vals), 1)
    v = vals.reshape(Q, R)
    rows = ["".join("\t" + n for n in ref.names)]
    rows += [qry.names[q] + "".join("\t%f" % x for x in v[q]) for q in range(Q)]
    assert "\n".join(rows) + "\n" == (g / f"{stem}.tsv").read_text()
