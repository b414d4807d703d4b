"""CPU tests of the product's host epilogue (libkssdhost.so: rows, calibrated
model, ranking, formatting) fed with counts from the oracle: the printed
text must be byte-identical to the reference binary's outputs under
tests/golden/ for each output path (index / stream / hash / matrices)."""
import numpy as np
import pytest

import kssdtest as kt

GEN = kt.GOLDEN / "genomes"
COUNT_CASES = sorted((kt.GOLDEN / "counts").glob("case_*"))


@pytest.fixture(scope="module")
def host():
    return kt.Host()


def counts(oracle, ref, qry, flags=0, records=None):
    rec = records if records is not None else (ref.records if ref.records is not None else
                                               kt.build_index_np(ref.keys, ref.index, ref.obj_bits))
    return oracle.count_block(rec, ref.n, ref.obj_bits, qry.keys, qry.index, flags=flags, threads=2)


def ctx_totals(oracle, sk, ignoreconflict=False):
    if not sk.conflict:
        return np.diff(sk.index).astype(np.uint32)
    return np.array([oracle.lib.orc_ctx_runs(kt._p(sk.sample(i)), len(sk.sample(i)), sk.obj_bits,
                                             int(ignoreconflict)) for i in range(sk.n)], np.uint32)


def render(host, path, opts, ref, qry, c, qctx, rctx, header=True):
    ctx, obj, uctx, uobj = c
    out = [kt.DETAIL_HEADER + "\n"] if header else []
    for q in range(qry.n):
        out.append(host.emit(path, opts, ctx[q], obj[q], qctx[q], rctx, qry.names[q], ref.names,
                             uctx=uctx[q] if opts.unique_ctx else None,
                             uobj=uobj[q] if opts.unique_ctx else None))
    return "".join(out)


@pytest.mark.parametrize("golden,kw", [
    *[(f"detail_all_s{s}_index.tsv", dict(s=s, anicut=0.0, afcut=0.0, ctxcut=1)) for s in (2, 5, 6, 7, 8)],
    ("detail_all_s3_index.tsv", dict(s=3, anicut=0.0, afcut=0.0, ctxcut=1)),
    ("detail_all_s9_index.tsv", dict(s=9, anicut=0.0, afcut=0.0, ctxcut=1)),
    ("detail_best_index.tsv", dict()),
    ("detail_qraw_index.tsv", dict(unassembled=True, anicut=0.0, afcut=0.0, ctxcut=1)),
])
def test_index_path_detail(host, oracle, golden, kw):
    ref, qry = kt.read_sketch_dir(GEN / "ref"), kt.read_sketch_dir(GEN / "qry")
    c = counts(oracle, ref, qry)
    opts = kt.ReportOpts.defaults(**kw)
    got = render(host, "index", opts, ref, qry, c, ctx_totals(oracle, qry), ctx_totals(oracle, ref))
    assert got == (GEN / golden).read_text()


@pytest.mark.parametrize("golden,kw", [
    ("detail_best_hash.tsv", dict()),
    ("detail_all_s3_hash.tsv", dict(s=3, anicut=0.0, afcut=0.0, ctxcut=1)),
])
def test_hash_path_detail(host, oracle, golden, kw):
    ref, qry = kt.read_sketch_dir(GEN / "ref"), kt.read_sketch_dir(GEN / "qry")
    c = counts(oracle, ref, qry)
    opts = kt.ReportOpts.defaults(**kw)
    got = render(host, "hash", opts, ref, qry, c, ctx_totals(oracle, qry), ctx_totals(oracle, ref))
    assert got == (GEN / golden).read_text()


@pytest.mark.parametrize("case", COUNT_CASES, ids=lambda p: p.name)
def test_qraw_index_paths_on_conflict_fixtures(host, oracle, case):
    ref, qry = kt.read_sketch_dir(case / "ref"), kt.read_sketch_dir(case / "qry")
    base = dict(unassembled=True, anicut=0.0, afcut=0.0, ctxcut=1)
    c = counts(oracle, ref, qry, flags=kt.FLAGS["UNIQUE"])
    qctx = ctx_totals(oracle, qry)
    rctx = ctx_totals(oracle, ref)
    assert render(host, "index", kt.ReportOpts.defaults(**base), ref, qry, c, qctx, rctx) == \
        (case / "detail_qraw.tsv").read_text()
    assert render(host, "index", kt.ReportOpts.defaults(unique_ctx=True, **base), ref, qry, c, qctx, rctx) == \
        (case / "detail_qraw_unique.tsv").read_text()
    ci = counts(oracle, ref, qry, flags=kt.FLAGS["IGNORECONFLICT"])
    assert render(host, "index", kt.ReportOpts.defaults(ignoreconflict=True, **base), ref, qry, ci, qctx,
                  ctx_totals(oracle, ref, True)) == (case / "detail_qraw_ignoreconflict.tsv").read_text()


def test_calibrated_model_special_cases(host):
    lib = host.lib
    # no shared context: the raw ANI comes back unchanged
    assert lib.kssd_refaf_predict_ani(0.93, 0.7, 0, 0, 0, 0, 8) == 0.93
    # output is a clamped probability-like ANI
    for raw in (0.0, 0.5, 0.9, 0.99, 1.0):
        v = lib.kssd_refaf_predict_ani(raw, 0.6, 500, 100, 120, 10, 8)
        assert 0.0 <= v <= 1.0


def runs_totals(oracle, sk, ignoreconflict=False):
    """Stream paths count distinct context runs whatever the conflict flag
    (command_ani.c:2117-2123, :2905-2916)."""
    return np.array([oracle.lib.orc_ctx_runs(kt._p(sk.sample(i)), len(sk.sample(i)), sk.obj_bits,
                                             int(ignoreconflict)) for i in range(sk.n)], np.uint32)


STREAM_CASES = [
    ("best", dict()),
    ("all_s3", dict(s=3, anicut=0.0, afcut=0.0, ctxcut=1)),
    ("all_s9_top2", dict(s=9, anicut=0.0, afcut=0.0, ctxcut=1, ntop=2)),
    ("qraw", dict(unassembled=True, anicut=0.0, ctxcut=1)),
    ("maxpdist", dict(s=4, anicut=0.0, afcut=0.0, ctxcut=1, max_pdist_set=True, max_pdist=0.06)),
    ("maxsection", dict(s=4, anicut=0.0, afcut=0.0, ctxcut=1, max_diff_obj_section=70)),
]


@pytest.mark.parametrize("tag", ["qry1", "qry3"])
@pytest.mark.parametrize("name,kw", STREAM_CASES, ids=[c[0] for c in STREAM_CASES])
def test_stream_paths_detail(host, oracle, tag, name, kw):
    ref, qry = kt.read_sketch_dir(GEN / "ref"), kt.read_sketch_dir(GEN / tag)
    if kw.get("unassembled"):
        kw = dict(kw, afcut=0.2)  # --qraw without -f (command_ani_wrapper.c:1087-1088)
    c = counts(oracle, ref, qry)
    opts = kt.ReportOpts.defaults(**kw)
    path = "stream" if qry.n == 1 else "multistream"
    got = render(host, path, opts, ref, qry, c, runs_totals(oracle, qry), runs_totals(oracle, ref))
    assert got == (GEN / f"stream_{tag}_{name}.tsv").read_text()


@pytest.mark.parametrize("s", [5, 6, 7, 8, -9])
def test_direct_rect_matrix_cells(host, oracle, s):
    """Direct rectangular matrix (comb_sortedsketch64Xcomb_sortedsketch64,
    command_ani.c:3220-3245): every cell through the host's matrix value."""
    ref, qry = kt.read_sketch_dir(GEN / "ref"), kt.read_sketch_dir(GEN / "qry")
    ctx, obj, _, _ = counts(oracle, ref, qry)
    qctx, rctx = ctx_totals(oracle, qry), ctx_totals(oracle, ref)
    opts = kt.ReportOpts.defaults(fmt=1, s=s)
    lines = ["".join("\t" + n for n in ref.names) + "\n"]
    for q in range(qry.n):
        cells = [host.lib.kssd_matrix_value(opts, int(ctx[q, r, 0]), int(obj[q, r, 0]),
                                            int(obj[q, r, 1]), int(ctx[q, r, 1]),
                                            int(qctx[q]), int(rctx[r]), None, None) for r in range(ref.n)]
        lines.append(qry.names[q] + "".join("\t%f" % v for v in cells) + "\n")
    assert "".join(lines) == (GEN / f"matrix_rq_s{s}.tsv").read_text()


def _fmt(host, fn, *args):
    b = kt.Buf()
    getattr(host.lib, fn)(kt.ctypes.byref(b), *args)
    text = kt.ctypes.string_at(b.s, b.l).decode() if b.l else ""
    host.lib.kssd_buf_free(kt.ctypes.byref(b))
    return text


MATRIX_CMD = [(f"matrix_triangle_{m}.tsv", "triangle", m, 1.0, None) for m in ("ctx-moe", "p_dist", "mash", "aaf")] + \
    [(f"matrix_full_{m}.tsv", "full", m, 1.0, None) for m in ("ctx-moe", "p_dist", "mash", "aaf")] + [
    ("matrix_triangle_default.tsv", "triangle", "ctx-naive", 1.0, None),
    ("matrix_full_default.tsv", "full", "ctx-naive", 1.0, None),
    ("matrix_triangle_diag.tsv", "triangle", "ctx-naive", 2.0, 0.25),
    ("matrix_full_diag.tsv", "full", "aaf", 1.0, 0.25),
]


@pytest.mark.parametrize("golden,fmt,metric,e,diag", MATRIX_CMD, ids=[c[0] for c in MATRIX_CMD])
def test_matrix_command_self(host, oracle, golden, fmt, metric, e, diag):
    """`matrix` on one sketch: the indexed lower-triangle path
    (matrix_write_indexed_self_matrix, command_matrix.c:439-619)."""
    sk = kt.read_sketch_dir(GEN / "all")
    rec = kt.build_index_np(sk.keys, sk.index, sk.obj_bits)
    ctx, obj, _, _ = oracle.count_block(rec, sk.n, sk.obj_bits, sk.keys, sk.index, flags=kt.FLAGS["LOWER"],
                                        threads=2)
    tot = ctx_totals(oracle, sk)
    mm = host.lib.kssd_matrix_metric(metric.encode())
    n = sk.n

    def val(q, r):
        return host.lib.kssd_matrix_context_value(mm, e, int(ctx[q, r, 0]), int(obj[q, r, 0]), int(obj[q, r, 1]),
                                                  int(ctx[q, r, 1]), int(tot[q]), int(tot[r]), 8, 20, 44)
    dv = diag if diag is not None else 0.0
    lines = []
    if fmt == "full":
        lines.append("".join("\t" + x for x in sk.names))
        for q in range(n):
            cells = []
            for r in range(n):
                v = dv if q == r else (val(q, r) if q > r else val(r, q))
                cells.append("\t" + _fmt(host, "kssd_matrix_put_scaled6", host.lib.kssd_matrix_scale_fixed6(v)))
            lines.append(sk.names[q] + "".join(cells))
    else:
        for q in range(n):
            cells = ["\t" + _fmt(host, "kssd_matrix_put_fixed6", val(q, r)) for r in range(q)]
            if diag is not None:
                cells.append("\t" + _fmt(host, "kssd_matrix_put_fixed6", dv))
            lines.append(sk.names[q] + "".join(cells))
    assert "\n".join(lines) + "\n" == (GEN / golden).read_text()


@pytest.mark.parametrize("golden,metric,e", [("matrix_rq_default.tsv", "ctx-naive", 1.0),
                                             *[(f"matrix_rq_{m}.tsv", m, 1.0) for m in ("ctx-moe", "p_dist",
                                                                                       "mash", "aaf")],
                                             ("matrix_rq_e2.tsv", "p_dist", 2.0)])
def test_matrix_command_rect(host, oracle, golden, metric, e):
    """`matrix -r -q`: the dense report (matrix_compute_dense,
    command_matrix.c:258-315, pairwise_eval_arrays per cell)."""
    ref, qry = kt.read_sketch_dir(GEN / "ref"), kt.read_sketch_dir(GEN / "qry")
    ctx, obj, _, _ = counts(oracle, ref, qry)
    qt, rt_ = ctx_totals(oracle, qry), ctx_totals(oracle, ref)
    mm = host.lib.kssd_matrix_metric(metric.encode())
    lines = ["".join("\t" + x for x in ref.names)]
    for q in range(qry.n):
        cells = []
        for r in range(ref.n):
            v = host.lib.kssd_pairwise_value(mm, e, len(qry.sample(q)), len(ref.sample(r)), int(ctx[q, r, 0]),
                                             int(obj[q, r, 0]), int(obj[q, r, 1]), int(ctx[q, r, 1]), int(qt[q]),
                                             int(rt_[r]), 8, 20, 44)
            cells.append("\t%f" % v)
        lines.append(qry.names[q] + "".join(cells))
    assert "\n".join(lines) + "\n" == (GEN / golden).read_text()
