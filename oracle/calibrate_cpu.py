#!/usr/bin/env python3
"""CPU calibration of the oracle against the reference binary.

Times, on the same synthetic sketches and the same host threads, the
reference checker (oracle/_ref/kssd3a_ref, built from /root/reference by
`make -C oracle ref`) on its indexed matrix path
(KSSD3A_FORCE_REF_INDEX=1 ani --format matrix --values ani -s3, i.e.
mem_eff_sorted_ctxgidobj_arrXcomb_sortedsketch64, command_ani.c:494-692)
and the oracle's restatement of the same block loop + FP64 matrix values
(orc_count_block + orc_matrix_values).  The ratio t_oracle / t_reference
lets the oracle's timings on the GPU box be quoted as reference-equivalent.
The reference time includes reading the sketch and index files and writing
the TSV (its command does that); the oracle's excludes them, so the ratio
is a lower bound on the reference's extra cost.

Also checks that both produce the same matrix bytes (the oracle's values
formatted as the reference's "%lf" cells).  Run in the container that has
/root/reference:

    python oracle/calibrate_cpu.py [--threads 8]
"""
import argparse
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "tests"))
import kssdtest as kt  # noqa: E402

BIN = ROOT / "oracle" / "_ref" / "kssd3a_ref"


def sketches(host, nref, nqry, seed, threads, clades=False):
    rng = np.random.default_rng(seed)
    if clades:
        nclade = max(1, nref // 20)
        clen = np.clip(rng.lognormal(np.log(4e6), 0.5, size=nclade), 1e6, 12e6).astype(np.uint64)
        ref = [kt.SimSpec(seed * 1000 + c % nclade + 1, int(clen[c % nclade]), seed * 7919 + i + 1,
                          float(rng.uniform(0.01, 0.15)), 0.0, 0, 0.0)
               for i, c in enumerate(range(nref))]
        qry = [kt.SimSpec(seed * 1000 + int(c) + 1, int(clen[int(c)]), seed * 104729 + j + 1,
                          float(rng.uniform(0.01, 0.15)), 0.0, 0, 0.0)
               for j, c in enumerate(rng.integers(0, nclade, size=nqry))]
    else:
        lens = np.clip(rng.lognormal(np.log(4e6), 0.5, size=nref), 1e6, 12e6).astype(np.uint64)
        ref = [kt.SimSpec(seed * 1000 + i + 1, int(lens[i]), 0, 0.0, 0.0, 0, 0.0) for i in range(nref)]
        par = rng.integers(0, nref, size=nqry)
        qry = [kt.SimSpec(ref[p].anc_seed, ref[p].anc_len, seed * 104729 + j + 1,
                          float(rng.uniform(0.001, 0.10)), 0.01, 0, 0.0) for j, p in enumerate(par)]
    rk, ri = host.simsketch(ref, threads=threads)
    qk, qi = host.simsketch(qry, threads=threads)
    return rk, ri, qk, qi


def split(keys, idx):
    return [keys[idx[i]:idx[i + 1]] for i in range(len(idx) - 1)]


def run_case(name, host, nref, nqry, seed, threads, clades, work):
    rk, ri, qk, qi = sketches(host, nref, nqry, seed, threads, clades)
    hid = int(host.lib.kssd_sketching_id(11, 0, 0, 8))
    R = kt.write_sketch_dir(work / f"{name}_ref", split(rk, ri), conflict=False, drfold=8, hash_id=hid)
    Q = kt.write_sketch_dir(work / f"{name}_qry", split(qk, qi), conflict=False, drfold=8, hash_id=hid)
    env = dict(os.environ, KSSD3A_FORCE_REF_INDEX="1", OMP_NUM_THREADS=str(threads))
    subprocess.run([str(BIN), "sketch", "-i", str(work / f"{name}_ref")], check=True, capture_output=True, env=env)
    out = work / f"{name}.tsv"
    cmd = [str(BIN), "ani", "-r", str(work / f"{name}_ref"), "-q", str(work / f"{name}_qry"), "--format", "matrix",
           "--values", "ani", "-s3", "-p", str(threads), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, env=env)  # warm the page cache
    t = time.perf_counter()
    subprocess.run(cmd, check=True, capture_output=True, env=env)
    t_ref = time.perf_counter() - t
    orc = kt.Oracle()
    rec = np.fromfile(work / f"{name}_ref" / "sortedcomb_ctxgid64obj32", dtype=kt.REC_DTYPE)
    qctx = np.diff(qi).astype(np.uint32)
    rctx = np.diff(ri).astype(np.uint32)
    best = None
    for _ in range(2):
        ctx = np.zeros((nqry, nref, 2), np.uint32)
        obj = np.zeros((nqry, nref, 2), np.uint32)
        vals = np.zeros(nqry * nref, np.float64)
        t = time.perf_counter()
        orc.lib.orc_count_block(kt._p(rec), len(rec), nref, 20, kt._p(qk), kt._p(qi), nqry, 0, 0, kt._p(ctx),
                                kt._p(obj), None, None, threads)
        orc.lib.orc_matrix_values(kt._p(ctx), kt._p(obj), nqry, nref, kt._p(qctx), kt._p(rctx), 3, 1, 0.5, 1.0,
                                  8, 20, 44, kt._p(vals), threads)
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    # same bytes: header + rows of "%lf" cells
    text = out.read_text()
    ref_names = [n for n in R.names]
    rows = ["".join("\t" + n for n in ref_names)]
    v = vals.reshape(nqry, nref)
    for q in range(nqry):
        rows.append(Q.names[q] + "".join("\t%f" % x for x in v[q]))
    same = text == "\n".join(rows) + "\n"
    return {"case": name, "refs": nref, "queries": nqry, "pairs": nref * nqry, "threads": threads,
            "t_reference_s": t_ref, "t_oracle_s": best, "ratio_oracle_over_reference": best / t_ref,
            "same_matrix_bytes": same}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    if not BIN.exists():
        raise SystemExit(f"{BIN} missing: run `make -C oracle ref`")
    host = kt.Host()
    with tempfile.TemporaryDirectory() as d:
        work = Path(d)
        for name, nref, nqry, clades in (("C2_200x200", 200, 200, False), ("C3_1000x1000_clades", 1000, 1000, True)):
            r = run_case(name, host, nref, nqry, 5, args.threads, clades, work)
            print(r, flush=True)


if __name__ == "__main__":
    main()
