#!/usr/bin/env python3
"""Benchmark: ANI pairs/s of the sketch comparison on one MI355X.

Headline workload (C3): a 10,000-genome reference index (500 clades of 20;
each clade a random ancestor, lognormal length median 4 Mbp, sigma 0.5,
clipped to 1-12 Mbp, whose members carry U(1%, 15%) substitutions) against
100,000 distinct query genomes (new members of uniformly chosen clades),
everything sketched at -f8 with the coden pattern by the package's sketcher
(bit-exact with the reference's).  One step = one query-major select over
all 100,000 queries with inputs resident in HBM: every one of the 1e9
(query, reference) pairs is counted and the index path's pass gate
(ctxcut 3, afcut 0.5, anicut 0.95: the reference's defaults) is applied on
the device; each passing pair is emitted with its counts and FP64 CtxMoE
ANI.  Beside it: the HIP-event duration of the launch against the HBM
roofline, the CPU oracle (a C/OpenMP restatement of the reference's block
loop and gate) on the first 4,096 queries, and bit-exact parity of the
passing pairs on that sample.

Side figures in the same line: C2 (1,000 x 1,000 genomes, dense output,
partitioned join) with the plain one-pair-per-launch kernel and per-query
latency; dense C3 blocks and C4-like read sets with unique-owner counters;
the C3 queries streamed from host memory (PCIe-inclusive); the serving loop
under Poisson load on a 20,000-genome index.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "kssd3a-public_amd"))
import kssdhip as kh  # noqa: E402
import tune  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(nref, nqry, seed, threads, host):
    rng = np.random.default_rng(seed)
    lens = np.clip(rng.lognormal(np.log(4e6), 0.5, size=nref), 1e6, 12e6).astype(np.uint64)
    ref_specs = [kh.SimSpec(int(seed * 1_000_003 + i + 1), int(lens[i]), 0, 0.0, 0.0, 0, 0.0)
                 for i in range(nref)]
    parents = rng.integers(0, nref, size=nqry)
    subs = rng.uniform(0.001, 0.10, size=nqry)
    qry_specs = [kh.SimSpec(ref_specs[p].anc_seed, ref_specs[p].anc_len, int(seed * 7_000_001 + j + 1),
                            float(subs[j]), 0.01, 0, 0.0) for j, p in enumerate(parents)]
    t = time.time()
    rkeys, ridx = host.simsketch(ref_specs, threads=threads)
    qkeys, qidx = host.simsketch(qry_specs, threads=threads)
    log(f"simulated+sketched {nref}+{nqry} genomes ({int(lens.sum())/1e9:.2f} Gbp refs) "
        f"in {time.time() - t:.1f}s; entries ref={len(rkeys)} qry={len(qkeys)}")
    return rkeys, ridx, qkeys, qidx


def matched_records(records, qkeys, obj_bits):
    """M: index records whose context occurs in at least one query."""
    qctx = np.unique(qkeys >> np.uint64(obj_bits))
    rctx = records["ctxgid"] >> np.uint64(20)
    return int(np.isin(rctx, qctx, assume_unique=False).sum())


def cpu_baseline(records, nref, obj_bits, qkeys, qidx, qctx, rctx, budget_s, threads):
    """Time the oracle (C/OpenMP restatement of the reference block loop,
    fenceposts rebuilt per block, guided schedule) + its FP64 epilogue on a
    prefix of the queries sized to about budget_s seconds."""
    sys.path.insert(0, str(ROOT / "tests"))
    import kssdtest as kt
    orc = kt.Oracle()

    def run(nq):
        ctx = np.zeros((nq, nref, 2), np.uint32)
        obj = np.zeros((nq, nref, 2), np.uint32)
        out = np.zeros(nq * nref, np.float64)
        keys = np.ascontiguousarray(qkeys[: int(qidx[nq])])
        t = time.perf_counter()
        orc.lib.orc_count_block(kt._p(records), len(records), nref, obj_bits, kt._p(keys),
                                kt._p(qidx[: nq + 1]), nq, 0, 0, kt._p(ctx), kt._p(obj), None, None,
                                threads)
        orc.lib.orc_matrix_values(kt._p(ctx), kt._p(obj), nq, nref, kt._p(qctx), kt._p(rctx), 3, 1,
                                  0.5, 1.0, 8, obj_bits, 44, kt._p(out), threads)
        return time.perf_counter() - t, ctx, obj, out

    nq_total = len(qidx) - 1
    nq = min(8, nq_total)
    dt, *_ = run(nq)
    while dt < budget_s / 8 and nq < nq_total:
        nq = min(nq_total, nq * 4)
        dt, *_ = run(nq)
    if dt < budget_s / 2 and nq < nq_total:
        nq = min(nq_total, max(nq + 1, int(nq * budget_s / max(dt, 1e-3) * 0.8)))
    dt, ctx, obj, out = run(nq)
    return nq, dt, ctx, obj, out


def tiling_key(tiling):
    """The fields that define a join tiling: the query-major join (mode 3)
    by its workgroup size and LDS slots, the other modes by chunk, workgroup
    size and partition group."""
    if tiling["mode"] == 3:
        return (3, tiling.get("query_threads", 0), tiling.get("query_slots", 0))
    return (tiling["mode"], tiling["chunk_records"], tiling["block_threads"], tiling.get("group_partitions", 1))


def profile_tiling(text):
    """The tiling a committed profile summary names (its 'mode ...' line)."""
    import re
    m = re.search(r"mode 3, query_threads (\d+), query_slots (\d+)", text)
    if m:
        return (3, int(m.group(1)), int(m.group(2)))
    m = re.search(r"mode (\d+), chunk_records (\d+), block_threads (\d+), group_partitions (\d+)", text)
    return tuple(map(int, m.groups())) if m else None


def measured_traffic(tiling, nref, nqry):
    """HBM bytes per step of the counting call (k_split + k_pjoin, or the
    query-major join) from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload, when it was measured with the tiling this run
    uses; (bytes or None, summary file name)."""
    import re
    files = [f for f in sorted((ROOT / "profiles").glob(f"*_c2_{nref}x{nqry}*.md"))
             if "traffic_bytes_per_step=" in f.read_text()]
    if not files:
        return None, None
    for f in files:  # one summary per measured tiling
        text = f.read_text()
        m = re.search(r"traffic_bytes_per_step=(\d+)", text)
        if m and profile_tiling(text) == tiling_key(tiling):
            return int(m.group(1)), f"profiles/{f.name}"
    return None, f"profiles/{files[0].name} (measured with other tilings)"


def measured_lds(tiling, nref, nqry):
    """The counting kernel's LDS figures per dispatch from the committed
    rocprofv3 summary of this workload when measured with this run's tiling:
    (dict or None, summary file name)."""
    import re
    files = [f for f in sorted((ROOT / "profiles").glob(f"*_c2_{nref}x{nqry}*.md"))
             if "lds_bytes_per_dispatch=" in f.read_text()]
    if not files:
        return None, None
    for f in files:  # one summary per measured tiling
        text = f.read_text()
        if profile_tiling(text) == tiling_key(tiling):
            vals = {k: float(v) for k, v in re.findall(r"^(lds_\w+)=([0-9.eE+-]+)$", text, re.M)}
            return vals, f"profiles/{f.name}"
    return None, f"profiles/{files[0].name} (measured with other tilings)"


def counters_note(tiling):
    """What bounds the counting kernel at C2, from the committed counter
    summary when it was measured at this run's tiling."""
    if tiling["mode"] == 3:
        import re
        for f in sorted((ROOT / "profiles").glob("qjoin_c2_1000x1000_*.md")):
            text = f.read_text()
            m = re.search(r"line requests per second over `k_qjoin` \| ([0-9.]+) G/s", text)
            if m and profile_tiling(text) == tiling_key(tiling):
                occ = re.search(r"MeanOccupancyPerCU \| ([0-9.]+)", text)
                return {"source": f"profiles/{f.name}", "k_qjoin_line_requests_G_per_s": float(m.group(1)),
                        "waves_per_cu": float(occ.group(1)) if occ else None,
                        "reading": "random 128-byte line fetches of the record table (one home line per query "
                                   "key); the line requests run near the probe's random-line rate"}
        return None
    vals, src = measured_profile("pjoin_c2_1000x1000_counters")
    if not vals or tiling_key(tiling) != (2, 5696, 1024, 4):
        return None
    return {"source": src, "k_pjoin_us": vals.get("k_pjoin_avg_us"),
            "dram_bytes_per_dispatch": vals.get("fetch_bytes_per_dispatch"),
            "wave_cycles_waiting": vals.get("wait_any_frac"), "waves_per_cu": vals.get("mean_occupancy_per_cu"),
            "reading": "latency of dependent phases (stage, scan, bucket build, per-key probe chains) with one "
                       "16-wave workgroup per CU; DRAM at about 29% of peak"}


def matched_events(records, qkeys, obj_bits):
    """H: (query key, index record) pairs sharing a context, the join's
    unit of work (one object-mismatch minimum and one counter update each)."""
    rctx = records["ctxgid"] >> np.uint64(20)
    qctx = qkeys >> np.uint64(obj_bits)
    return int((np.searchsorted(rctx, qctx, "right") - np.searchsorted(rctx, qctx, "left")).sum())


def scratch_dir(need_bytes):
    """A directory for large temporary files: memory-backed /dev/shm when it
    has room for need_bytes, else /tmp."""
    import shutil
    for d in ("/dev/shm", "/tmp"):
        if os.path.isdir(d) and shutil.disk_usage(d).free > need_bytes:
            return d
    return "/tmp"


def real_readsets(host, specs, rng, threads, seed, cov_range=(10.0, 30.0)):
    """Read-set sketches made from reads, as the reference makes them: for
    each genome spec, 2 x 150 bp pairs (insert N(300, 30)) at U(cov_range)x
    with 1% per-base errors are written as plain FASTQ to memory-backed files
    (kssd_sim_fastq_pair) and sketched by the package's `sketch --asone
    --conflict -A` code path in process (kssd_sketch_pooled, -f8), one
    genome per worker thread.  Returns (keys, offsets, read bases)."""
    import shutil
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    covs = rng.uniform(*cov_range, size=len(specs))
    tmp = tempfile.mkdtemp(dir=scratch_dir(threads * 400 << 20))

    def one(i):
        g = host.sim_genome(specs[i])
        a, b = f"{tmp}/s{i}_1.fq", f"{tmp}/s{i}_2.fq"
        n = host.sim_fastq_pair(g, covs[i], seed * 1_000_003 + i, a, b)
        try:
            k, _ = host.sketch_pooled([a, b], threads=1)
        finally:
            os.unlink(a)
            os.unlink(b)
        return k, 2 * 150 * n

    try:
        with ThreadPoolExecutor(max(1, threads)) as ex:
            res = list(ex.map(one, range(len(specs))))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    keys = np.concatenate([k for k, _ in res]) if res else np.zeros(0, np.uint64)
    idx = np.zeros(len(res) + 1, np.uint64)
    idx[1:] = np.cumsum([len(k) for k, _ in res])
    return keys, idx, int(sum(b for _, b in res))


def clade_specs(rng, nref, seed):
    """nref genomes = clades of 20: a random ancestor per clade (lognormal
    length, median 4 Mbp, sigma 0.5, clipped to 1-12 Mbp) whose members carry
    U(1%, 15%) substitutions.  Returns (specs, clade seeds, clade lengths)."""
    nclade = nref // 20
    clen = np.clip(rng.lognormal(np.log(4e6), 0.5, size=nclade), 1e6, 12e6).astype(np.uint64)
    cseed = [int(seed * 9_000_011 + c + 1) for c in range(nclade)]
    specs = [kh.SimSpec(cseed[c], int(clen[c]), int(seed * 5_000_003 + 20 * c + m + 1),
                        float(rng.uniform(0.01, 0.15)), 0.0, 0, 0.0)
             for c in range(nclade) for m in range(20)]
    return specs, cseed, clen


def index_build_bench(host, hip, rk, ri, threads, reps=1):
    """The inverted index of sketch (rk, ri) built by the host builder and by
    kssdhip_build_index (host buffers in and out, so its time includes both
    PCIe copies; best of `reps`); the GPU records must be the host's bytes."""
    t = time.perf_counter()
    rec = host.build_index(rk, ri, 20, threads=threads)
    host_s = time.perf_counter() - t
    hip.build_index(rk[: int(ri[1])], ri[:2])  # warm the kernels
    gpu_s = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        grec, gbits = hip.build_index(rk, ri)
        gpu_s = min(gpu_s, time.perf_counter() - t)
    build = {"records": int(len(rec)), "gpu_ms_host_buffers": gpu_s * 1e3,
             "gpu_records_per_s": len(rec) / gpu_s, "host_ms": host_s * 1e3, "host_threads": threads,
             "host_what": "kssd_build_index (bucketed parallel sort, records only, no unique bits)",
             "gpu_what": "kssdhip_build_index: upload keys, records + LSD radix sort + unique bits, download both",
             "records_identical": bool(grec.tobytes() == rec.tobytes()), "unique_bits_bytes": len(gbits)}
    return rec, build


def edges_bench(args, host, threads):
    """`matrix --format edges` end to end on a GENOMES-genome clade sketch
    written to disk with its index: kssd3a-hip (votes and pair counts on the
    GPU) against the reference's own binary (oracle/_ref/kssd3a_ref, when
    present) with the same threads; the edge sets must agree (the
    reference's row order depends on its thread schedule)."""
    import subprocess
    import tempfile
    rng = np.random.default_rng(args.seed + 300)
    nref = args.edges_genomes - args.edges_genomes % 20
    specs, _, _ = clade_specs(rng, nref, args.seed + 300)
    rk, ri = host.simsketch(specs, threads=threads)
    hip = kh.KssdHip(0)
    rec, bits = hip.build_index(rk, ri)
    out = {"genomes": nref, "records": int(len(rec)), "cut": args.edges_cut, "threads": threads}
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        sk = Path(d) / "sketch"
        host.write_sketch(sk, rk, ri, [f"genome{i:05d}" for i in range(nref)], records=rec, unique_bits=bits)
        del rk, rec, bits
        runs = {"gpu": [str(kh.PKG / "kssd3a-hip")]}
        refbin = ROOT / "oracle" / "_ref" / "kssd3a_ref"
        if refbin.exists():
            runs["reference"] = [str(refbin)]
        text = {}
        for name, cmd in runs.items():
            argv = [*cmd, "matrix", "--format", "edges", "--cut", str(args.edges_cut), "-p", str(threads), str(sk)]
            subprocess.run(argv, capture_output=True, check=True)  # warm the page cache
            t = time.perf_counter()
            p = subprocess.run(argv, capture_output=True, text=True, check=True)
            out[f"{name}_s"] = time.perf_counter() - t
            text[name] = p.stdout
        lines = text["gpu"].splitlines()
        out["edges"] = len(lines) - 1
        if "reference" in text:
            ref_lines = text["reference"].splitlines()
            out["same_edges"] = lines[0] == ref_lines[0] and sorted(lines[1:]) == sorted(ref_lines[1:])
            out["reference_over_gpu"] = out["reference_s"] / out["gpu_s"]
    return out


def write_fastq_pair(path1, path2, rng, genome, coverage, read_len=150, insert=300, sd=30, err=0.01):
    """Paired-end reads of `genome` (2-bit codes): fragments of N(insert, sd)
    bp from either strand, read 1 the fragment's first read_len bases, read 2
    the reverse complement of its last; i.i.d. per-base errors to one of the
    other three bases; a constant quality line.  Plain 4-line FASTQ."""
    L = len(genome)
    n = int(coverage * L / (2 * read_len))
    frag = np.clip(rng.normal(insert, sd, size=n).astype(np.int64), read_len, L - 1)
    start = (rng.random(n) * (L - frag)).astype(np.int64)
    cols = np.arange(read_len)
    r1 = genome[start[:, None] + cols]
    r2 = 3 - genome[(start + frag - 1)[:, None] - cols]
    flip = rng.random(n) < 0.5  # fragments from the reverse strand: swap the mates
    r1[flip], r2[flip] = r2[flip].copy(), r1[flip].copy()
    letters = np.frombuffer(b"ACGT", np.uint8)
    digits = np.empty((n, 9), np.uint8)  # "@r%09d/%c\n" headers, vectorised
    idx = np.arange(n, dtype=np.int64)
    for k in range(9):
        digits[:, 8 - k] = 48 + (idx // 10 ** k) % 10
    for path, reads, tag in ((path1, r1, b"1"), (path2, r2, b"2")):
        flat = reads.reshape(-1)  # errors: a binomial count of distinct positions
        pos = np.unique(rng.integers(0, flat.size, size=int(rng.binomial(flat.size, err)), dtype=np.int64))
        flat[pos] = (flat[pos] + rng.integers(1, 4, size=len(pos), dtype=np.uint8)) % 4
        rec = np.empty((n, 14 + read_len + 3 + read_len + 1), np.uint8)
        rec[:, 0], rec[:, 1] = ord("@"), ord("r")
        rec[:, 2:11] = digits
        rec[:, 11], rec[:, 12], rec[:, 13] = ord("/"), tag[0], ord("\n")
        rec[:, 14:14 + read_len] = letters[reads]
        rec[:, 14 + read_len:17 + read_len] = np.frombuffer(b"\n+\n", np.uint8)
        rec[:, 17 + read_len:17 + 2 * read_len] = ord("I")
        rec[:, -1] = ord("\n")
        Path(path).write_bytes(rec.tobytes())
    return n


def readset_bench(args, host, threads):
    """The read-set pipeline on files, kssd3a-hip against the reference's own
    binary (oracle/_ref/kssd3a_ref, built from its sources) with the same
    threads, and their outputs compared byte for byte:
      genomes  args.readsets_bench clade genomes (as the index bench) as FASTA:
               `sketch -o`, then `sketch -i` (GPU index builder);
      samples  args.readsets_samples read sets, each a reference genome plus 0.5%
               substitutions read as 2x150 bp pairs (insert 300 +- 30) at
               U(10, 30)x with 1% per-base errors, plain FASTQ pairs:
               `sketch --asone --conflict -A` per sample;
      ani      each sample against the genomes: `--qraw --unique-ctx
               --estimate-coverage` (indexed path with the unique-owner
               counters; coverage columns from comblco.a and .unique_bits)."""
    import subprocess
    import tempfile
    rng = np.random.default_rng(args.seed + 400)
    nref = args.readsets_bench - args.readsets_bench % 20
    specs, _, _ = clade_specs(rng, nref, args.seed + 400)
    refbin = ROOT / "oracle" / "_ref" / "kssd3a_ref"
    runs = {"gpu": str(kh.PKG / "kssd3a-hip")}
    if refbin.exists():
        runs["reference"] = str(refbin)
    out = {"genomes": nref, "samples": args.readsets_samples, "threads": threads}
    letters = np.frombuffer(b"ACGT", np.uint8)

    def timed(name, argv, cwd):
        t = time.perf_counter()
        p = subprocess.run(argv, capture_output=True, cwd=cwd)
        if p.returncode:
            raise RuntimeError(f"{name}: {' '.join(argv[1:4])}...: {p.stderr.decode()[-400:]}")
        if name == "gpu" and os.environ.get("KSSDHIP_TIMING"):  # the CLI's stage times
            for ln in p.stderr.decode().splitlines():
                if ln.startswith("[kssd3a-hip"):
                    log(ln)
            log(f"[{argv[1]} wall] {time.perf_counter() - t:.3f} s")
        return time.perf_counter() - t, p.stdout

    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        d = Path(d)
        gnames = []
        for i, sp in enumerate(specs):
            g = host.sim_genome(sp)
            gnames.append(f"g{i:05d}.fa")
            (d / gnames[-1]).write_bytes(b">g%05d\n" % i + letters[g].tobytes() + b"\n")
        samples, bases = [], 0
        for k in range(args.readsets_samples):
            sp = specs[int(rng.integers(0, nref))]
            sample = kh.SimSpec(sp.anc_seed, sp.anc_len, int(args.seed * 3_000_017 + k + 1), 0.005, 0.0,
                                sp.mut_seed, sp.sub_rate)
            g = host.sim_genome(sample)
            cov = float(rng.uniform(10, 30))
            n = write_fastq_pair(d / f"s{k:03d}_1.fq", d / f"s{k:03d}_2.fq", rng, g, cov)
            bases += 2 * 150 * n
            samples.append((f"s{k:03d}_1.fq", f"s{k:03d}_2.fq"))
        out["read_bases"] = bases
        log(f"readsets bench: {nref} genomes and {len(samples)} read sets ({bases / 1e9:.2f} Gbp of reads) written")
        res = {}
        for name, exe in runs.items():
            r = {}
            p = str(threads)
            r["sketch_genomes_s"], _ = timed(name, [exe, "sketch", "-p", p, "-o", f"G_{name}", *gnames], d)
            r["index_s"], _ = timed(name, [exe, "sketch", "-i", f"G_{name}"], d)
            r["sketch_reads_s"], r["ani_s"], tsv = 0.0, 0.0, []
            for k, (a, b) in enumerate(samples):
                dt, _ = timed(name, [exe, "sketch", "--asone", "--conflict", "-A", "-p", p, "-o",
                                     f"R{k}_{name}", a, b], d)
                r["sketch_reads_s"] += dt
                dt, o = timed(name, [exe, "ani", "-r", f"G_{name}", "--qraw", f"R{k}_{name}", "--unique-ctx",
                                     "--estimate-coverage", "-p", p], d)
                r["ani_s"] += dt
                tsv.append(o)
            res[name] = (r, tsv)
            out[name] = r
            log(f"readsets bench {name}: " + ", ".join(f"{k} {v:.2f}" for k, v in r.items()))
        if "reference" in res:
            same = {}
            for f in ("comblco", "comblco.index", "lcofiles.infilemeta", "sortedcomb_ctxgid64obj32",
                      "sortedcomb_ctxgid64obj32.unique_bits"):
                same[f"genomes/{f}"] = (d / "G_gpu" / f).read_bytes() == (d / "G_reference" / f).read_bytes()
            for f in ("comblco", "comblco.index", "comblco.a", "lcofiles.qc", "lcofiles.infilemeta"):
                same[f"reads/{f}"] = all((d / f"R{k}_gpu" / f).read_bytes() == (d / f"R{k}_reference" / f).read_bytes()
                                         for k in range(len(samples)))
            same["ani_tsv"] = res["gpu"][1] == res["reference"][1]
            out["identical"] = same
            out["reference_over_gpu"] = {k: res["reference"][0][k] / res["gpu"][0][k] for k in res["gpu"][0]}
            log(f"readsets bench identical outputs: {all(same.values())} {same if not all(same.values()) else ''}")
    return out


def clade_queries(rng, n, nclade, cseed, clen, seed):
    """n new members of uniformly chosen clades: the clade's ancestor with
    U(1%, 15%) substitutions of its own."""
    clades = rng.integers(0, nclade, size=n)
    divs = rng.uniform(0.01, 0.15, size=n)
    return [kh.SimSpec(cseed[c], int(clen[c]), int(seed * 4_000_037 + i + 1), float(divs[i]), 0.0, 0, 0.0)
            for i, c in enumerate(clades)]


def simsketch_chunked(host, specs, threads, chunk=10000):
    keys, sizes = [], []
    for b in range(0, len(specs), chunk):
        k, idx = host.simsketch(specs[b:b + chunk], threads=threads)
        keys.append(k)
        sizes.append(np.diff(idx))
    qi = np.zeros(len(specs) + 1, np.uint64)
    qi[1:] = np.cumsum(np.concatenate(sizes)) if sizes else []
    out = np.concatenate(keys) if keys else np.zeros(0, np.uint64)
    return out, qi


def measured_profile(name):
    """Figures of a committed rocprofv3 summary (profiles/<name>.md):
    key=value lines; (dict or None, file name)."""
    import re
    f = ROOT / "profiles" / f"{name}.md"
    if not f.exists():
        return None, None
    vals = {k: float(v) for k, v in re.findall(r"^(\w+)=([0-9.eE+-]+)\s*$", f.read_text(), re.M)}
    return vals, f"profiles/{f.name}"


def c3_bench(args, host, threads):
    """The headline workload: a 10,000-genome clade index (500 clades of 20)
    against 100,000 distinct query genomes (new members of random clades),
    all sketched at -f8 by the package's sketcher; inputs resident in HBM.
    A step = one kssdhip_select_block_dev over all 100,000 queries: the
    query-major join counts every (query, reference) pair (1e9 pairs) and
    applies the index path's pass gate with the reference's defaults
    (ctxcut 3, afcut 0.5, anicut 0.95), emitting each passing pair with its
    counts and FP64 CtxMoE ANI."""
    rng = np.random.default_rng(args.seed + 200)
    nref = args.index_refs - args.index_refs % 20
    nclade = nref // 20
    specs, cseed, clen = clade_specs(rng, nref, args.seed + 200)
    qspecs = clade_queries(rng, args.c3_queries, nclade, cseed, clen, args.seed)
    cache = Path(args.data_cache) / f"c3_{nref}_{args.c3_queries}_{args.seed}" if args.data_cache else None
    if cache is not None and (cache / "qi.npy").exists():
        t = time.time()
        rk, ri, qk, qi = (np.load(cache / f"{x}.npy") for x in ("rk", "ri", "qk", "qi"))
        t_ref = t_qry = 0.0
        log(f"C3: sketches loaded from {cache} in {time.time() - t:.1f}s")
    else:
        t = time.time()
        rk, ri = host.simsketch(specs, threads=threads)
        t_ref = time.time() - t
        t = time.time()
        qk, qi = simsketch_chunked(host, qspecs, threads)
        t_qry = time.time() - t
        if cache is not None:
            cache.mkdir(parents=True, exist_ok=True)
            for x, v in (("rk", rk), ("ri", ri), ("qk", qk), ("qi", qi)):
                np.save(cache / f"{x}.npy", v)
    nq, total = len(qi) - 1, int(qi[-1])
    log(f"C3: {nref} reference genomes sketched in {t_ref:.1f}s ({len(rk)} entries); {nq} queries in {t_qry:.1f}s "
        f"({total} entries, {total * 8 / 1e9:.1f} GB)")
    hip = kh.KssdHip(0)
    rt = kh.Hip(0)
    rec, build = index_build_bench(host, hip, rk, ri, threads)
    rctx = host.ctx_counts(rk, ri, 20)
    qctx = host.ctx_counts(qk, qi, 20)
    h = hip.open_index(rec, nref, 20)
    stream = rt.stream()
    d_k, d_i = rt.upload(qk), rt.upload(qi)
    d_qc, d_rc = rt.upload(qctx), rt.upload(rctx)
    cap = 1 << 24
    d_rows, d_n = rt.malloc(56 * cap), rt.malloc(4)
    d_first, d_cnt = rt.malloc(4 * nq), rt.malloc(4 * nq)
    if args.copy_probe_mib:  # a copy of known size: calibrates the HBM counters of a profiled run
        nb_ = args.copy_probe_mib << 20
        a_, b_ = rt.malloc(nb_), rt.malloc(nb_)
        rt.check(rt.lib.hipMemcpy(b_, a_, nb_, rt.D2D))
        rt.free(a_)
        rt.free(b_)
    value = kh.EpilogueOpts(3, 1, 0, 1, 0, 0, 8, 20, 44, 0.5, 1.0)  # CtxMoE ANI, float AFs (index path)
    sopts = kh.SelectOpts(3, 0.5, 0.95, 0, value)
    # the kernel shape: the library's autotuner on a reference-sized block of
    # the queries (dense counts; the query-major shapes it tries are the
    # select's)
    nt = min(4096, nq)
    t = time.time()
    d_tc, d_to = rt.malloc(8 * nt * nref), rt.malloc(8 * nt * nref)
    tiling, tune_sum = tune.tune(hip, rt, h, d_k, d_i, nt, int(qi[nt]), nref, 0, d_tc, d_to, stream,
                                 force=args.retune, log=log)
    rt.free(d_tc)
    rt.free(d_to)
    if tiling["mode"] == 3 and nref * 16 > 64 * 1024:
        # the select spills what a query's LDS table cannot hold: its table size
        # timed on the select itself (the first 10,000 queries)
        ns = min(10000, nq)
        best = None
        for qs in (256, 512, 1024):
            hip.set_tiling(h, query_slots=qs)
            for rep in range(4):
                if rep == 1:
                    e0, e1 = rt.event(), rt.event()
                    rt.record(e0, stream)
                hip.select_block_dev(h, d_k, d_i, ns, int(qi[ns]), 0, d_qc, d_rc, sopts, d_rows, cap, d_n, d_first,
                                     d_cnt, stream)
            rt.record(e1, stream)
            rt.sync(stream)
            ms_qs = rt.elapsed_ms(e0, e1) / 3
            if best is None or ms_qs < best[1]:
                best = (qs, ms_qs)
        hip.set_tiling(h, query_slots=best[0])
        tiling["query_slots"] = best[0]
        tune_sum["select_query_slots_ms"] = best[1]
    log(f"C3 tiling {tiling} ({time.time() - t:.1f}s)")

    def step():
        hip.select_block_dev(h, d_k, d_i, nq, total, 0, d_qc, d_rc, sopts, d_rows, cap, d_n, d_first, d_cnt,
                             stream)

    step()  # sizes the workspace before any capture
    rt.sync(stream)
    for _ in range(args.warmup):
        step()
    rt.sync(stream)
    evs = [(rt.event(), rt.event()) for _ in range(args.steps)]
    e0, e1 = rt.event(), rt.event()
    t0 = time.perf_counter()
    rt.record(e0, stream)
    for a_, b_ in evs:
        rt.record(a_, stream)
        step()
        rt.record(b_, stream)
    rt.record(e1, stream)
    rt.sync(stream)
    wall = time.perf_counter() - t0
    eager_ms = rt.elapsed_ms(e0, e1) / args.steps
    launch_ms = [rt.elapsed_ms(a_, b_) for a_, b_ in evs]
    n_rows = int(rt.download(d_n, np.uint32, 1)[0])
    rows = rt.download(d_rows, np.uint8, 56 * min(n_rows, cap)).view(kh.ROW_DTYPE).copy()
    q_cnt = rt.download(d_cnt, np.uint32, nq)
    # the same step as a HIP graph
    g, gx = rt.capture(stream, step)
    for _ in range(args.warmup):
        rt.graph_launch(gx, stream)
    rt.sync(stream)
    rt.record(e0, stream)
    for _ in range(args.steps):
        rt.graph_launch(gx, stream)
    rt.record(e1, stream)
    rt.sync(stream)
    graph_ms = rt.elapsed_ms(e0, e1) / args.steps
    g_rows = rt.download(d_rows, np.uint8, 56 * min(n_rows, cap)).view(kh.ROW_DTYPE)
    graph_ok = int(rt.download(d_n, np.uint32, 1)[0]) == n_rows and np.array_equal(
        np.sort(g_rows, order=["query", "ref"]), np.sort(rows, order=["query", "ref"]))
    rt.graph_destroy(g, gx)
    pairs = nq * nref
    ms = graph_ms if graph_ok else eager_ms
    log(f"C3 step: {eager_ms:.2f} ms eager, {graph_ms:.2f} ms graph (agrees {graph_ok}), "
        f"{pairs / (ms / 1e3):.3e} pairs/s; {n_rows} passing pairs")
    props = {"rows_fit": n_rows <= cap, "segments_sum": int(q_cnt.sum()) == n_rows,
             "rows_pass_ctxcut": bool((rows["num_ctx"] >= 3).all()),
             "rows_pass_anicut": bool((rows["rank_ani"] >= np.float64(np.float32(0.95))).all())}
    # dense counts of one reference-sized block by both joins agree cell for cell
    nb = min(1024, nq)
    d_c1, d_o1, d_c2, d_o2 = (rt.malloc(8 * nb * nref) for _ in range(4))
    forms = {}
    for form, tiling, dc, do in (("persistent_query_major", dict(mode=3), d_c1, d_o1),
                                 ("static_partitioned", dict(mode=2, chunk_records=5696, block_threads=1024,
                                                             group_partitions=4), d_c2, d_o2)):
        hip.set_tiling(h, **tiling)
        hip.count_block_dev(h, d_k, d_i, nb, int(qi[nb]), 0, 0, dc, do, None, None, stream)  # warm
        t_p = rt.event(), rt.event()
        rt.record(t_p[0], stream)
        for _ in range(3):
            hip.count_block_dev(h, d_k, d_i, nb, int(qi[nb]), 0, 0, dc, do, None, None, stream)
        rt.record(t_p[1], stream)
        rt.sync(stream)
        forms[form] = {"ms": rt.elapsed_ms(*t_p) / 3, "tiling": tiling}
    forms["what"] = (f"dense counts of the first {nb} queries x {nref} refs: the persistent query-major join "
                     "(workgroups take queries from a device work queue, largest first) against the static "
                     "partitioned merge-join grid (one workgroup per index-partition group)")
    c1_cells = rt.download(d_c1, np.uint64, nb * nref)
    props["query_major_equals_partitioned_on_a_block"] = bool(
        np.array_equal(c1_cells, rt.download(d_c2, np.uint64, nb * nref))
        and np.array_equal(rt.download(d_o1, np.uint64, nb * nref), rt.download(d_o2, np.uint64, nb * nref)))
    p_touched_block = int(((c1_cells & np.uint64(0xffffffff)) > 0).sum())
    hip.set_tiling(h, mode=0)
    for d in (d_c1, d_o1, d_c2, d_o2):
        rt.free(d)
    # roofline: compulsory bytes of the launch, B = 8 N_Q + 12 M + 16 P_out
    M = hip.matched_records(h, d_k, d_i, nq, total)
    algo = 8 * total + 12 * M + 16 * n_rows
    prof, prof_src = measured_profile("qjoin_c3_10k_x_100k")
    kern_ms = prof.get("k_qjoin_avg_ms") if prof else None
    t_launch = float(np.mean(launch_ms)) / 1e3
    traffic = int(prof["traffic_bytes_per_launch"]) if prof and "traffic_bytes_per_launch" in prof else None
    dense_algo = 8 * total + 12 * M + 16 * pairs
    table_lines = hip.table_info(h)[0]
    line_rate = hip.probe_random_reads(nbytes=table_lines * 128)
    dram_req = prof.get("dram_read_requests_per_launch") if prof else None
    random_access = {
        "what": "k_qjoin is bound by independent random line fetches: one 128-byte line of the record table per "
                "query context looked up (plus overflow lines, long contexts and the streamed keys), not by "
                "streamed bytes. Its lookups per second (every query key of an assembly sketch heads its own "
                "context group) against the rate of random 128-byte line reads, in the kernel's load form, from "
                "a buffer of the record table's size (rounded down to a power of two) measured in this run "
                "(kssdhip_probe_random_reads), and its DRAM line requests from the committed profile",
        "probe_lines_per_s": line_rate,
        "probe_buffer_bytes": 1 << (int(table_lines * 128).bit_length() - 1),
        "lookups_per_launch": total,
        "lookups_per_s": total / t_launch,
        "frac": total / t_launch / line_rate,
        "k_qjoin_line_requests_per_launch": dram_req,
        "k_qjoin_lines_per_lookup": dram_req / total if dram_req else None,
        "k_qjoin_line_requests_per_s": dram_req / (kern_ms / 1e3) if dram_req and kern_ms else None,
        "line_requests_frac": dram_req / (kern_ms / 1e3) / line_rate if dram_req and kern_ms else None,
        "source": prof_src}
    roof = {"bound": "hbm", "achieved": algo / t_launch / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": algo / t_launch / 1e9 / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": prof_src,
            "traffic_GBs": traffic / t_launch / 1e9 if traffic else None, "kernel": "k_qjoin (select)",
            "algorithmic_bytes": algo, "bytes_model": "8*N_Q + 12*M + 16*P_out", "N_Q": total, "M": M,
            "P_out": n_rows, "index_records": int(len(rec)), "avg_launch_ms": t_launch * 1e3,
            "avg_launch_what": "HIP events around kssdhip_select_block_dev on its stream (order + k_qjoin + "
                               "row values)",
            "k_qjoin_avg_ms_rocprof": kern_ms,
            "P_out_note": "the select writes only passing pairs, so P_out = passing pairs; with a dense Q x R output "
                          "(P_out = Q*R) the same launch would move dense_output_bytes",
            "dense_output_bytes": dense_algo, "frac_dense_output_model": dense_algo / t_launch / 1e9 / HBM_PEAK_GBS,
            "P_touched_first_block": p_touched_block, "P_touched_block_queries": nb,
            "random_access": random_access}
    # LDS use of the launch (rocprofv3 counters of the committed profile)
    lds_prof, lds_src = measured_profile("qjoin_c3_10k_x_100k_lds")
    lds_peak = 256 * 128 * 2.4e9 / 1e9  # GB/s: 256 CUs x 128 B/clk x 2.4 GHz
    lds = {"what": "LDS bytes of one select launch (staged home lines written and scanned, walk reads, lane "
                   "shuffles, counter atomics) over its HIP-event time, against the LDS peak",
           "bytes_per_launch": lds_prof.get("lds_bytes_per_dispatch") if lds_prof else None,
           "achieved_GBs": lds_prof["lds_bytes_per_dispatch"] / t_launch / 1e9 if lds_prof else None,
           "peak_GBs": lds_peak,
           "frac": lds_prof["lds_bytes_per_dispatch"] / t_launch / 1e9 / lds_peak if lds_prof else None,
           "lds_util_pct": lds_prof.get("lds_util_pct") if lds_prof else None,
           "bank_conflict": lds_prof.get("lds_bank_conflict") if lds_prof else None,
           "source": lds_src}
    # CPU baseline (the oracle's block count with per-block fenceposts and
    # guided OpenMP, then its restatement of the gate) on the first block
    sys.path.insert(0, str(ROOT / "tests"))
    import kssdtest as kt
    orc = kt.Oracle()
    ncpu = min(args.c3_cpu_queries, nq)
    c_ctx = np.zeros((ncpu, nref, 2), np.uint32)
    c_obj = np.zeros((ncpu, nref, 2), np.uint32)
    sub = np.ascontiguousarray(qk[: int(qi[ncpu])])
    t = time.perf_counter()
    orc.lib.orc_count_block(kt._p(rec), len(rec), nref, 20, kt._p(sub), kt._p(qi[: ncpu + 1]), ncpu, 0, 0,
                            kt._p(c_ctx), kt._p(c_obj), None, None, threads)
    ranked = orc.index_rank(c_ctx, c_obj, c_ctx, c_obj, rctx, qctx[:ncpu], ctxcut=3, afcut=0.5, anicut=0.95)
    cpu_s = time.perf_counter() - t
    mine = np.sort(rows[rows["query"] < ncpu], order=["query", "ref"])
    want_q = np.concatenate([np.full(len(p[0]), q, np.uint32) for q, p in enumerate(ranked)])
    want_r = np.concatenate([p[0] for p in ranked])
    exact = bool(np.array_equal(mine["query"], want_q) and np.array_equal(mine["ref"], want_r))
    if exact and len(mine):
        qq, rr = mine["query"].astype(np.int64), mine["ref"].astype(np.int64)
        exact = bool(np.array_equal(mine["num_ctx"], c_ctx[qq, rr, 0])
                     and np.array_equal(mine["num_mut2_ctx"], c_ctx[qq, rr, 1])
                     and np.array_equal(mine["diff_obj"], c_obj[qq, rr, 0])
                     and np.array_equal(mine["diff_obj_section"], c_obj[qq, rr, 1])
                     and np.array_equal(mine["rank_ani"], np.concatenate([p[1] for p in ranked])))
    cpu_pps = ncpu * nref / cpu_s
    log(f"C3 CPU oracle: {ncpu} queries in {cpu_s:.2f}s = {cpu_pps:.3e} pairs/s; passing pairs equal: {exact} "
        f"({len(mine)} pairs)")
    res = {
        "value": pairs / (ms / 1e3), "ms_per_step": ms, "pairs": pairs, "queries": nq, "refs": nref,
        "query_entries": total, "index_records": int(len(rec)), "passing_pairs": n_rows,
        "step": {"timed": "HIP graph of one kssdhip_select_block_dev over all queries" if graph_ok
                 else "eager launches", "graph_ms": graph_ms, "eager_ms": eager_ms, "graph_agrees_with_eager": graph_ok,
                 "wall_ms_per_step_eager": wall / args.steps * 1e3},
        "gate": {"ctxcut": 3, "afcut": 0.5, "anicut": 0.95, "value": "CtxMoE ANI (FP64, float AFs)"},
        "parity": {"cpu_sample_queries": ncpu, "passing_pairs_bit_exact": exact, "properties": props},
        "forms": forms,
        "tiling": tiling, "tune": tune_sum,
        "roofline": roof, "lds": lds,
        "cpu_baseline": {"value": cpu_pps, "unit": "pairs/s", "cores": threads, "kind": "port",
                         "sample": f"first {ncpu} queries x {nref} refs: oracle block count (fenceposts per block, "
                                   "guided OpenMP) + its restatement of the pass gate (command_ani.h:577-621)"},
        "sketch_s": {"refs": t_ref, "queries": t_qry}, "index_build": build,
    }
    for d in (d_k, d_i, d_qc, d_rc, d_rows, d_n, d_first, d_cnt):
        rt.free(d)
    shared = {"hip": hip, "rt": rt, "h": h, "rec": rec, "rctx": rctx, "specs": specs, "qk": qk, "qi": qi,
              "nref": nref, "cseed": cseed, "clen": clen, "rk": rk, "ri": ri, "rows": rows}
    return res, shared


def c4_profile(tiling):
    """The committed rocprofv3 summary of the C4 count when it was measured
    at this run's tiling: (figures or None, summary file name)."""
    import re
    for f in sorted((ROOT / "profiles").glob("qjoin_c4_*.md")):
        text = f.read_text()
        if profile_tiling(text) == tiling_key(tiling):
            return ({k: float(v) for k, v in re.findall(r"^(\w+)=([0-9.eE+-]+)$", text, re.M)},
                    f"profiles/{f.name}")
    return None, None


def c4_index(args, host, threads):
    """The C3 reference index alone (10,000 clade genomes, no queries), for
    runs of the C4 count by itself."""
    rng = np.random.default_rng(args.seed + 200)
    nref = args.index_refs - args.index_refs % 20
    specs, cseed, clen = clade_specs(rng, nref, args.seed + 200)
    cache = Path(args.data_cache) / f"c4idx_{nref}_{args.seed}" if args.data_cache else None
    if cache is not None and (cache / "ri.npy").exists():
        rk, ri = (np.load(cache / f"{x}.npy") for x in ("rk", "ri"))
    else:
        rk, ri = host.simsketch(specs, threads=threads)
        if cache is not None:
            cache.mkdir(parents=True, exist_ok=True)
            np.save(cache / "rk.npy", rk)
            np.save(cache / "ri.npy", ri)
    hip = kh.KssdHip(0)
    rt = kh.Hip(0)
    rec = host.build_index(rk, ri, 20, threads=threads)
    h = hip.open_index(rec, nref, 20)
    if args.copy_probe_mib:  # a copy of known size: calibrates the HBM counters of a profiled run
        nb_ = args.copy_probe_mib << 20
        a_, b_ = rt.malloc(nb_), rt.malloc(nb_)
        rt.check(rt.lib.hipMemcpy(b_, a_, nb_, rt.D2D))
        rt.free(a_)
        rt.free(b_)
    return {"hip": hip, "rt": rt, "h": h, "rec": rec, "rctx": host.ctx_counts(rk, ri, 20), "specs": specs,
            "qk": np.zeros(0, np.uint64), "qi": np.zeros(1, np.uint64), "nref": nref}


def index_bench(args, host, threads, plain_s_per_pair, shared, names=("assemblies", "read_sets")):
    """The C3 index (resident) against two query sets counted in one batched
    launch each into dense blocks, with inputs resident in HBM:
      assemblies  the first args.index_queries of the C3 queries (one
                  reference-style block), dense counts + FP64 CtxMoE ANI;
      read sets   C4: args.index_readsets samples, each a reference genome
                  plus 0.5% substitutions read as 2 x 150 bp pairs at
                  U(10, 30)x with 1% errors and sketched with the package's
                  `sketch --asone --conflict -A` (see real_readsets), counted
                  with the unique-owner counters (--qraw --unique-ctx).
    Beside each: the oracle (C/OpenMP restatement of the reference's block
    loop) on a bounded prefix of the same queries, and the plain kernel's
    measured per-pair time."""
    rng = np.random.default_rng(args.seed + 300)
    hip, rt, h, rec, rctx, specs = (shared[k] for k in ("hip", "rt", "h", "rec", "rctx", "specs"))
    nref = shared["nref"]
    nq = min(args.index_queries, len(shared["qi"]) - 1)
    qi = shared["qi"][: nq + 1].copy()
    qk = shared["qk"][: int(qi[-1])]
    nrs = args.index_readsets
    rs = []
    for i in range(nrs):
        r = specs[int(rng.integers(0, nref))]
        rs.append(kh.SimSpec(r.anc_seed, r.anc_len, int(args.seed * 6_000_011 + i + 1), 0.005, 0.0,
                             r.mut_seed, r.sub_rate))
    t = time.time()
    rkeys, ridx, read_bases = real_readsets(host, rs, rng, threads, args.seed + 300)
    log(f"C4: {nrs} read sets ({read_bases / 1e9:.1f} Gbp of simulated reads) sketched in {time.time() - t:.1f}s, "
        f"median {int(np.median(np.diff(ridx)))} entries")
    stream = rt.stream()
    out = {"index_genomes": nref, "index_records": int(len(rec)),
           "read_sets": {"what": "C4: read sets sketched from simulated 2x150 bp reads (U(10,30)x, 1% errors) by "
                                 "the package's sketch --asone --conflict -A path", "read_bases": read_bases,
                         "sketch_s": time.time() - t}}
    orc = None
    for name, keys, idx, flags in (("assemblies", qk, qi, 0), ("read_sets", rkeys, ridx, kh.UNIQUE)):
        if name not in names:
            continue
        n = len(idx) - 1
        cells = n * nref
        unique = bool(flags & kh.UNIQUE)
        d_k, d_i = rt.upload(keys), rt.upload(idx)
        qctx = host.ctx_counts(keys, idx, 20, conflict=unique)
        d_qc, d_rc = rt.upload(qctx), rt.upload(rctx)
        d_c, d_o, d_out = rt.malloc(8 * cells), rt.malloc(8 * cells), rt.malloc(8 * cells)
        d_uc = rt.malloc(8 * cells) if unique else None
        d_uo = rt.malloc(8 * cells) if unique else None
        # --qraw reads: naive metric (s4); assemblies: CtxMoE (s3)
        opts = kh.EpilogueOpts(4 if unique else 3, 1, int(unique), 0, 0, 0, 8, 20, 44, 0.2 if unique else 0.5, 1.0)
        total = int(idx[-1])

        def step():
            hip.count_block_dev(h, d_k, d_i, n, total, 0, flags, d_c, d_o, d_uc, d_uo, stream)
            hip.epilogue_dev(d_c, d_o, n, nref, d_qc, d_rc, opts, d_out, stream)

        tiling, _ = tune.tune(hip, rt, h, d_k, d_i, n, total, nref, len(rec), d_c, d_o, stream,
                              flags=flags, force=args.retune, log=log)
        step()
        rt.sync(stream)
        e0, e1 = rt.event(), rt.event()
        reps = 3
        rt.record(e0, stream)
        for _ in range(reps):
            step()
        rt.record(e1, stream)
        ms = rt.elapsed_ms(e0, e1) / reps
        pairs_per_s = cells / (ms / 1e3)
        # the counting call alone: its roofline (B = 8 N_Q + 12 M + 16 P_out,
        # 32 bytes per output cell with the unique copies)
        rt.record(e0, stream)
        for _ in range(reps):
            hip.count_block_dev(h, d_k, d_i, n, total, 0, flags, d_c, d_o, d_uc, d_uo, stream)
        rt.record(e1, stream)
        rt.sync(stream)
        count_ms = rt.elapsed_ms(e0, e1) / reps
        M = hip.matched_records(h, d_k, d_i, n, total)
        algo = 8 * total + 12 * M + (32 if unique else 16) * cells
        prof, prof_src = c4_profile(tiling) if unique else (None, None)
        traffic = int(prof["traffic_bytes_per_launch"]) if prof and "traffic_bytes_per_launch" in prof else None
        roof = {"bound": "hbm", "kernel": "k_qjoin (key-range items)" if tiling["mode"] == 3 else "count",
                "achieved": algo / (count_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": algo / (count_ms / 1e3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": algo,
                "bytes_model": "8*N_Q + 12*M + %d*P_out (dense output)" % (32 if unique else 16),
                "N_Q": total, "M": M, "P_out": cells, "count_ms": count_ms,
                "traffic": traffic, "traffic_source": prof_src,
                "traffic_GBs": traffic / (count_ms / 1e3) / 1e9 if traffic else None,
                "line_requests_per_launch": prof.get("dram_read_requests_per_launch") if prof else None,
                "lines_per_key": prof["dram_read_requests_per_launch"] / total
                if prof and "dram_read_requests_per_launch" in prof else None}
        # oracle on a bounded prefix (fenceposts rebuilt per block, as the reference)
        if orc is None:
            sys.path.insert(0, str(ROOT / "tests"))
            import kssdtest as kt
            orc = kt.Oracle()
        ncpu = 1
        while True:
            c_ctx = np.zeros((ncpu, nref, 2), np.uint32)
            c_obj = np.zeros((ncpu, nref, 2), np.uint32)
            c_uc = np.zeros_like(c_ctx)
            c_uo = np.zeros_like(c_obj)
            sub = np.ascontiguousarray(keys[: int(idx[ncpu])])
            t = time.perf_counter()
            orc.lib.orc_count_block(kt._p(rec), len(rec), nref, 20, kt._p(sub), kt._p(idx[: ncpu + 1]), ncpu,
                                    0, flags, kt._p(c_ctx), kt._p(c_obj), kt._p(c_uc), kt._p(c_uo), threads)
            cdt = time.perf_counter() - t
            if cdt > args.cpu_budget / 6 or ncpu >= n:
                break
            ncpu = min(n, ncpu * 4)
        g_ctx = rt.download(d_c, np.uint32, 2 * ncpu * nref).reshape(ncpu, nref, 2)
        g_obj = rt.download(d_o, np.uint32, 2 * ncpu * nref).reshape(ncpu, nref, 2)
        exact = bool(np.array_equal(g_ctx, c_ctx) and np.array_equal(g_obj, c_obj))
        if unique:
            exact = exact and bool(np.array_equal(rt.download(d_uc, np.uint32, 2 * ncpu * nref).reshape(ncpu, nref, 2),
                                                  c_uc))
        cpu_pps = ncpu * nref / cdt
        plain_pps = 1.0 / plain_s_per_pair if plain_s_per_pair else None
        out[name] = {**out.get(name, {}), "queries": n, "query_entries": total, "pairs": cells, "ms": ms,
                     "pairs_per_s": pairs_per_s,
                     "over_plain": pairs_per_s / plain_pps if plain_pps else None, "cpu_pairs_per_s": cpu_pps,
                     "cpu_queries": ncpu, "over_cpu": pairs_per_s / cpu_pps, "counts_bit_exact_on_cpu_sample": exact,
                     "flags": "unique-ctx" if unique else "none", "tiling": tiling, "roofline": roof}
        log(f"index bench {name}: {n} x {nref} in {ms:.2f} ms = {pairs_per_s:.3e} pairs/s "
            f"({pairs_per_s / plain_pps if plain_pps else 0:.0f}x plain, {pairs_per_s / cpu_pps:.0f}x CPU on "
            f"{ncpu} queries), bit-exact {exact}; count {count_ms:.3f} ms, {roof['achieved']:.0f} GB/s algorithmic")
        for d in (d_k, d_i, d_qc, d_rc, d_c, d_o, d_out, d_uc, d_uo):
            if d:
                rt.free(d)
        hip.set_tiling(h, mode=0)
    if args.index_stream_queries and "assemblies" in names:
        out["streamed"] = stream_bench(args, host, hip, rt, h, shared["qk"], shared["qi"], nref, rctx)
    out["cpu_threads"] = threads
    return out


def stream_bench(args, host, hip, rt, h, qk, qi, nref, rctx):
    """The C3 queries with their keys in host memory (PCIe-inclusive): in
    reference-sized blocks of 4,096 queries, each block's keys go over PCIe
    from pinned memory on a copy stream, double-buffered against the select
    (count + pass gate + FP64 values, graph-captured per buffer and block
    size) on the compute stream; the passing pairs come back to the host.
    Timed from the first copy to the last rows on the host."""
    nq_all = min(args.index_stream_queries, len(qi) - 1)
    bq = 4096
    blocks = [(b, min(nq_all, b + bq)) for b in range(0, nq_all, bq)]
    maxkeys = max(int(qi[e] - qi[b]) for b, e in blocks)
    pin = rt.host_malloc(8 * int(qi[nq_all]))
    ctypes.memmove(pin, qk.ctypes.data, 8 * int(qi[nq_all]))
    qctx = host.ctx_counts(qk[: int(qi[nq_all])], qi[: nq_all + 1], 20)
    d_keys = [rt.malloc(8 * maxkeys), rt.malloc(8 * maxkeys)]
    d_idx = [rt.upload((qi[b:e + 1] - qi[b]).astype(np.uint64)) for b, e in blocks]
    d_qc = [rt.upload(qctx[b:e]) for b, e in blocks]
    d_rc = rt.upload(rctx)
    cap = 1 << 20
    d_rows = [rt.malloc(56 * cap), rt.malloc(56 * cap)]
    d_n = [rt.malloc(4), rt.malloc(4)]
    d_first, d_cnt = rt.malloc(4 * bq), rt.malloc(4 * bq)
    h_rows = [rt.host_malloc(56 * cap), rt.host_malloc(56 * cap)]
    h_n = [rt.host_malloc(4), rt.host_malloc(4)]
    value = kh.EpilogueOpts(3, 1, 0, 1, 0, 0, 8, 20, 44, 0.5, 1.0)
    sopts = kh.SelectOpts(3, 0.5, 0.95, 0, value)
    cs, ks = rt.stream(), rt.stream()
    copied = [rt.event(), rt.event()]
    done = [rt.event(), rt.event()]

    def compute(buf, k):
        b, e = blocks[k]
        hip.select_block_dev(h, d_keys[buf], d_idx[k], e - b, int(qi[e] - qi[b]), 0, d_qc[k], d_rc, sopts,
                             d_rows[buf], cap, d_n[buf], d_first, d_cnt, ks)

    # eager warm-up sizes the workspace; graphs per (buffer, block)
    rt.copy_async(d_keys[0], ctypes.c_void_p(pin.value + 8 * int(qi[0])), 8 * int(qi[blocks[0][1]]), rt.H2D, ks)
    compute(0, 0)
    rt.sync(ks)
    graphs = {}
    for k in range(len(blocks)):
        for buf in (0, 1):
            graphs[(buf, k)] = rt.capture(ks, lambda: compute(buf, k))
    t0 = time.perf_counter()
    total_rows, overflow = 0, False

    def collect(k):
        nonlocal total_rows, overflow
        pb = k % 2
        rt.lib.hipEventSynchronize(done[pb])
        n = ctypes.cast(h_n[pb], ctypes.POINTER(ctypes.c_uint32))[0]
        overflow |= n > cap
        if n:
            rt.copy_async(h_rows[pb], d_rows[pb], 56 * min(n, cap), rt.D2H, cs)
        total_rows += n

    for k, (b, e) in enumerate(blocks):
        buf = k % 2
        if k >= 2:
            rt.wait_event(cs, done[buf])  # block k - 2 is done with this buffer
        rt.copy_async(d_keys[buf], ctypes.c_void_p(pin.value + 8 * int(qi[b])), 8 * int(qi[e] - qi[b]), rt.H2D, cs)
        rt.record(copied[buf], cs)
        rt.wait_event(ks, copied[buf])
        rt.graph_launch(graphs[(buf, k)][1], ks)
        rt.copy_async(h_n[buf], d_n[buf], 4, rt.D2H, ks)
        rt.record(done[buf], ks)
        if k >= 1:  # block k - 1's rows while block k runs
            collect(k - 1)
    collect(len(blocks) - 1)
    rt.sync(cs)
    rt.sync(ks)
    wall = time.perf_counter() - t0
    pairs = nq_all * nref
    key_bytes = 8 * int(qi[nq_all])
    res = {"queries": nq_all, "blocks": len(blocks), "block_queries": bq, "pairs": pairs, "wall_s": wall,
           "pairs_per_s": pairs / wall, "h2d_bytes": key_bytes, "h2d_GBs_sustained": key_bytes / wall / 1e9,
           "passing_pairs": int(total_rows), "rows_overflowed_cap": bool(overflow),
           "what": "PCIe-inclusive: distinct queries' keys streamed from pinned host memory in blocks of 4096, "
                   "double-buffered; select (count + pass gate + FP64 values) graph-captured per block; passing "
                   "pairs copied back"}
    log(f"streamed: {nq_all} queries x {nref} refs in {len(blocks)} blocks, {wall:.2f} s = {pairs / wall:.3e} "
        f"pairs/s incl. {key_bytes / 1e9:.1f} GB of H2D keys; {total_rows} passing pairs")
    for g, ex in graphs.values():
        rt.graph_destroy(g, ex)
    for p in (*d_keys, *d_idx, *d_qc, d_rc, *d_rows, *d_n, d_first, d_cnt):
        rt.free(p)
    for p in (pin, *h_rows, *h_n):
        rt.host_free(p)
    return res


def cli_e2e(args, host, shared, threads):
    """`kssd3a-hip ani -r REF -q QRY` end to end on the C3 sketches written as
    sketch directories (the reference index with its sorted index file, the
    first args.cli_queries queries): the drop-in CLI reads the files, picks
    the index path, runs the device pass gate block by block and ranks and
    prints on the host with the reference's defaults.  The printed pairs must
    be among the device gate's passing pairs of the headline step (the host
    adds the selected-ANI cut, command_ani.c:3517)."""
    import shutil
    import subprocess
    import tempfile
    nq = min(args.cli_queries, len(shared["qi"]) - 1)
    qi = shared["qi"][: nq + 1]
    qk = shared["qk"][: int(qi[-1])]
    rk, ri, rec, nref = shared["rk"], shared["ri"], shared["rec"], shared["nref"]
    need = 8 * (len(qk) + len(rk)) + 14 * len(rec) + (4 << 30)
    d = Path(tempfile.mkdtemp(dir=scratch_dir(need)))
    try:
        t = time.time()
        host.write_sketch(d / "ref", rk, ri, [f"ref{i:05d}" for i in range(nref)], records=rec)
        host.write_sketch(d / "qry", qk, qi, [f"qry{i:06d}" for i in range(nq)])
        write_s = time.time() - t
        argv = [str(kh.PKG / "kssd3a-hip"), "ani", "-r", str(d / "ref"), "-q", str(d / "qry"), "-p", str(threads),
                "-o", str(d / "out.tsv")]
        t = time.perf_counter()
        p = subprocess.run(argv, capture_output=True, text=True, env=dict(os.environ, KSSDHIP_TIMING="1"),
                           timeout=900)
        wall = time.perf_counter() - t
        if p.returncode:
            raise RuntimeError(f"kssd3a-hip ani: {p.stderr[-500:]}")
        stages = [ln for ln in p.stderr.splitlines() if ln.startswith("[kssd3a-hip")]
        lines = (d / "out.tsv").read_text().splitlines()
        rows = shared["rows"]
        keep = rows["query"] < nq
        gate = set(zip(rows["query"][keep].tolist(), rows["ref"][keep].tolist()))
        printed = {(int(c[0][3:]), int(c[1][3:])) for c in (ln.split("\t") for ln in lines[1:])}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    res = {"command": "kssd3a-hip ani -r REF -q QRY -p THREADS -o OUT (reference defaults: -s1 Best, ctxcut 3, "
                      "afcut 0.5, anicut 0.95)",
           "queries": nq, "refs": nref, "pairs": nq * nref, "wall_s": wall, "pairs_per_s": nq * nref / wall,
           "files_written_s": write_s, "printed_rows": len(printed),
           "printed_within_device_gate": printed <= gate, "device_gate_pairs": len(gate),
           "stages": stages[-12:]}
    log(f"CLI end to end: {nq} queries x {nref} refs in {wall:.2f} s = {nq * nref / wall:.3e} pairs/s; "
        f"{len(printed)} rows printed, within the device gate's {len(gate)}: {printed <= gate}")
    return res


def c1_bench(host, hip, rt, threads, seed):
    """C1: the all-vs-all self matrix of 64 genomes (8 clades of 8, 2 Mbp
    ancestors, members U(1%, 15%) substitutions), by the indexed path
    (lower-triangle counts against the self index + FP64 epilogue,
    command_ani.c:1657-1760) and the direct path (one merge per pair,
    command_ani.c:3314-3365), each on the GPU and in the oracle."""
    rng = np.random.default_rng(seed)
    n = 64
    specs = [kh.SimSpec(7000 + c, 2_000_000, 8000 + 8 * c + m, float(rng.uniform(0.01, 0.15)), 0.0, 0, 0.0)
             for c in range(8) for m in range(8)]
    k, i = host.simsketch(specs, threads=threads)
    rec = host.build_index(k, i, 20, threads=threads)
    cnt = np.diff(i).astype(np.uint32)
    total, cells = int(i[-1]), n * n
    h = hip.open_index(rec, n, 20)
    s = rt.stream()
    d_k, d_i, d_cnt = rt.upload(k), rt.upload(i), rt.upload(cnt)
    d_c, d_o, d_out = rt.malloc(8 * cells), rt.malloc(8 * cells), rt.malloc(8 * cells)
    opts = kh.EpilogueOpts(3, 1, 0, 0, 1, 0, 8, 20, 44, 0.5, 1.0)
    e0, e1 = rt.event(), rt.event()

    def indexed():
        hip.count_block_dev(h, d_k, d_i, n, total, 0, kh.LOWER, d_c, d_o, None, None, s)
        hip.epilogue_dev(d_c, d_o, n, n, d_cnt, d_cnt, opts, d_out, s)

    indexed()
    rt.record(e0, s)
    for _ in range(5):
        indexed()
    rt.record(e1, s)
    rt.sync(s)
    gpu_indexed_ms = rt.elapsed_ms(e0, e1) / 5
    g_ctx = rt.download(d_c, np.uint32, 2 * cells).reshape(n, n, 2)
    g_obj = rt.download(d_o, np.uint32, 2 * cells).reshape(n, n, 2)
    pairs = [(q, r) for q in range(n) for r in range(q)]
    d_pout = rt.malloc(20 * len(pairs))
    d_ws = rt.malloc(kh.PAIR_WORKSPACE_BYTES)
    hip.check(hip.lib.kssdhip_pair_workspace_init_dev(d_ws, s))

    def direct():
        for j, (q, r) in enumerate(pairs):
            hip.pair_dev(d_k.value + 8 * int(i[q]), int(cnt[q]), d_k.value + 8 * int(i[r]), int(cnt[r]), 20, False,
                         d_pout.value + 20 * j, d_ws, s)

    direct()
    rt.sync(s)
    rt.record(e0, s)
    direct()
    rt.record(e1, s)
    rt.sync(s)
    gpu_direct_ms = rt.elapsed_ms(e0, e1)
    pres = rt.download(d_pout, np.uint32, 5 * len(pairs)).reshape(len(pairs), 5)
    qq = np.array([q for q, _ in pairs])
    rr = np.array([r for _, r in pairs])
    direct_ok = bool(np.array_equal(pres[:, 0], g_ctx[qq, rr, 0]) and np.array_equal(pres[:, 4], g_obj[qq, rr, 0])
                     and np.array_equal(pres[:, 2], g_obj[qq, rr, 1]) and np.array_equal(pres[:, 3], g_ctx[qq, rr, 1]))
    sys.path.insert(0, str(ROOT / "tests"))
    import kssdtest as kt
    orc = kt.Oracle()
    c_ctx = np.zeros((n, n, 2), np.uint32)
    c_obj = np.zeros((n, n, 2), np.uint32)
    vals = np.zeros(cells, np.float64)
    t = time.perf_counter()
    orc.lib.orc_count_block(kt._p(rec), len(rec), n, 20, kt._p(k), kt._p(i), n, 0, kt.FLAGS["LOWER"], kt._p(c_ctx),
                            kt._p(c_obj), None, None, threads)
    orc.lib.orc_matrix_values(kt._p(c_ctx), kt._p(c_obj), n, n, kt._p(cnt), kt._p(cnt), 3, 1, 0.5, 1.0, 8, 20, 44,
                              kt._p(vals), threads)
    cpu_indexed_ms = (time.perf_counter() - t) * 1e3
    f = kt.Features()
    t = time.perf_counter()
    for q, r in pairs:
        orc.lib.orc_pair_features(kt._p(k[i[q]:i[q + 1]]), int(cnt[q]), kt._p(k[i[r]:i[r + 1]]), int(cnt[r]), 20, 0,
                                  ctypes.byref(f))
    cpu_direct_ms = (time.perf_counter() - t) * 1e3
    low = np.tril(np.ones((n, n), bool), -1)
    exact = bool(np.array_equal(g_ctx[low], c_ctx[low]) and np.array_equal(g_obj[low], c_obj[low]))
    for p_ in (d_k, d_i, d_cnt, d_c, d_o, d_out, d_pout, d_ws):
        rt.free(p_)
    hip.close_index(h)
    res = {"genomes": n, "pairs": len(pairs), "entries": total,
           "gpu_indexed_ms": gpu_indexed_ms, "gpu_direct_ms": gpu_direct_ms,
           "cpu_indexed_ms": cpu_indexed_ms, "cpu_direct_ms": cpu_direct_ms, "cpu_threads": threads,
           "cpu_direct_threads": 1,
           "what": "indexed: lower-triangle counts against the self index + FP64 CtxMoE epilogue (one launch each); "
                   "direct: one merge launch per pair (GPU) / one oracle merge per pair (CPU, one thread)",
           "counts_bit_exact": exact, "direct_equals_indexed": direct_ok}
    log(f"C1: indexed {gpu_indexed_ms:.3f} ms GPU vs {cpu_indexed_ms:.1f} ms CPU; direct {gpu_direct_ms:.1f} ms GPU "
        f"vs {cpu_direct_ms:.1f} ms CPU; bit-exact {exact}, direct == indexed {direct_ok}")
    return res


def serve_bench(args, host, hip, threads, plain_s_per_pair):
    """Serving loop under Poisson load against a large resident index.

    Index: args.serve_refs genomes = clades of 20, each clade a random
    ancestor (lognormal length as above) whose members carry U(1%, 15%)
    substitutions.  Requests: 70% assemblies (a new member of a random clade)
    and 30% read sets (drawn from 64 read-set sketches of reference genomes
    plus 0.5% substitutions, read as 2 x 150 bp pairs at U(5, 30)x and
    sketched with the package's `sketch --asone --conflict -A`, see
    real_readsets), arriving as a Poisson stream at args.serve_rate per
    second.  The server returns results: each batch runs the pass gate
    (ctxcut 3, afcut 0.2, anicut 0.95) and the FP64 CtxMoE ANI on the device
    and each request gets its ranked passing rows.  Latency = arrival ->
    ranked hits on the host, measured per request.  CPU baselines per
    request: the oracle's restatement of the reference's default one-query
    path (stream_ref_sketches_one_qraw_lookup) and its forced-index path;
    the plain kernel's per-query figure is its measured per-pair time times
    the index size."""
    import threading
    rng = np.random.default_rng(args.seed + 100)
    nref = args.serve_refs - args.serve_refs % 20
    nclade = nref // 20
    specs, cseed, clen = clade_specs(rng, nref, args.seed)
    t = time.time()
    rk, ri = host.simsketch(specs, threads=threads)
    rec = host.build_index(rk, ri, 20, threads=threads)
    rctx = host.ctx_counts(rk, ri, 20)
    log(f"serving index: {nref} genomes ({nclade} clades of 20), {len(rec)} records, built in "
        f"{time.time() - t:.1f}s")
    h = hip.open_index(rec, nref, 20)
    nq = args.serve_queries
    is_asm = rng.random(nq) < 0.7
    asm_specs = []
    for i in range(nq):
        c = int(rng.integers(0, nclade))
        asm_specs.append(kh.SimSpec(cseed[c], int(clen[c]), int(args.seed * 3_000_017 + i + 1),
                                    float(rng.uniform(0.01, 0.15)), 0.0, 0, 0.0))
    qk, qi = host.simsketch([asm_specs[i] for i in range(nq) if is_asm[i]], threads=threads)
    n_rs = 64
    rs_specs = []
    for k in range(n_rs):  # a reference genome itself, then 0.5% sample substitutions
        r = specs[int(rng.integers(0, nref))]
        rs_specs.append(kh.SimSpec(r.anc_seed, r.anc_len, int(args.seed * 7_000_001 + k + 1), 0.005, 0.0,
                                   r.mut_seed, r.sub_rate))
    t = time.time()
    sk, si, _ = real_readsets(host, rs_specs, rng, threads, args.seed + 100, cov_range=(5.0, 30.0))
    log(f"serving read sets: {n_rs} sketched from simulated reads in {time.time() - t:.1f}s")
    queries, a = [], 0
    pick = rng.integers(0, n_rs, size=nq)
    for i in range(nq):
        if is_asm[i]:
            queries.append(qk[qi[a]:qi[a + 1]])
            a += 1
        else:
            queries.append(sk[si[pick[i]]:si[pick[i] + 1]])
    sizes = np.array([len(q) for q in queries])
    log(f"serving requests: {nq} ({int(is_asm.sum())} assemblies, median {int(np.median(sizes[is_asm]))} keys; "
        f"{int((~is_asm).sum())} read sets, median {int(np.median(sizes[~is_asm])) if (~is_asm).any() else 0} keys)")
    budget = 2 << 30
    value = kh.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, 20, 44, 0.2, 1.0)
    sel = kh.SelectOpts(3, 0.2, 0.95, 0, value)
    srv = hip.serve_open(h, budget_bytes=budget, slots=4, max_batch=32, max_queued=1 << 20, select_opts=sel,
                         ref_ctx=rctx)
    # warm the server (first launches load kernels)
    for q in queries[:8]:
        hip.serve_wait_rows(srv, hip.serve_submit(srv, q))
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, size=nq))
    tickets = [None] * nq
    late = np.zeros(nq)
    t0 = time.perf_counter() + 0.05

    def producer():
        for i in range(nq):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            late[i] = max(0.0, time.perf_counter() - (t0 + arrivals[i]))
            tickets[i] = hip.serve_submit(srv, queries[i])

    th = threading.Thread(target=producer)
    th.start()
    lat = np.zeros(nq)
    bsz = np.zeros(nq, np.int64)
    hits = np.zeros(nq, np.int64)
    for i in range(nq):
        while tickets[i] is None:
            time.sleep(0.0002)
        if tickets[i] < 0:
            raise RuntimeError(f"serve_submit failed: {tickets[i]}")
        rows, st = hip.serve_wait_rows(srv, tickets[i], cap=4096)
        lat[i] = st.total_ms + late[i] * 1e3
        bsz[i] = st.batch_queries
        hits[i] = len(rows)
    t_end = time.perf_counter()
    th.join()
    info = hip.serve_info(srv)
    wall = t_end - t0
    # saturation: every request submitted at once, so admission fills batches
    tb = time.perf_counter()
    burst_t = [hip.serve_submit(srv, q) for q in queries]
    blat, bbsz = np.zeros(nq), np.zeros(nq, np.int64)
    for i, tk in enumerate(burst_t):
        _, st = hip.serve_wait_rows(srv, tk, cap=4096)
        blat[i], bbsz[i] = st.total_ms, st.batch_queries
    burst_wall = time.perf_counter() - tb
    hip.serve_close(srv)
    # CPU: the reference's forced-index per-query path (fenceposts over the
    # whole index rebuilt per request, then the lookups), on 2 requests
    sys.path.insert(0, str(ROOT / "tests"))
    import kssdtest as kt
    orc = kt.Oracle()
    cpu_ms, cpu_stream_ms = [], []
    for i in (int(np.flatnonzero(is_asm)[0]), int(np.flatnonzero(~is_asm)[0]) if (~is_asm).any() else 1):
        q = queries[i]
        ctx = np.zeros((1, nref, 2), np.uint32)
        obj = np.zeros((1, nref, 2), np.uint32)
        t = time.perf_counter()
        orc.lib.orc_count_block(kt._p(rec), len(rec), nref, 20, kt._p(q),
                                kt._p(np.array([0, len(q)], np.uint64)), 1, 0, 0, kt._p(ctx), kt._p(obj),
                                None, None, threads)
        cpu_ms.append((time.perf_counter() - t) * 1e3)
        t = time.perf_counter()
        orc.stream_one_query(q, rk, ri, ctxcut=3, afcut=0.2, threads=threads)
        cpu_stream_ms.append((time.perf_counter() - t) * 1e3)
    del rk
    hip.close_index(h)
    res = {
        "index_genomes": nref, "index_records": int(len(rec)), "requests": nq,
        "arrival_rate_per_s": args.serve_rate, "mix": "70% assemblies / 30% read sets",
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "p99_assembly": float(np.percentile(lat[is_asm], 99)),
                       "p99_readset": float(np.percentile(lat[~is_asm], 99)) if (~is_asm).any() else None,
                       "max": float(lat.max()),
                       "what": "arrival -> ranked hits on the host (device pass gate ctxcut 3 / afcut 0.2 / anicut "
                               "0.95 + FP64 CtxMoE ANI per batch, passing rows ranked per request)"},
        "hits_per_request": {"mean": float(hits.mean()), "max": int(hits.max()),
                             "requests_with_hits": int((hits > 0).sum())},
        "read_set_requests": "drawn from 64 read-set sketches of simulated 2x150 bp reads (U(5,30)x, 1% errors)",
        "sustained_pairs_per_s": nq * nref / wall,
        "batches": int(info.batches), "mean_batch": float(bsz.mean()), "max_batch_seen": int(info.max_batch_seen),
        "saturated": {"what": "all requests submitted at once", "sustained_pairs_per_s": nq * nref / burst_wall,
                      "queries_per_s": nq / burst_wall, "mean_batch": float(bbsz.mean()),
                      "latency_ms_p50": float(np.percentile(blat, 50)),
                      "latency_ms_p99": float(np.percentile(blat, 99))},
        "budget_bytes": budget, "slots": 4,
        "plain_kernel_per_query_ms": plain_s_per_pair * nref * 1e3,
        "cpu_forced_index_per_query_ms": float(np.mean(cpu_ms)),
        "cpu_default_path_per_query_ms": float(np.mean(cpu_stream_ms)),
        "cpu_default_path_what": "oracle restatement of stream_ref_sketches_one_qraw_lookup (query hash, every "
                                 "reference streamed) on one assembly and one read-set request",
        "cpu_threads": threads,
    }
    log(f"serving: p50 {res['latency_ms']['p50']:.2f} ms, p99 {res['latency_ms']['p99']:.2f} ms, "
        f"{res['sustained_pairs_per_s']:.3e} pairs/s over {wall:.1f}s, mean batch {bsz.mean():.2f}; "
        f"plain per query {res['plain_kernel_per_query_ms']:.0f} ms, CPU per query {np.mean(cpu_ms):.0f} ms "
        f"(default path {np.mean(cpu_stream_ms):.0f} ms); "
        f"saturated {res['saturated']['sustained_pairs_per_s']:.3e} pairs/s, mean batch {bbsz.mean():.1f}")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nref", type=int, default=1000)
    ap.add_argument("--nqry", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--plain-pairs", type=int, default=4000)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--threads", type=int, default=0, help="host threads (default: all)")
    ap.add_argument("--retune", action="store_true", help="ignore the tuning cache")
    ap.add_argument("--edges-genomes", type=int, default=0, metavar="GENOMES",
                    help="only time `matrix --format edges` on GENOMES clade genomes (GPU vs the reference "
                         "binary) and exit")
    ap.add_argument("--edges-cut", type=float, default=0.05)
    ap.add_argument("--readsets-bench", type=int, default=0, metavar="GENOMES",
                    help="only run the read-set pipeline on files (sketch, index, read-set sketches, "
                         "--qraw --unique-ctx --estimate-coverage) against the reference binary and exit")
    ap.add_argument("--readsets-samples", type=int, default=6)
    ap.add_argument("--index-build-only", type=int, default=0, metavar="GENOMES",
                    help="only time the index build of GENOMES clade genomes (host builder vs GPU) and exit")
    ap.add_argument("--tiling", default="",
                    help="fixed join tiling MODE,CHUNK,BLOCK[,GROUP], or 3,THREADS[,SLOTS] for the query-major "
                         "join, instead of the autotuner (for profiling runs)")
    ap.add_argument("--latency-queries", type=int, default=64)
    ap.add_argument("--index-refs", type=int, default=10000,
                    help="reference genomes of the large-index bench (500 clades of 20); 0 skips it")
    ap.add_argument("--index-queries", type=int, default=4096)
    ap.add_argument("--index-readsets", type=int, default=200)
    ap.add_argument("--index-stream-queries", type=int, default=100000,
                    help="C3 queries streamed from host memory against the resident index in blocks (0: skip)")
    ap.add_argument("--c3-queries", type=int, default=100000,
                    help="distinct query genomes of the headline workload (C3)")
    ap.add_argument("--c3-cpu-queries", type=int, default=4096,
                    help="C3 queries timed on the CPU oracle (the first reference block)")
    ap.add_argument("--c3-only", action="store_true",
                    help="run only the headline workload (for profiling runs)")
    ap.add_argument("--data-cache", default="",
                    help="directory to keep the C3 sketches in between runs (profiling passes)")
    ap.add_argument("--c2-only", action="store_true", help="only the C2 step (for profiling runs)")
    ap.add_argument("--c4-only", action="store_true",
                    help="only the C4 count (200 read sets vs the 10k index; for profiling runs)")
    ap.add_argument("--cli-queries", type=int, default=100000,
                    help="C3 queries run end to end through kssd3a-hip ani (0 = skip)")
    ap.add_argument("--serve-refs", type=int, default=20000,
                    help="reference genomes of the serving index (1000 clades of 20); 0 skips serving")
    ap.add_argument("--serve-queries", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrivals per second")
    ap.add_argument("--copy-probe-mib", type=int, default=256,
                    help="size of a device-to-device copy run once before the timed region; its "
                         "HBM counters calibrate the join kernel's (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this path shards "
              "by reference gid range across GPUs; only the one-GPU kernel is built)", file=sys.stderr)
        sys.exit(2)
    threads = args.threads or min(16, os.cpu_count() or 1)
    host = kh.Host()
    if args.edges_genomes:
        print(json.dumps({"matrix_edges": edges_bench(args, host, threads)}), flush=True)
        return
    if args.readsets_bench:
        print(json.dumps({"readset_pipeline": readset_bench(args, host, threads)}), flush=True)
        return
    if args.index_build_only:
        # the index build alone: clade genomes as in the index bench
        rng = np.random.default_rng(args.seed + 200)
        nref = args.index_build_only - args.index_build_only % 20
        specs, _, _ = clade_specs(rng, nref, args.seed + 200)
        rk, ri = host.simsketch(specs, threads=threads)
        _, build = index_build_bench(host, kh.KssdHip(0), rk, ri, threads, reps=3)
        print(json.dumps({"index_build": build, "index_genomes": nref}), flush=True)
        return
    if args.c4_only:  # profiling runs of the C4 count alone
        shared = c4_index(args, host, threads)
        out = index_bench(args, host, threads, None, shared, names=("read_sets",))
        shared["hip"].close_index(shared["h"])
        print(json.dumps({"c4": out["read_sets"]}), flush=True)
        return
    if args.c2_only:  # profiling runs of the C2 step alone
        c3, shared, headline = None, None, {}
    else:
        c3, shared = c3_bench(args, host, threads)
    headline = {} if args.c2_only else {
        "metric": "ANI pairs/sec on 1 MI355X, batched vs plain kernel; p50/p99 query latency",
        "value": c3["value"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": c3["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 counts, f32 pass gate, f64 ANI",
        "data": "synthetic genomes (seeded simulator) sketched at -f8 by the package sketcher",
        "config": {"workload": f"C3: {c3['refs']}-genome clade index (500 clades of 20, lognormal median 4 Mbp, "
                               f"members U(1%,15%) substitutions from the ancestor) x {c3['queries']} distinct query "
                               "genomes (new members of random clades); every pair counted, pass gate ctxcut 3 / "
                               "afcut 0.5 / anicut 0.95, passing pairs with FP64 CtxMoE ANI",
                   "pairs": c3["pairs"], "queries": c3["queries"], "refs": c3["refs"],
                   "query_entries": c3["query_entries"], "index_records": c3["index_records"],
                   "output": "passing (query, ref) pairs: counts + ranking ANI + FP64 CtxMoE ANI"},
        "roofline": c3["roofline"], "lds": c3["lds"], "cpu_baseline": c3["cpu_baseline"],
        "gpu_over_cpu": c3["value"] / c3["cpu_baseline"]["value"],
        "c3": c3,
    }
    if args.c3_only:
        shared["hip"].close_index(shared["h"])
        print(json.dumps(headline), flush=True)
        return
    rkeys, ridx, qkeys, qidx = make_workload(args.nref, args.nqry, args.seed, threads, host)
    obj_bits = 20
    t = time.time()
    records = host.build_index(rkeys, ridx, obj_bits, threads=threads)
    log(f"index built: {len(records)} records in {time.time() - t:.1f}s")
    nref, nqry = args.nref, args.nqry
    qctx = host.ctx_counts(qkeys, qidx, obj_bits)
    rctx = host.ctx_counts(rkeys, ridx, obj_bits)

    # ---------------- GPU setup: everything resident before timing
    rt = kh.Hip(0)
    hip = kh.KssdHip(0)
    h = hip.open_index(records, nref, obj_bits)
    stream = rt.stream()
    d_qkeys = rt.upload(qkeys)
    d_qidx = rt.upload(qidx)  # block starts at 0: offsets are already relative
    d_qctx, d_rctx = rt.upload(qctx), rt.upload(rctx)
    cells = nqry * nref
    d_ctx, d_obj, d_out = rt.malloc(8 * cells), rt.malloc(8 * cells), rt.malloc(8 * cells)
    opts = kh.EpilogueOpts(3, 1, 0, 0, 0, 0, 8, obj_bits, 44, 0.5, 1.0)
    total = int(qidx[-1])
    rt.lib.hipMemsetAsync.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]

    def step(ev_join=None):
        rt.check(rt.lib.hipMemsetAsync(d_ctx, 0, 8 * cells, stream))
        rt.check(rt.lib.hipMemsetAsync(d_obj, 0, 8 * cells, stream))
        if ev_join:
            rt.record(ev_join[0], stream)
        hip.count_block_dev(h, d_qkeys, d_qidx, nqry, total, 0, kh.ACCUMULATE, d_ctx, d_obj,
                            None, None, stream)
        if ev_join:
            rt.record(ev_join[1], stream)
        hip.epilogue_dev(d_ctx, d_obj, nqry, nref, d_qctx, d_rctx, opts, d_out, stream)

    t = time.time()
    if args.tiling:
        v = [int(x) for x in args.tiling.split(",")]
        if v[0] == 3:
            tiling = {"mode": 3, "query_threads": v[1], "query_slots": v[2] if len(v) > 2 else 0}
        else:
            tiling = {"mode": v[0], "chunk_records": v[1], "block_threads": v[2],
                      "group_partitions": v[3] if len(v) > 3 else 1}
        hip.set_tiling(h, **tiling)
        tune_ms = {}
        log(f"fixed tiling {tiling}")
    else:
        tiling, tune_ms = tune.tune(hip, rt, h, d_qkeys, d_qidx, nqry, total, nref, len(records), d_ctx, d_obj,
                                    stream, flags=0, force=args.retune, log=log)
        log(f"autotuned tiling {tiling} in {time.time() - t:.1f}s")
    if args.copy_probe_mib:
        nb = args.copy_probe_mib << 20
        a_, b_ = rt.malloc(nb), rt.malloc(nb)
        rt.check(rt.lib.hipMemcpy(b_, a_, nb, rt.D2D))
        rt.free(a_)
        rt.free(b_)
    for _ in range(args.warmup):
        step()
    rt.sync(stream)
    evs = [(rt.event(), rt.event()) for _ in range(args.steps)]
    e0, e1 = rt.event(), rt.event()
    rt.sync(stream)
    t0 = time.perf_counter()
    rt.record(e0, stream)
    for k in range(args.steps):
        step(evs[k])
    rt.record(e1, stream)
    rt.sync(stream)
    wall = time.perf_counter() - t0
    total_ms = rt.elapsed_ms(e0, e1)
    join_ms = [rt.elapsed_ms(a, b) for a, b in evs]
    ms_per_step = total_ms / args.steps
    pairs_per_s = cells / (ms_per_step / 1e3)
    log(f"batched: {ms_per_step:.3f} ms/step ({wall / args.steps * 1e3:.3f} ms wall), "
        f"join {np.mean(join_ms):.3f} ms, {pairs_per_s:.3e} pairs/s")

    # ---------------- results back + parity on the CPU-baseline sample
    ctx = rt.download(d_ctx, np.uint32, 2 * cells).reshape(nqry, nref, 2).copy()
    obj = rt.download(d_obj, np.uint32, 2 * cells).reshape(nqry, nref, 2).copy()
    ani = rt.download(d_out, np.float64, cells).reshape(nqry, nref).copy()

    # ---------------- the same step captured in a HIP graph (counter zeroing by
    # fill kernels inside the counting call, no memset nodes)
    def graph_step():
        hip.count_block_dev(h, d_qkeys, d_qidx, nqry, total, 0, 0, d_ctx, d_obj, None, None, stream)
        hip.epilogue_dev(d_ctx, d_obj, nqry, nref, d_qctx, d_rctx, opts, d_out, stream)

    graph_step()
    rt.sync(stream)
    g, gx = rt.capture(stream, graph_step)
    for _ in range(args.warmup):
        rt.graph_launch(gx, stream)
    rt.sync(stream)
    rt.record(e0, stream)
    for _ in range(args.steps):
        rt.graph_launch(gx, stream)
    rt.record(e1, stream)
    graph_ms = rt.elapsed_ms(e0, e1) / args.steps
    graph_ok = bool(np.array_equal(rt.download(d_ctx, np.uint32, 2 * cells).reshape(nqry, nref, 2), ctx)
                    and np.array_equal(rt.download(d_obj, np.uint32, 2 * cells).reshape(nqry, nref, 2), obj)
                    and np.array_equal(rt.download(d_out, np.float64, cells).reshape(nqry, nref), ani,
                                       equal_nan=True))
    rt.graph_destroy(g, gx)
    log(f"graph: {graph_ms:.3f} ms/step, agrees with eager: {graph_ok}")
    p_touched = int((ctx[:, :, 0] > 0).sum())

    # ---------------- the persistent query-major form on the same block
    hip.set_tiling(h, mode=3)
    hip.count_block_dev(h, d_qkeys, d_qidx, nqry, total, 0, 0, d_ctx, d_obj, None, None, stream)
    rt.record(e0, stream)
    for _ in range(3):
        hip.count_block_dev(h, d_qkeys, d_qidx, nqry, total, 0, 0, d_ctx, d_obj, None, None, stream)
    rt.record(e1, stream)
    rt.sync(stream)
    qmajor_ms = rt.elapsed_ms(e0, e1) / 3
    qmajor_ok = bool(np.array_equal(rt.download(d_ctx, np.uint32, 2 * cells).reshape(nqry, nref, 2), ctx)
                     and np.array_equal(rt.download(d_obj, np.uint32, 2 * cells).reshape(nqry, nref, 2), obj))
    hip.set_tiling(h, **tiling)
    forms_c2 = {"static_partitioned": {"ms": float(np.mean(join_ms)), "tiling": tiling},
                "persistent_query_major": {"ms": qmajor_ms, "tiling": {"mode": 3}, "counts_equal": qmajor_ok},
                "what": "the counting call alone (split + join, or the query-major join) on the C2 block"}
    log(f"C2 forms: static partitioned {np.mean(join_ms):.3f} ms, persistent query-major {qmajor_ms:.3f} ms "
        f"(equal {qmajor_ok})")

    # ---------------- per-query latency of the batched path (one query per launch);
    # auto mode sends such a small block to the lookup kernel
    hip.set_tiling(h, mode=0)
    lat = []
    sample_q = np.random.default_rng(1).choice(nqry, size=min(args.latency_queries, nqry), replace=False)
    ea, eb = rt.event(), rt.event()
    for q in sample_q:
        nk = int(qidx[q + 1] - qidx[q])
        d_sk = ctypes.c_void_p(d_qkeys.value + 8 * int(qidx[q]))
        d_one = rt.upload(np.array([0, nk], np.uint64))
        rt.record(ea, stream)
        hip.count_block_dev(h, d_sk, d_one, 1, nk, 0, 0, d_ctx, d_obj, None, None, stream)
        hip.epilogue_dev(d_ctx, d_obj, 1, nref, ctypes.c_void_p(d_qctx.value + 4 * int(q)), d_rctx,
                         opts, d_out, stream)
        rt.record(eb, stream)
        lat.append(rt.elapsed_ms(ea, eb))
        rt.free(d_one)
    lat = np.array(lat)

    # ---------------- plain kernel: one launch per pair
    rng = np.random.default_rng(3)
    npl = min(args.plain_pairs, cells)
    pq = rng.integers(0, nqry, size=npl)
    pr = rng.integers(0, nref, size=npl)
    d_rkeys = rt.upload(rkeys)
    d_pout = rt.malloc(20 * npl)
    d_ws = rt.malloc(kh.PAIR_WORKSPACE_BYTES)
    hip.check(hip.lib.kssdhip_pair_workspace_init_dev(d_ws, stream))

    def plain(i):
        q, r = int(pq[i]), int(pr[i])
        hip.pair_dev(d_qkeys.value + 8 * int(qidx[q]), int(qidx[q + 1] - qidx[q]),
                     d_rkeys.value + 8 * int(ridx[r]), int(ridx[r + 1] - ridx[r]), obj_bits, False,
                     d_pout.value + 20 * i, d_ws, stream)

    for i in range(min(50, npl)):
        plain(i)
    rt.sync(stream)
    rt.record(e0, stream)
    for i in range(npl):
        plain(i)
    rt.record(e1, stream)
    plain_ms = rt.elapsed_ms(e0, e1)
    plain_pairs_per_s = npl / (plain_ms / 1e3)
    pres = rt.download(d_pout, np.uint32, 5 * npl).reshape(npl, 5)
    plain_ok = bool(np.array_equal(pres[:, 0], ctx[pq, pr, 0]) and np.array_equal(pres[:, 4], obj[pq, pr, 0])
                    and np.array_equal(pres[:, 2], obj[pq, pr, 1]) and np.array_equal(pres[:, 3], ctx[pq, pr, 1]))
    log(f"plain: {plain_ms / npl * 1e3:.1f} us/pair, {plain_pairs_per_s:.3e} pairs/s, "
        f"agrees with batched: {plain_ok}")

    # ---------------- roofline of the join kernel
    M = matched_records(records, qkeys, obj_bits)
    algo_bytes = 8 * total + 12 * M + 16 * cells  # dense output: P_out = Q x R
    join_s = float(np.mean(join_ms)) / 1e3
    achieved = algo_bytes / join_s / 1e9
    traffic, traffic_src = measured_traffic(tiling, nref, nqry)
    H = matched_events(records, qkeys, obj_bits)
    lds, lds_src = measured_lds(tiling, nref, nqry)
    lds_peak = 256 * 128 * 2.4e9 / 1e9  # GB/s: 256 CUs x 128 B/clk x 2.4 GHz

    # ---------------- CPU baseline (oracle) + parity on its sample
    nq_cpu, cpu_s, cctx, cobj, cout = cpu_baseline(records, nref, obj_bits, qkeys, qidx, qctx, rctx,
                                                   args.cpu_budget, threads)
    counts_exact = bool(np.array_equal(cctx, ctx[:nq_cpu]) and np.array_equal(cobj, obj[:nq_cpu]))
    g = ani[:nq_cpu].ravel()
    rel = np.abs(g - cout) / np.maximum(np.maximum(np.abs(g), np.abs(cout)), 1e-300)
    max_rel = float(rel.max()) if len(rel) else 0.0
    cpu_pairs_per_s = nq_cpu * nref / cpu_s
    log(f"cpu oracle: {nq_cpu} queries x {nref} refs in {cpu_s:.2f}s on {threads} threads "
        f"= {cpu_pairs_per_s:.3e} pairs/s; counts bit-exact {counts_exact}, max rel ANI err {max_rel:.2e}")

    # ---------------- serving loop on a large resident index
    big_index = cli = c1 = serving = None
    if not args.c2_only:
        big_index = index_bench(args, host, threads, plain_ms / npl / 1e3, shared)
        cli = cli_e2e(args, host, shared, threads) if args.cli_queries else None
        c1 = c1_bench(host, hip, rt, threads, args.seed + 500)
        shared["hip"].close_index(shared["h"])
        del shared
        serving = serve_bench(args, host, hip, threads, plain_ms / npl / 1e3) if args.serve_refs else None

    eager_pairs_per_s = pairs_per_s
    if graph_ok:
        ms_per_step = graph_ms
        pairs_per_s = cells / (graph_ms / 1e3)
    c2 = {
        "what": "C2: batched ref x query with dense output, side figure",
        "value": pairs_per_s,
        "unit": "pairs/s",
        "ms_per_step": ms_per_step,
        "config": {
            "workload": f"batched ref x query, {nref} refs x {nqry} queries, lognormal median 4 Mbp "
                        "(sigma 0.5, clip 1-12 Mbp), queries = ref + 0.1-10% subs + 1% indels, -f8",
            "pairs": cells, "ref_entries": int(ridx[-1]), "query_entries": total,
            "index_records": len(records), "output": "dense counts + f64 CtxMoE ANI per pair",
        },
        "step": {"timed": ("HIP graph of one step (fill kernels + "
                           + {2: "k_split + k_pjoin", 3: "k_qjoin"}.get(tiling["mode"], "k_join")
                           + " + FP64 epilogue)") if graph_ok else "eager launches",
                 "graph_ms": graph_ms, "graph_agrees_with_eager": graph_ok,
                 "eager_ms": total_ms / args.steps, "eager_pairs_per_s": eager_pairs_per_s},
        "plain_kernel": {"pairs_per_s": plain_pairs_per_s, "sample_pairs": npl,
                         "us_per_pair": plain_ms / npl * 1e3, "agrees_with_batched": plain_ok},
        "batched_over_plain": pairs_per_s / plain_pairs_per_s,
        "tiling": tiling, "tuner_ms_per_block": tune_ms, "forms": forms_c2,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "what": f"one query ({len(lat)} sampled) vs the {nref}-genome resident index, "
                               "join + epilogue, device time"},
        "parity": {"counts_bit_exact": counts_exact, "queries_checked": nq_cpu,
                   "max_rel_ani_err": max_rel, "ani_tolerance": 1e-12},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "traffic_GBs": traffic / join_s / 1e9 if traffic else None,
                     "kernel": {2: "k_split+k_pjoin", 3: "k_qjoin (key-range items)"}.get(tiling["mode"], "k_join"),
                     "algorithmic_bytes": algo_bytes,
                     "bytes_model": "8*N_Q + 12*M + 16*P_out, P_out = Q*R (dense output)",
                     "N_Q": total, "M": M, "P_out": cells, "P_touched": p_touched,
                     "avg_launch_ms": float(np.mean(join_ms)),
                     "bound_by": counters_note(tiling)},
        "lds": {"events": H, "events_per_s": H / join_s,
                "events_what": "(query key, index record) pairs sharing a context, each one object-mismatch "
                               "minimum and one counter update",
                "bytes_per_launch": lds.get("lds_bytes_per_dispatch") if lds else None,
                "achieved_GBs": lds["lds_bytes_per_dispatch"] / join_s / 1e9 if lds else None,
                "peak_GBs": lds_peak,
                "frac": lds["lds_bytes_per_dispatch"] / join_s / 1e9 / lds_peak if lds else None,
                "lds_util_pct": lds.get("lds_util_pct") if lds else None,
                "bank_conflict": lds.get("lds_bank_conflict") if lds else None,
                "source": lds_src},
        "cpu_baseline": {"value": cpu_pairs_per_s, "unit": "pairs/s", "cores": threads, "kind": "port",
                         "sample": f"first {nq_cpu} queries x {nref} refs: oracle block count "
                                   "(fenceposts per block, guided OpenMP) + FP64 CtxMoE epilogue"},
        "gpu_over_cpu": pairs_per_s / cpu_pairs_per_s,
        "index_10k": big_index,
        "serving": serving,
    }
    if args.c2_only:
        hip.close_index(h)
        print(json.dumps({"c2": c2}), flush=True)
        return
    rec = dict(headline)
    rec["c2"] = c2
    rec["batched_over_plain"] = c3["value"] / c2["plain_kernel"]["pairs_per_s"]
    rec["latency_ms"] = c2["latency_ms"]
    rec["index_10k"] = c2.pop("index_10k")
    rec["serving"] = c2.pop("serving")
    rec["c1"] = c1
    rec["cli_end_to_end"] = cli
    hip.close_index(h)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
