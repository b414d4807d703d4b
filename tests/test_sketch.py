"""CPU tests of the host components: the sketcher against the reference's own
sketches of the same FASTA files (tests/golden/sketch, -f8, -f4, --conflict),
the sketching id, the inverted-index builder and the simulator."""
import numpy as np
import pytest

import kssdtest as kt

SK = kt.GOLDEN / "sketch"


@pytest.fixture(scope="module")
def host():
    return kt.Host()


@pytest.mark.parametrize("sub,drfold,conflict", [("f8", 8, False), ("f4", 4, False), ("f8c", 8, True)])
def test_sketcher_matches_reference(host, sub, drfold, conflict):
    ref = kt.read_sketch_dir(SK / sub)
    assert ref.drfold == drfold and bool(ref.conflict) == conflict
    for i, name in enumerate(ref.names):
        got = host.sketch_file(SK / name, drfold=drfold, keep_conflict=conflict)
        assert np.array_equal(got, ref.sample(i)), (sub, name)
    # coden mode hashes NUM_CODENS in the first slot (command_sketch_wrapper.c:924)
    assert host.lib.kssd_sketching_id(11, 0, 0, drfold) == ref.hash_id


def test_index_builder_matches_reference_file(host):
    g = kt.read_sketch_dir(kt.GOLDEN / "genomes" / "ref")
    out = host.build_index(g.keys, g.index, g.obj_bits, threads=3)
    assert out.tobytes() == g.records.tobytes()
    # and on conflict-rich fixtures against the independent numpy build
    for case in sorted((kt.GOLDEN / "counts").glob("case_*")):
        r = kt.read_sketch_dir(case / "ref")
        assert host.build_index(r.keys, r.index, r.obj_bits).tobytes() == r.records.tobytes()


def test_ctx_counts(host, oracle):
    c = kt.read_sketch_dir(kt.GOLDEN / "counts" / "case_09" / "ref")
    for ign in (0, 1):
        got = host.ctx_counts(c.keys, c.index, c.obj_bits, conflict=True, ignoreconflict=bool(ign))
        want = [oracle.lib.orc_ctx_runs(kt._p(c.sample(i)), len(c.sample(i)), c.obj_bits, ign)
                for i in range(c.n)]
        assert list(got) == want
        wi = np.zeros(c.n, np.uint32)
        oracle.lib.orc_ref_ctx_counts_from_index(kt._p(c.records), len(c.records), c.n,
                                                 kt._p(c.index), ign, kt._p(wi))
        assert list(got) == list(wi)


def test_simulator_is_deterministic_and_mutation_scaled(host):
    specs = [kt.SimSpec(11, 300_000, 0, 0.0, 0.0, 0, 0.0),
             kt.SimSpec(11, 300_000, 5, 0.01, 0.0, 0, 0.0),
             kt.SimSpec(11, 300_000, 6, 0.05, 0.01, 0, 0.0)]
    a, ia = host.simsketch(specs, threads=1)
    b, ib = host.simsketch(specs, threads=3)
    assert np.array_equal(a, b) and np.array_equal(ia, ib)
    # expected ~L/256 entries at -f8 and sorted unique keys per sample
    n0 = int(ia[1] - ia[0])
    assert 900 < n0 < 1500
    for i in range(3):
        s = a[ia[i]:ia[i + 1]]
        assert np.all(np.diff(s.astype(np.uint64)) > 0)
    # shared contexts drop as divergence grows
    shared = [len(np.intersect1d(a[ia[0]:ia[1]] >> np.uint64(20), a[ia[i]:ia[i + 1]] >> np.uint64(20)))
              for i in (1, 2)]
    assert n0 > shared[0] > shared[1] > 0


def test_sim_fastq_pair_reads_come_from_the_genome(tmp_path):
    """The paired-read simulator behind the read-set workloads: with no
    errors every read is a substring of the genome or of its reverse
    complement, mates face each other within the insert size; with 1%
    errors about 1% of the bases differ from the best placement."""
    host = kt.Host()
    g = host.sim_genome(kt.SimSpec(77, 20000, 3, 0.0, 0.0, 0, 0.0))
    letters = np.frombuffer(b"ACGT", np.uint8)
    fwd = letters[g].tobytes().decode()
    rev = letters[3 - g[::-1]].tobytes().decode()

    def reads(err):
        a, b = tmp_path / f"r{err}_1.fq", tmp_path / f"r{err}_2.fq"
        n = host.sim_fastq_pair(g, 10.0, 5, a, b, err=err)
        ra = a.read_text().splitlines()[1::4]
        rb = b.read_text().splitlines()[1::4]
        assert len(ra) == len(rb) == n == int(10.0 * len(g) / 300)
        assert all(len(x) == 150 for x in ra + rb)
        return ra, rb

    ra, rb = reads(0.0)
    for x, y in zip(ra[:200], rb[:200]):
        if x in fwd:
            i, j = fwd.find(x), rev.find(y)
            assert j >= 0 and 150 <= (len(g) - j) - i <= 600  # mate on the other strand, downstream
        else:
            assert x in rev and y in fwd
    ra, rb = reads(0.01)
    diff = total = 0
    for x in ra[:300]:
        place = [sum(p != q for p, q in zip(x, s[i:i + 150]))
                 for s in (fwd, rev) for o in range(0, 121, 30) for i in [s.find(x[o:o + 30]) - o]
                 if s.find(x[o:o + 30]) >= 0 and 0 <= i <= len(s) - 150]
        if place:  # an error-free 30-mer anchors the read
            diff += min(place)
            total += 150
    assert total >= 150 * 290 and 0.005 < diff / total < 0.02


def test_pooled_sketch_equals_cli_asone(tmp_path):
    """Host.sketch_pooled (kssd_sketch_pooled, in process) gives the keys and
    counts `kssd3a-hip sketch --asone --conflict -A` writes to comblco and
    comblco.a for the same read files."""
    import subprocess
    host = kt.Host()
    g = host.sim_genome(kt.SimSpec(91, 60000, 4, 0.0, 0.0, 0, 0.0))
    a, b = tmp_path / "s_1.fq", tmp_path / "s_2.fq"
    host.sim_fastq_pair(g, 15.0, 9, a, b)
    keys, counts = host.sketch_pooled([a, b], threads=2)
    p = subprocess.run([str(kt.PKG / "kssd3a-hip"), "sketch", "--asone", "--conflict", "-A", "-p", "2", "-o",
                        str(tmp_path / "R"), str(a), str(b)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    assert np.array_equal(keys, np.fromfile(tmp_path / "R" / "comblco", "<u8"))
    assert np.array_equal(counts, np.fromfile(tmp_path / "R" / "comblco.a", "<u4"))
    assert len(keys) > 1000 and (np.diff(keys >> np.uint64(20)) == 0).any()  # conflicts kept
