"""Pins the CPU oracle (oracle/crispy_oracle.c) against the reference's own
outputs (tests/golden, written by the reference binaries built from source)
and against closed forms that a plausible mistake would violate."""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O
from crispy_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CLEAN = os.path.join(GOLD, "V2Mice_sample1000_Clean")
ADV = os.path.join(GOLD, "adversarial.fa")

ENC = {ord("A"): 0, ord("G"): 1, ord("T"): 2, ord("C"): 3}


def canon(p, d):
    p = np.asarray(p).astype(np.int64)
    d = np.asarray(d, np.float32).view(np.uint32).astype(np.int64)
    key = np.lexsort((p[:, 1], p[:, 0], d))
    return np.stack([d[key], p[key, 0], p[key, 1]], 1)


def all_pairs(n):
    i, j = np.triu_indices(n, 1)
    return np.stack([i, j], 1).astype(np.int32)


def closed_form_codes(seq: bytes, K: int):
    e = np.array([ENC.get(c, 0) for c in seq], np.int64)
    T = len(seq) - K + 1
    codes = np.zeros(T, np.int64)
    for q in range(K):
        codes = (codes << 2) | e[q:q + T]
    return codes


def rand_seq(rng, alphabet, n):
    return np.array(list(alphabet), np.uint8)[rng.integers(0, len(alphabet), int(n))].tobytes()


def sub_fasta(path, n):
    lines = open(path, "rb").read().split(b"\n")
    seqs = [lines[k] for k in range(1, 2 * n, 2)]
    return O.Reads.from_list(seqs)


# --------------------------------------------------------------------- reader
def test_reader_clean_sample():
    r = O.read_fasta(CLEAN, "kmer")
    g = O.read_fasta(CLEAN, "gen")
    lines = open(CLEAN, "rb").read().split(b"\n")
    seqs = [lines[k] for k in range(1, len(lines) - 1, 2)]
    assert r.n == g.n == len(seqs) == 585
    assert [r.seq(i) for i in range(r.n)] == seqs
    assert r.maxlen == max(len(s) for s in seqs)


def test_reader_edge_cases():
    r = O.read_fasta(ADV, "kmer")
    g = O.read_fasta(ADV, "gen")
    assert r.n == g.n == 14
    # trailing N's are cut only by the kmerDist reader (ATGCatgc vs nNATGCatgc)
    assert r.seq(2).endswith(b"GGCCTTAA") and g.seq(2).endswith(b"GGCCTTAANNN")
    # an all-whitespace line is left as "\n" by the trim and lands inside the read
    assert b"\n" in r.seq(3)
    # the last record has no final newline and still becomes a read
    assert r.seq(13) == b"TTGACCAGGTACCATTGACCAGGTACCA"
    # ... but does not update maxLen (kmerMain.cpp:382-393)
    assert r.maxlen == max(len(r.seq(i)) for i in range(13))


def test_reader_rejects_non_fasta(tmp_path):
    p = tmp_path / "x.fa"
    p.write_bytes(b"ACGT\n")
    with pytest.raises(IOError):
        O.read_fasta(str(p))


# ---------------------------------------------------------------- thresholds
def test_default_thresholds():
    # kmerMain.cpp:43-48 / genMain.cpp:42-47 evaluated by hand
    assert O.kmer_threshold(585) == np.float32(2 / np.log2(585.0))
    assert O.nw_threshold(585) == np.float32(1 / np.log2(585.0))
    assert O.kmer_threshold(400) == np.float32(2 / np.log10(400.0))
    assert O.nw_threshold(400) == np.float32(1 / np.log10(400.0))
    assert O.kmer_threshold(20) == np.float32(1.0)  # 2/log10(20) > 1 is clamped
    assert abs(O.kmer_threshold(20000) - 0.1400) < 1e-4
    assert abs(O.nw_threshold(20000) - 0.0700) < 1e-4


def test_pass_predicate_edges():
    thr = np.float32(0.2)
    assert O.passes(0.1, thr)
    assert O.passes(float(thr), thr)
    assert O.passes(float(np.float32(0.2 + 0.9e-5)), thr)
    assert not O.passes(float(np.float32(0.2 + 1.1e-5)), thr)
    assert not O.passes(0.3, thr)


# ------------------------------------------------------------------- k-mers
def test_kmer_known_codes():
    assert list(O.kmer_tuples(b"AAAAAA", 6)) == [0]
    assert list(O.kmer_tuples(b"CCCCCC", 6)) == [4095]
    assert list(O.kmer_tuples(b"GATTAC", 6)) == [0x4A3]
    assert list(O.kmer_tuples(b"NNNNNN", 6)) == [0]
    assert list(O.kmer_tuples(b"gattac", 6)) == [0]  # lowercase encodes as A


@pytest.mark.parametrize("K", range(1, 9))
def test_kmer_tuples_closed_form(K):
    rng = np.random.default_rng(K)
    for _ in range(20):
        s = rand_seq(rng, b"ACGTN", rng.integers(K, 80))
        got = O.kmer_tuples(s, K)
        cf = closed_form_codes(s, K)
        # signed-short order (compareTwoTuples): codes >= 0x8000 sort first
        exp = np.sort(cf.astype(np.uint16).view(np.int16)).view(np.uint16)
        assert np.array_equal(got, exp)


@pytest.mark.parametrize("K", [3, 5, 6, 7, 8])
def test_kmer_matches_closed_form(K):
    """K<=7: matches = sum_v min(c_a[v], c_b[v]).  K=8: the merge over
    signed-sorted lists gives I_hi + [hinfo_a == hinfo_b] * I_lo, where hi
    codes are >= 0x8000 and hinfo is (max hi code, its multiplicity)."""
    rng = np.random.default_rng(100 + K)
    for _ in range(300):
        la, lb = rng.integers(K, 60, 2)
        alpha = b"ACGT" if rng.random() < 0.7 else b"TC"
        a = rand_seq(rng, alpha, la)
        b = rand_seq(rng, alpha, lb)
        ta, tb = O.kmer_tuples(a, K), O.kmer_tuples(b, K)
        d, m = O.kmer_pairs(O.Reads.from_list([a, b]), K, np.array([[0, 1]]))
        ca = np.bincount(closed_form_codes(a, K), minlength=4 ** K)
        cb = np.bincount(closed_form_codes(b, K), minlength=4 ** K)
        mins = np.minimum(ca, cb)
        if K < 8:
            exp = int(mins.sum())
        else:
            i_hi, i_lo = int(mins[0x8000:].sum()), int(mins[:0x8000].sum())

            def hinfo(c):
                nz = np.nonzero(c[0x8000:])[0]
                return (0, 0) if len(nz) == 0 else (nz[-1], c[0x8000 + nz[-1]])
            exp = i_hi + (i_lo if hinfo(ca) == hinfo(cb) else 0)
        assert m[0] == exp
        # dist = 1.0f - (float)matches / min(T_a, T_b) in fp32 (kmerCPU.cpp:43)
        q = np.float32(exp) / np.float32(min(len(ta), len(tb)))
        assert d[0] == np.float32(np.float32(1.0) - q)


@pytest.mark.parametrize("K", range(4, 9))
def test_kmer_clean_sample_golden(K):
    g = np.load(os.path.join(GOLD, "kmer_v2mice.npz"))
    r = O.read_fasta(CLEAN, "kmer")
    p, d = O.kmer_all(r, K, O.kmer_threshold(r.n), nthreads=1)
    assert np.array_equal(p, g[f"k{K}_pairs"])
    assert np.array_equal(d.view(np.uint32), g[f"k{K}_dists"].view(np.uint32))


@pytest.mark.parametrize("K", range(1, 9))
def test_kmer_all_pairs_golden(K):
    g = np.load(os.path.join(GOLD, "kmer_sub120.npz"))
    r = sub_fasta(CLEAN, 120)
    d, _ = O.kmer_pairs(r, K, all_pairs(120))
    assert np.array_equal(d.view(np.uint32), g[f"k{K}_dists"].view(np.uint32))


@pytest.mark.parametrize("K", [3, 6, 8])
def test_kmer_adversarial_golden(K):
    g = np.load(os.path.join(GOLD, "adversarial.npz"))
    r = O.read_fasta(ADV, "kmer")
    p, d = O.kmer_all(r, K, 1.0, nthreads=1)
    assert np.array_equal(p, g[f"k{K}_pairs"])
    assert np.array_equal(d.view(np.uint32), g[f"k{K}_dists"].view(np.uint32))


# ----------------------------------------------------------------------- NW
@pytest.mark.parametrize("band", [1, 2, 5, 10])
def test_nw_adversarial_golden(band):
    g = np.load(os.path.join(GOLD, "adversarial.npz"))
    r = O.read_fasta(ADV, "gen")
    pairs = all_pairs(r.n)
    d, _, _ = O.nw_pairs(r, pairs, band)
    assert np.array_equal(canon(pairs, d), canon(g[f"b{band}_pairs"], g[f"b{band}_dists"]))


@pytest.mark.parametrize("band", [1, 2, 3, 5, 10, 20])
def test_nw_all_pairs_golden(band):
    g = np.load(os.path.join(GOLD, "gen_sub60.npz"))
    r = sub_fasta(CLEAN, 60)
    pairs = all_pairs(60)
    d, _, _ = O.nw_pairs(r, pairs, band)
    assert np.array_equal(canon(pairs, d), canon(g[f"b{band}_pairs"], g[f"b{band}_dists"]))


@pytest.mark.parametrize("band", [10, 5, 2, 1])
def test_gendist_clean_sample_golden(band, tmp_path):
    gk = np.load(os.path.join(GOLD, "kmer_v2mice.npz"))
    g = np.load(os.path.join(GOLD, "gen_v2mice.npz"))
    r = O.read_fasta(CLEAN, "gen")
    prefix = str(tmp_path / "out")
    nf, tot = O.gendist_files(r, gk["k6_pairs"], band, O.nw_threshold(r.n), prefix)
    assert nf == 1
    p = np.fromfile(prefix + ".npair_0", np.uint32).reshape(-1, 2)
    d = np.fromfile(prefix + ".ndist_0", np.float32)
    assert np.all(np.diff(d) >= 0) and np.all(np.diff(g[f"b{band}_dists"]) >= 0)
    assert np.array_equal(canon(p, d), canon(g[f"b{band}_pairs"], g[f"b{band}_dists"]))


def test_nw_closed_forms():
    L = 40
    rng = np.random.default_rng(7)
    s = rand_seq(rng, b"ACGT", L)
    for band in (1, 2, 10):
        # identical reads: all-diagonal path, ML = 0
        d, ml, al, sc = O.nw_pair(s, s, band)
        assert (d, ml, al, sc) == (0.0, 0, L, 5 * L)
        # one interior substitution: ML = 1, AL = L
        t = bytearray(s)
        t[20] = ord("A") if t[20] != ord("A") else ord("C")
        d, ml, al, sc = O.nw_pair(s, bytes(t), band)
        assert (ml, al, sc) == (1, L, 5 * (L - 1) - 4) and d == np.float32(1) / np.float32(L)
    # a run of mismatches: at every cell of the run the diagonal (-4) beats
    # opening a gap (-10), so the path stays diagonal: ML = 3, AL = 13
    a, b = b"AAAAAGGGAAAAA", b"AAAAATTTAAAAA"
    for band in (1, 10):
        d, ml, al, sc = O.nw_pair(a, b, band)
        assert (ml, al, sc) == (3, 13, 50 - 12) and d == np.float32(3) / np.float32(13)
    # one inserted base: one gap opened after a diagonal (GO = -10), ML = 1
    d, ml, al, sc = O.nw_pair(b"ACGTACGTAACGTACGT", b"ACGTACGTACGTACGT", 1)
    assert (ml, al, sc) == (1, 17, 5 * 16 - 10)
    # kB = 0: only diagonal cells; the extra base of the longer read is free
    assert O.nw_pair(b"ACGT", b"ACGA", 10)[:3] == (0.25, 1, 4)
    assert O.nw_pair(b"ACG", b"ACGT", 10)[:3] == (0.0, 0, 3)
    # leading overhang counts in AL, not ML; trailing overhang is free
    assert O.nw_pair(b"TTTTACGTACGT", b"ACGTACGT", 2)[:3] == (0.0, 0, 12)
    assert O.nw_pair(b"ACGTACGT", b"ACGTACGTAAAA", 2)[:3] == (0.0, 0, 8)


def test_nw_brute_force_full_matrix():
    """A full-matrix DP that stores every cell and visits them in anti-
    diagonal order (a different loop structure from the two-row oracle)
    must give the same (ML, AL) on random tiny pairs."""
    GO, GE, NEG = -10, -1, -1000000
    rng = np.random.default_rng(11)

    def brute(s1, s2, band):
        L1, L2 = len(s1), len(s2)
        full = band <= 1
        kb = (L1 + L2) // (2 * band) if not full else L1 + L2
        M = {}
        for i in range(0, L2 + 1):
            for j in range(0, L1 + 1):
                if (i == 0 or j == 0) and (i + j <= kb or full):
                    M[(i, j)] = (0, 0, 0, i + j)  # M, U, ML, AL
        bestc, bestr = (NEG, 0, 0), (NEG, 0, 0)
        for s in range(2, L1 + L2 + 1):
            for i in range(max(1, s - L1), min(L2, s - 1) + 1):
                j = s - i
                if not full and abs(i - j) > kb:
                    continue
                m = s2[i - 1] == s1[j - 1]
                dM, _, dML, dAL = M[(i - 1, j - 1)]
                diag = dM + (5 if m else -4)
                up = NEG
                if (i, j - 1) in M:
                    uM, uU, _, _ = M[(i, j - 1)]
                    up = uM + (uU & 1) * GO + ((uU >> 2) & 1) * GE
                left = NEG
                if (i - 1, j) in M:
                    lM, lU, _, _ = M[(i - 1, j)]
                    left = lM + (lU & 1) * GO + ((lU >> 1) & 1) * GE
                best = max(diag, up, left)
                if best == diag:
                    cell = (best, 1, dML - m + 1, dAL + 1)
                elif best == up:
                    cell = (best, 4, M[(i, j - 1)][2] + 1, M[(i, j - 1)][3] + 1)
                else:
                    cell = (best, 2, M[(i - 1, j)][2] + 1, M[(i - 1, j)][3] + 1)
                M[(i, j)] = cell
        for j in range(1, L1 + 1):  # last row, ascending j, >= keeps the last
            if (L2, j) in M and M[(L2, j)][0] >= bestc[0]:
                bestc = (M[(L2, j)][0], M[(L2, j)][2], M[(L2, j)][3])
        for i in range(1, L2 + 1):  # last column, ascending i
            if (i, L1) in M and M[(i, L1)][0] >= bestr[0]:
                bestr = (M[(i, L1)][0], M[(i, L1)][2], M[(i, L1)][3])
        c = bestc if bestc[0] > bestr[0] else bestr
        return c[1], c[2]

    for t in range(400):
        l1, l2 = (int(x) for x in rng.integers(1, 14, 2))
        alpha = b"AC" if t % 2 else b"ACGT"
        s1 = rand_seq(rng, alpha, l1)
        s2 = rand_seq(rng, alpha, l2)
        band = int(rng.choice([1, 2, 3, 10]))
        _, ml, al, _ = O.nw_pair(s1, s2, band)
        assert (ml, al) == brute(s1, s2, band), (s1, s2, band)


# ----------------------------------------------------------------- synthetic
def test_synth_reads1k_golden(tmp_path):
    g = np.load(os.path.join(GOLD, "synth_reads1k.npz"))
    fa = str(tmp_path / "reads1k.fa")
    synth.make("reads1k", fa)
    assert hashlib.sha256(open(fa, "rb").read()).hexdigest() == str(g["fasta_sha256"])
    r = O.read_fasta(fa, "kmer")
    p, d = O.kmer_all(r, 6, O.kmer_threshold(r.n), nthreads=1)
    assert np.array_equal(p, g["k6_pairs"])
    assert np.array_equal(d.view(np.uint32), g["k6_dists"].view(np.uint32))
    rg = O.read_fasta(fa, "gen")
    prefix = str(tmp_path / "reads1k.fa")
    nf, _ = O.gendist_files(rg, p, 10, O.nw_threshold(rg.n), prefix)
    assert nf == int(g["b10_nfiles"])
    for f in range(nf):
        fp = np.fromfile(f"{prefix}.npair_{f}", np.uint32).reshape(-1, 2)
        fd = np.fromfile(f"{prefix}.ndist_{f}", np.float32)
        c = canon(fp, fd)
        rec = np.zeros(len(c), dtype=[("d", "<u4"), ("i", "<u4"), ("j", "<u4")])
        rec["d"], rec["i"], rec["j"] = c[:, 0], c[:, 1], c[:, 2]
        assert hashlib.sha256(rec.tobytes()).hexdigest() == str(g[f"b10_f{f}_sha256"])


@pytest.mark.slow
def test_gendist_file_split_golden(tmp_path):
    g = np.load(os.path.join(GOLD, "synth_multifile.npz"))
    fa = str(tmp_path / "reads1200.fa")
    seqs, _, _ = synth.make("reads1k", fa, n=1200)
    assert hashlib.sha256(open(fa, "rb").read()).hexdigest() == str(g["fasta_sha256"])
    r = O.read_fasta(fa, "gen")
    nf, _ = O.gendist_files(r, all_pairs(r.n), 10, float(g["thr"]), fa)
    assert nf == int(g["nfiles"]) == 2
    for f in range(nf):
        fp = np.fromfile(f"{fa}.npair_{f}", np.uint32).reshape(-1, 2)
        fd = np.fromfile(f"{fa}.ndist_{f}", np.float32)
        assert len(fd) == int(g[f"f{f}_count"])
        c = canon(fp, fd)
        rec = np.zeros(len(c), dtype=[("d", "<u4"), ("i", "<u4"), ("j", "<u4")])
        rec["d"], rec["i"], rec["j"] = c[:, 0], c[:, 1], c[:, 2]
        assert hashlib.sha256(rec.tobytes()).hexdigest() == str(g[f"f{f}_sha256"])


PRE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pre")
PRE_CASES = [("pre_adv.fa", "pre_adv.fa_Clean", False, None),
             ("pre_dups.fa", "pre_dups.fa_Clean", False, None),
             ("pre_dups.fa", "pre_dups.fa_s1", False, 1.0),
             ("pre_reads.fq", "pre_reads.fq_Clean", True, None)]


@pytest.mark.parametrize("inp,out,fastq,sd", PRE_CASES + [("../V2Mice_sample1000", "../V2Mice_sample1000_Clean", False, None)])
def test_oracle_preprocess_matches_reference(tmp_path, inp, out, fastq, sd):
    """The preprocess restatement against the reference binary's _Clean/.frq."""
    o = str(tmp_path / "out_Clean")
    O.preprocess_file(os.path.join(PRE, inp), o, fastq, sd)
    assert open(o, "rb").read() == open(os.path.join(PRE, out), "rb").read()
    assert open(o + ".frq", "rb").read() == open(os.path.join(PRE, out) + ".frq", "rb").read()


PRIMER_CASES = [("pre_primer.fa_Clean", 2, None), ("pre_primer.fa_m0", 0, None), ("pre_primer.fa_m4s1", 4, 1.0)]


@pytest.mark.parametrize("out,max_mm,sd", PRIMER_CASES)
def test_oracle_primer_filter_matches_reference(tmp_path, out, max_mm, sd):
    """-p (with -m and -s) against the reference binary's outputs."""
    primer = open(os.path.join(PRE, "pre_primer.txt"), "rb").readline().rstrip(b"\n\r \t\v\f\b")
    o = str(tmp_path / "p_Clean")
    O.preprocess_file(os.path.join(PRE, "pre_primer.fa"), o, False, sd, primer, max_mm)
    assert open(o, "rb").read() == open(os.path.join(PRE, out), "rb").read()
    assert open(o + ".frq", "rb").read() == open(os.path.join(PRE, out) + ".frq", "rb").read()


def test_oracle_primer_closed_forms():
    """Cases the alignment fixes by hand: an exact primer anywhere in the
    window aligns with 0 mismatches; k substitutions in the middle cost k; a
    primer cut short by c bases costs c (unaligned end); the window is
    2 |primer| bytes, and a window without a single matching byte leaves the
    all-zero first cell as the best, a one-base alignment whose unaligned
    primer end costs |primer| - 1; empty primer or read always fails."""
    P = b"GTGCCAGCAGCCGCGGTAA"
    tail = b"T" * 30
    assert O.primer_keep(P, P + tail, 0)
    assert O.primer_keep(P, b"ACG" + P + tail, 0)
    sub2 = P[:6] + b"T" + P[7:12] + b"T" + P[13:]
    assert O.primer_keep(P, sub2 + tail, 2) and not O.primer_keep(P, sub2 + tail, 1)
    cut3 = P[3:] + tail
    assert O.primer_keep(P, cut3, 3) and not O.primer_keep(P, cut3, 2)
    far = b"N" * (2 * len(P)) + P  # the primer lies past the window
    assert O.primer_keep(P, far, len(P) - 1) and not O.primer_keep(P, far, len(P) - 2)
    assert not O.primer_keep(b"", P, 999) and not O.primer_keep(P, b"", 999)
    assert O.primer_keep(b"A", b"A", 0) and O.primer_keep(b"A", b"C", 0)


def test_oracle_preprocess_merge_is_stable_last_wins(tmp_path):
    """Closed-form check: duplicates collapse onto their last occurrence with
    the run length as frequency, output in (length, bytes) order."""
    fa = tmp_path / "d.fa"
    recs = [("a", b"CCCC"), ("b", b"AAAA"), ("c", b"CCCC"), ("d", b"AC"), ("e", b"CCCC"), ("f", b"AAAA")]
    fa.write_bytes(b"".join(b">" + l.encode() + b"\n" + s + b"\n" for l, s in recs))
    o = str(tmp_path / "d_Clean")
    assert O.preprocess_file(str(fa), o) == 3
    assert open(o, "rb").read() == b"> d\nAC\n> f\nAAAA\n> e\nCCCC\n"
    assert open(o + ".frq", "rb").read() == b"d 1\nf 2\ne 3\n"


def test_nw_as_shipped_settings_same_distances():
    """The timing-only form under genDist's shipped thread settings (threads
    confined to a few CPUs, static,1 schedule, a critical section per pair)
    gives the same distances as the plain oracle."""
    seqs, _, _ = synth.make("reads1k", n=300)
    r = O.Reads.from_list(seqs)
    p = np.array([[i, j] for i in range(0, 40) for j in range(i + 1, 80)], np.int32)
    a, _, _ = O.nw_pairs(r, p, 10)
    b = O.nw_pairs_as_shipped(r, p, 10, 4, 2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("K", [4, 6, 8])
def test_kmer_rows_is_kmer_all_on_those_rows(K):
    """The row-list form (used to sample full-size runs) gives exactly the
    pairs and distances of the all-pairs form whose first read is listed;
    repeated and unsorted rows are taken once."""
    seqs, _, _ = synth.make("reads1k", n=400)
    r = O.Reads.from_list(seqs)
    thr = 0.3
    ap, ad = O.kmer_all(r, K, thr, nthreads=1)
    rows = np.array([399, 0, 7, 7, 200, 398, 123], np.int32)
    p, d = O.kmer_rows(r, K, thr, rows, nthreads=3)
    order = np.lexsort((p[:, 1], p[:, 0]))
    sel = np.isin(ap[:, 0], rows)
    assert sel.sum() > 0
    assert np.array_equal(p[order], ap[sel]) and np.array_equal(d[order].view(np.uint32), ad[sel].view(np.uint32))


def _bucket(codes, buckets):
    """The k = 6 candidate filter's bucket of a code: h = code * 2654435761
    mod 2^32 scaled to the bucket range, (h * buckets) >> 32
    (kmer_bucket in crispy_amd/csrc/kmer.hip)."""
    h = (codes.astype(np.uint64) * 2654435761) & 0xFFFFFFFF
    return (h * np.uint64(buckets)) >> 32


def test_bucket_of_2048_is_the_top_11_bits():
    """(h * 2048) >> 32 is h >> 21 (2,048 buckets: the hash's top 11 bits);
    with 1,280 or 1,536 buckets every code lands in [0, buckets) and the
    bucket is monotone in h."""
    c = np.arange(4096, dtype=np.int64)
    h = (c.astype(np.uint64) * 2654435761) & 0xFFFFFFFF
    assert np.array_equal(_bucket(c, 2048), h >> 21)
    hs = np.sort(h)
    for buckets in (1280, 1536):
        b = _bucket(c, buckets)
        assert b.min() >= 0 and b.max() <= buckets - 1
        # the scaled range is monotone in h: the buckets split the hash range
        # into runs of 2^32 / buckets values
        assert np.all(np.diff(_bucket_of_hash(hs, buckets).astype(np.int64)) >= 0)


def _bucket_of_hash(h, buckets):
    return (h * np.uint64(buckets)) >> 32


@pytest.mark.parametrize("buckets", [1280, 1536, 2048])
@pytest.mark.parametrize("K", [6, 7, 8])
def test_hashed_filter_bound_is_sound(K, buckets):
    """The hashed k = 6 candidate filter may keep extra pairs but must never
    drop one: for every pair, popcount(bucket presence AND) + min(T - distinct
    buckets) >= the oracle's exact K-mer matches (kmerCPU.cpp:29-43), since
    merging codes into buckets only raises sum_b min(count_i, count_j) and
    K-mer matches never exceed the 6-mer multiset intersection.  Brute force
    over random, low-complexity, T/C-rich and near-copy reads."""
    rng = np.random.default_rng(600 + K)
    seqs = []
    alph = np.frombuffer(b"ACGT", np.uint8)
    base = alph[rng.integers(0, 4, 900)]
    for k in range(90):
        L = int(rng.integers(K + 2, 900))
        kind = k % 4
        if kind == 0:
            s = np.resize(alph[rng.integers(0, 4, int(rng.integers(1, 6)))], L)
        elif kind == 1:
            s = np.frombuffer(b"TC", np.uint8)[rng.integers(0, 2, L)]
        elif kind == 2:
            s = base[:L].copy()
            for _ in range(L // 50):
                s[int(rng.integers(0, L))] = alph[int(rng.integers(0, 4))]
        else:
            s = alph[rng.integers(0, 4, L)]
        seqs.append(s.tobytes())
    reads = O.Reads.from_list(seqs)
    n = reads.n
    presence, hexcess = [], []
    for s in seqs:
        c = O.kmer_tuples(s, 6).astype(np.int64)
        b = np.unique(_bucket(c, buckets))
        presence.append(set(b.tolist()))
        hexcess.append(len(c) - len(b))
    pairs = np.array([[i, j] for i in range(n) for j in range(i + 1, n)], np.int32)
    _, matches = O.kmer_pairs(reads, K, pairs)
    tight = 0
    for (i, j), m in zip(pairs, matches):
        bound = len(presence[i] & presence[j]) + min(hexcess[i], hexcess[j])
        assert bound >= m, (i, j, bound, m)
        tight += bound == m
    assert tight > 0  # the bound is attained (identical or near-identical reads)
