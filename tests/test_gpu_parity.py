"""GPU parity: the HIP path (through the C ABI) against the reference's
golden outputs and the CPU oracle, bit for bit (integer counts and the
float32 distances from the same formulas)."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import crispy_amd as ca
from crispy_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CLEAN = os.path.join(GOLD, "V2Mice_sample1000_Clean")
ADV = os.path.join(GOLD, "adversarial.fa")
BIN = os.path.join(ROOT, "bin")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ca.Context(0)
    yield c
    c.close()


def canon(p, d):
    p = np.asarray(p).astype(np.int64).reshape(-1, 2)
    d = np.asarray(d, np.float32).view(np.uint32).astype(np.int64)
    key = np.lexsort((p[:, 1], p[:, 0], d))
    return np.stack([d[key], p[key, 0], p[key, 1]], 1)


def digest(p, d):
    c = canon(p, d)
    rec = np.zeros(len(c), dtype=[("d", "<u4"), ("i", "<u4"), ("j", "<u4")])
    rec["d"], rec["i"], rec["j"] = c[:, 0], c[:, 1], c[:, 2]
    return hashlib.sha256(rec.tobytes()).hexdigest()


def all_pairs(n):
    i, j = np.triu_indices(n, 1)
    return np.stack([i, j], 1).astype(np.int32)


def sub_reads(path, n, variant):
    r = ca.read_fasta(path, variant)
    return ca.Reads(r.bases[:r.offsets[n]], r.offsets[:n + 1])


def bits(d):
    return np.asarray(d, np.float32).view(np.uint32)


# ---------------------------------------------------------------- arithmetic
def test_division_matches_host(ctx):
    say("division table 2048")
    k, w = ctx.selftest_division(2048)
    m = np.arange(2049, dtype=np.float32)[None, :]
    a = np.arange(2049, dtype=np.float32)[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        ek = np.float32(1.0) - m / a
        ew = m / a
    ek[0, :] = 0
    ew[0, :] = 0
    assert np.array_equal(bits(k), bits(ek.astype(np.float32)))
    assert np.array_equal(bits(w), bits(ew.astype(np.float32)))


# ------------------------------------------------------------------ kmerDist
@pytest.mark.parametrize("K", range(4, 9))
def test_kmer_clean_sample_golden(ctx, K):
    say("kmer golden K", K)
    g = np.load(os.path.join(GOLD, "kmer_v2mice.npz"))
    r = ca.read_fasta(CLEAN, ca.KMER)
    ctx.upload_reads(ca.KMER, r)
    c = ctx.kmer_run(K, ca.kmer_default_threshold(r.n))
    p, d = ctx.kmer_fetch()
    assert c == len(g[f"k{K}_pairs"])
    assert np.array_equal(p, g[f"k{K}_pairs"])
    assert np.array_equal(bits(d), bits(g[f"k{K}_dists"]))


@pytest.mark.parametrize("K", range(1, 9))
def test_kmer_all_pairs_golden(ctx, K):
    say("kmer all pairs K", K)
    g = np.load(os.path.join(GOLD, "kmer_sub120.npz"))
    ctx.upload_reads(ca.KMER, sub_reads(CLEAN, 120, ca.KMER))
    c = ctx.kmer_run(K, 1.0)
    p, d = ctx.kmer_fetch()
    assert c == 120 * 119 // 2
    assert np.array_equal(p, all_pairs(120))
    assert np.array_equal(bits(d), bits(g[f"k{K}_dists"]))


@pytest.mark.parametrize("K", [3, 6, 8])
def test_kmer_adversarial_golden(ctx, K):
    say("kmer adversarial K", K)
    g = np.load(os.path.join(GOLD, "adversarial.npz"))
    ctx.upload_reads(ca.KMER, ca.read_fasta(ADV, ca.KMER))
    ctx.kmer_run(K, 1.0)
    p, d = ctx.kmer_fetch()
    assert np.array_equal(p, g[f"k{K}_pairs"])
    assert np.array_equal(bits(d), bits(g[f"k{K}_dists"]))


def random_reads(rng, n, lo, hi, alphabet=b"ACGT", repeats=True):
    seqs = []
    alph = np.frombuffer(alphabet, np.uint8)
    for k in range(n):
        L = int(rng.integers(lo, hi + 1))
        kind = k % 5
        if kind == 0 and repeats:      # low complexity: tandem repeat
            unit = alph[rng.integers(0, len(alph), int(rng.integers(1, 5)))]
            s = np.resize(unit, L)
        elif kind == 1:                # T/C rich (k = 8 hi codes)
            s = np.frombuffer(b"TC", np.uint8)[rng.integers(0, 2, L)]
        else:
            s = alph[rng.integers(0, len(alph), L)]
        seqs.append(s.tobytes())
    # a family of near copies so that some pairs pass
    base = seqs[2]
    for k in range(3, n, 7):
        t = bytearray(base[: len(seqs[k])] if len(seqs[k]) <= len(base) else base)
        for _ in range(int(rng.integers(0, 6))):
            t[int(rng.integers(0, len(t)))] = alph[int(rng.integers(0, len(alph)))]
        seqs[k] = bytes(t)
    return seqs


@pytest.mark.parametrize("K", range(1, 9))
def test_kmer_random_vs_oracle(ctx, K):
    say("kmer random K", K)
    rng = np.random.default_rng(1000 + K)
    seqs = random_reads(rng, 333, max(K, 8), 300)
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    for thr in (0.3, 1.0):
        ctx.upload_reads(ca.KMER, r)
        ctx.kmer_run(K, thr)
        p, d = ctx.kmer_fetch()
        op, od = O.kmer_all(ro, K, thr, nthreads=1)
        assert np.array_equal(p, op)
        assert np.array_equal(bits(d), bits(od))


@pytest.mark.parametrize("K", [7, 8])
def test_kmer_long_reads_vs_oracle(ctx, K):
    """K-mer lists of every sort form: T = L-K+1 codes up to 512 and up to
    1,024 sort in registers (8 and 16 per lane), longer reads in LDS; lengths
    straddle both limits and reach the 2,048 bp maximum, with near copies so
    that pairs pass."""
    say("kmer long reads K", K)
    rng = np.random.default_rng(4000 + K)
    lens = [T + K - 1 for T in (511, 512, 513, 1023, 1024, 1025)] + [2048, 2047, 1500, 700]
    lens += [int(x) for x in rng.integers(400, 2049, 30)]
    alph = np.frombuffer(b"ACGT", np.uint8)
    base = alph[rng.integers(0, 4, 2048)]
    seqs = []
    for k, L in enumerate(lens):
        if k % 3 == 0:
            s = alph[rng.integers(0, 4, L)]
        elif k % 3 == 1:
            s = base[:L].copy()
            for _ in range(L // 40):
                s[int(rng.integers(0, L))] = alph[int(rng.integers(0, 4))]
        else:
            s = np.frombuffer(b"TC", np.uint8)[rng.integers(0, 2, L)]
        seqs.append(s.tobytes())
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    ctx.upload_reads(ca.KMER, r)
    for thr in (0.3, 1.0):
        ctx.kmer_run(K, thr)
        p, d = ctx.kmer_fetch()
        op, od = O.kmer_all(ro, K, thr, nthreads=1)
        assert np.array_equal(p, op)
        assert np.array_equal(bits(d), bits(od))


@pytest.mark.parametrize("K", range(1, 9))
def test_kmer_exact_pass_forms(ctx, K):
    """Every exact-count form that applies to K gives the oracle's pairs and
    distances: in the tile kernel (K <= 6), in a pass of its own over the
    tile's candidates (hashed for K >= 6, into 1,280, 1,536 and 2,048 buckets)
    and over exact-bitmap candidates (K >= 6), at a threshold where most
    candidates pass and at one where most fail."""
    say("kmer exact pass forms K", K)
    rng = np.random.default_rng(3000 + K)
    seqs = random_reads(rng, 300, max(K, 8), 300, repeats=True)
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    ctx.upload_reads(ca.KMER, r)
    forms = [(0, 0)] if K <= 6 else []
    forms += [(1, 1280), (1, 1536), (1, 2048)] if K >= 6 else [(1, 0)]
    forms += [(2, 0)] if K >= 6 else []
    try:
        for thr in (0.25, 0.6):
            op, od = O.kmer_all(ro, K, thr, nthreads=1)
            for mode, buckets in forms:
                ctx.kmer_exact_pass(mode)
                ctx.kmer_hash_buckets(buckets)
                ctx.kmer_run(K, thr)
                if mode == 1 and K >= 6:
                    assert ctx.kmer_tile_bits() == buckets
                p, d = ctx.kmer_fetch()
                assert np.array_equal(p, op), (thr, mode, buckets)
                assert np.array_equal(bits(d), bits(od)), (thr, mode, buckets)
    finally:
        ctx.kmer_exact_pass(-1)
        ctx.kmer_hash_buckets(0)


@pytest.mark.parametrize("K", [6, 8])
def test_kmer_autotune(ctx, K):
    """The k-mer autotuner picks a form that applies to K and leaves the
    oracle's result."""
    say("kmer autotune K", K)
    seqs, _, _ = synth.make("reads1k")
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    thr = ca.kmer_default_threshold(r.n)
    ctx.upload_reads(ca.KMER, r)
    try:
        mode = ctx.kmer_autotune(K, thr, repeats=1)
        assert mode in ((0, 1, 2) if K == 6 else (1, 2))
        p, d = ctx.kmer_fetch()
    finally:
        ctx.kmer_exact_pass(-1)
    op, od = O.kmer_all(ro, K, thr, nthreads=O.num_threads())
    order = np.lexsort((op[:, 1], op[:, 0]))
    assert np.array_equal(p, op[order]) and np.array_equal(bits(d), bits(od[order]))


def test_kmer_tile_ranges_partition(ctx):
    say("kmer tile ranges")
    seqs, _, _ = synth.make("reads1k")
    r = ca.Reads.from_list(seqs)
    ctx.upload_reads(ca.KMER, r)
    ctx.kmer_run(6, 0.25)
    fp, fd = ctx.kmer_fetch()
    nt = ca.api.load_library().crispy_kmer_num_tiles(r.n)
    parts = []
    for t0, t1 in [(0, 5), (5, 17), (17, nt)]:
        ctx.kmer_run(6, 0.25, tiles=(t0, t1))
        parts.append(ctx.kmer_fetch())
    p = np.concatenate([x[0] for x in parts])
    d = np.concatenate([x[1] for x in parts])
    assert np.array_equal(canon(p, d), canon(fp, fd))


def test_trim_releases_every_buffer():
    """After runs that touch every device buffer family (the hashed k = 6
    candidates and their segments, K = 8 lists, the fused NW, the device
    output sort, the CSR), trimming a context gives all its device memory
    back, and the context still runs."""
    say("trim")
    seqs, _, _ = synth.make("reads1k")
    r = ca.Reads.from_list(seqs)
    thr = ca.kmer_default_threshold(r.n)
    with ca.Context(0) as c:
        c.upload_reads(ca.KMER, r)
        c.upload_reads(ca.NW, r)
        c.kmer_exact_pass(1)
        c.kmer_run(8, thr)
        c.kmer_run(6, thr)
        c.nw_run_on_kmer(10, ca.nw_default_threshold(r.n))
        files = ca.pipeline.fused_output_files(c)
        c.nw_csr(r.n)
        assert c.device_bytes() > 0
        c.trim()
        assert c.device_bytes() == 0
        c.upload_reads(ca.KMER, r)
        c.upload_reads(ca.NW, r)
        c.kmer_run(6, thr)
        c.nw_run_on_kmer(10, ca.nw_default_threshold(r.n))
        again = ca.pipeline.fused_output_files(c)
    assert len(again) == len(files)
    for (p0, d0), (p1, d1) in zip(files, again):
        assert np.array_equal(p0, p1) and np.array_equal(bits(d0), bits(d1))


# ------------------------------------------------------------------- genDist
@pytest.mark.parametrize("band", [10, 5, 2, 1])
def test_nw_clean_sample_golden(ctx, band):
    say("nw golden band", band)
    gk = np.load(os.path.join(GOLD, "kmer_v2mice.npz"))
    g = np.load(os.path.join(GOLD, "gen_v2mice.npz"))
    r = ca.read_fasta(CLEAN, ca.NW)
    ctx.upload_reads(ca.NW, r)
    c = ctx.nw_run(gk["k6_pairs"], band, ca.nw_default_threshold(r.n))
    p, d = ctx.nw_fetch()
    assert c == len(g[f"b{band}_dists"])
    assert np.array_equal(canon(p, d), canon(g[f"b{band}_pairs"], g[f"b{band}_dists"]))


@pytest.mark.parametrize("band", [1, 2, 3, 5, 10, 20])
def test_nw_all_pairs_golden(ctx, band):
    say("nw all pairs band", band)
    g = np.load(os.path.join(GOLD, "gen_sub60.npz"))
    ctx.upload_reads(ca.NW, sub_reads(CLEAN, 60, ca.NW))
    pairs = all_pairs(60)
    d = ctx.nw_dists(pairs, band)
    assert np.array_equal(canon(pairs, d), canon(g[f"b{band}_pairs"], g[f"b{band}_dists"]))


@pytest.mark.parametrize("band", [1, 2, 5, 10])
def test_nw_adversarial_golden(ctx, band):
    say("nw adversarial band", band)
    g = np.load(os.path.join(GOLD, "adversarial.npz"))
    r = ca.read_fasta(ADV, ca.NW)
    ctx.upload_reads(ca.NW, r)
    pairs = all_pairs(r.n)
    d = ctx.nw_dists(pairs, band)
    assert np.array_equal(canon(pairs, d), canon(g[f"b{band}_pairs"], g[f"b{band}_dists"]))


@pytest.mark.parametrize("band", [2, 3, 5, 7, 10, 16, 25])
def test_nw_random_vs_oracle(ctx, band):
    """Lengths 1..400 with every kind of length disparity, both orientations,
    clean and non-ACGT reads: covers every register-kernel kB up to the
    maximum and the generic kernel."""
    say("nw random band", band)
    rng = np.random.default_rng(2000 + band)
    seqs = random_reads(rng, 120, 1, 400, repeats=True)
    for k in range(0, 120, 13):  # some non-ACGT reads
        s = bytearray(seqs[k])
        if s:
            s[int(rng.integers(0, len(s)))] = ord("N") if k % 2 else ord("a")
        seqs[k] = bytes(s)
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    pairs = rng.integers(0, 120, (4000, 2)).astype(np.int32)
    ctx.upload_reads(ca.NW, r)
    d = ctx.nw_dists(pairs, band)
    od, _, _ = O.nw_pairs(ro, pairs, band)
    bad = np.nonzero(bits(d) != bits(od))[0]
    assert len(bad) == 0, [(int(pairs[b, 0]), int(pairs[b, 1]), float(d[b]), float(od[b])) for b in bad[:5]]


def test_nw_register_kernel_every_kb(ctx):
    """Pairs whose kB = (L1+L2)/20 spans 0..32 at band 10 (lengths up to
    329 each), near-identical and unrelated, against the oracle."""
    say("nw every kB")
    rng = np.random.default_rng(77)
    seqs = []
    for L in range(1, 330, 3):
        seqs.append(rng.choice(np.frombuffer(b"ACGT", np.uint8), L).tobytes())
    base = seqs[-1]
    for L in range(5, 329, 11):
        t = bytearray(base[:L])
        for _ in range(L // 30):
            t[int(rng.integers(0, L))] = ord("ACGT"[int(rng.integers(0, 4))])
        seqs.append(bytes(t))
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    n = r.n
    pairs = all_pairs(n)
    pairs = np.concatenate([pairs, pairs[:, ::-1]])
    ctx.upload_reads(ca.NW, r)
    d = ctx.nw_dists(pairs, 10)
    od, _, _ = O.nw_pairs(ro, pairs, 10)
    assert np.array_equal(bits(d), bits(od))


@pytest.mark.parametrize("width", [24, 32, 0])
def test_nw_sliced_kernel_shapes(ctx, width):
    """Band 2 and 3 on reads of 60..2048 bp: kB from 33 to beyond 767 covers
    every lane-sliced instance of both slice widths (kB mod 12 and kB mod 16),
    group sizes 3..64 lanes, and the generic kernel above; near copies and
    unrelated reads, both orders; each width forced, then the cost model."""
    say("nw sliced shapes")
    rng = np.random.default_rng(31)
    alph = np.frombuffer(b"ACGT", np.uint8)
    base = rng.choice(alph, 2048).tobytes()
    seqs = []
    for L in list(range(60, 2049, 53)) + [2048, 1535, 1534, 767, 768]:
        t = bytearray(base[:L])
        for _ in range(L // 25):
            t[int(rng.integers(0, L))] = int(alph[int(rng.integers(0, 4))])
        if int(rng.integers(0, 2)):
            del t[int(rng.integers(0, L)):int(rng.integers(0, L)) + 3]
        seqs.append(bytes(t) if len(t) > 0 else b"ACGT")
    for L in range(100, 2049, 211):
        seqs.append(rng.choice(alph, L).tobytes())
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    pairs = all_pairs(r.n)
    sel = rng.choice(len(pairs), size=min(700, len(pairs)), replace=False)
    pairs = np.concatenate([pairs[sel], pairs[sel[:200], ::-1]])
    lens = r.lengths()
    for band in (2, 3):  # the selection reaches every slice end of both widths
        kb = (lens[pairs[:, 0]] + lens[pairs[:, 1]]) // (2 * band)
        sl = kb[(kb > 32) & (kb <= 767)]
        assert set(sl % 12) == set(range(12)) and set(sl % 16) == set(range(16))
    ctx.upload_reads(ca.NW, r)
    ctx.nw_slice_width(width)
    try:
        for band in (2, 3):
            d = ctx.nw_dists(pairs, band)
            od, _, _ = O.nw_pairs(ro, pairs, band)
            bad = np.nonzero(bits(d) != bits(od))[0]
            assert len(bad) == 0, (band, [(int(pairs[b, 0]), int(pairs[b, 1]), float(d[b]), float(od[b]))
                                          for b in bad[:5]])
    finally:
        ctx.nw_slice_width(0)


def test_nw_autotune(ctx):
    """The autotuner picks one of the settings, the context keeps it, the
    result is the same under every setting, and an earlier NW result stays
    fetchable."""
    seqs, _, _ = synth.make("reads1k")
    rng = np.random.default_rng(4)
    seqs = [s + s[: int(rng.integers(100, 400))] for s in seqs[:600]]  # 300-700 bp: kB 33..117 at band 3
    r = ca.Reads.from_list(seqs)
    ctx.upload_reads(ca.KMER, r)
    ctx.upload_reads(ca.NW, r)
    ctx.kmer_run(6, 0.5)
    ctx.nw_run_on_kmer(3, 0.4)
    before = ctx.nw_fetch()
    try:
        w = ctx.nw_autotune(3, repeats=2)
        say("autotune chose", w)
        assert w in (0, 24, 32)
        after = ctx.nw_fetch()
        assert np.array_equal(after[0], before[0]) and np.array_equal(bits(after[1]), bits(before[1]))
        ctx.nw_run_on_kmer(3, 0.4)
        tuned = ctx.nw_fetch()
        assert np.array_equal(tuned[0], before[0]) and np.array_equal(bits(tuned[1]), bits(before[1]))
    finally:
        ctx.nw_slice_width(0)


def test_nw_graph_replays(ctx):
    """The k-mer passes and the NW DP phase and filter replayed from their
    captured graphs equal the same kernels launched one by one, on every
    replay: two NW inputs alternate (each switch updates the graph), with
    eager work and synchronisations between replays, timing toggled (new
    arguments), register and sliced families, and reads with non-ACGT bytes
    (generic kernel after the graph)."""
    say("nw graph replays")
    seqs, _, _ = synth.make("reads1k")
    rng = np.random.default_rng(8)
    seqs = [s + s[: int(rng.integers(0, 400))] for s in seqs[:500]]  # 200-700 bp
    for k in range(0, 500, 97):
        t = bytearray(seqs[k])
        t[int(rng.integers(0, len(t)))] = ord("N")
        seqs[k] = bytes(t)
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    ctx.upload_reads(ca.KMER, r)
    ctx.upload_reads(ca.NW, r)
    extra = rng.integers(0, r.n, (3000, 2)).astype(np.int32)
    try:
        for band in (10, 3):
            ctx.graphs(False)
            ctx.kmer_run(6, 0.5)
            ref_k = ctx.kmer_fetch()
            ctx.nw_run_on_kmer(band, 0.45)
            ref_a = ctx.nw_fetch()
            ref_b = ctx.nw_dists(extra, band)
            od, _, _ = O.nw_pairs(ro, extra, band)
            assert np.array_equal(bits(ref_b), bits(od))
            ctx.graphs(True)
            for rep in range(4):
                ctx.timing(rep == 2)
                ctx.kmer_run(6, 0.5)
                k = ctx.kmer_fetch()
                assert np.array_equal(k[0], ref_k[0]) and np.array_equal(bits(k[1]), bits(ref_k[1])), (band, rep)
                ctx.nw_run_on_kmer(band, 0.45)
                got = ctx.nw_fetch()
                assert np.array_equal(got[0], ref_a[0]) and np.array_equal(bits(got[1]), bits(ref_a[1])), (band, rep)
                ctx.kmer_fetch(dists=False)  # eager work between replays
                ctx.sync()
                d = ctx.nw_dists(extra, band)
                assert np.array_equal(bits(d), bits(ref_b)), (band, rep)
            say("band", band, "pairs", len(ref_a[1]))
    finally:
        ctx.timing(False)
        ctx.graphs(True)


def test_sort_by_dist(ctx):
    rng = np.random.default_rng(5)
    n = 100000
    p = rng.integers(0, 5000, (n, 2)).astype(np.int32)
    d = (rng.integers(0, 50, n) / 64.0).astype(np.float32)
    sp, sd = ctx.sort_by_dist(p, d)
    assert np.array_equal(canon(sp, sd), canon(p, d))
    c = canon(p, d)
    assert np.array_equal(bits(sd).astype(np.int64), c[:, 0])
    assert np.array_equal(sp[:, 0].astype(np.int64), c[:, 1])
    assert np.array_equal(sp[:, 1].astype(np.int64), c[:, 2])


# --------------------------------------------------------------- end to end
def test_fused_reads1k_golden(ctx):
    say("fused reads1k")
    g = np.load(os.path.join(GOLD, "synth_reads1k.npz"))
    seqs, _, _ = synth.make("reads1k")
    r = ca.Reads.from_list(seqs)
    ctx.upload_reads(ca.KMER, r)
    ctx.upload_reads(ca.NW, r)
    c = ctx.kmer_run(6, ca.kmer_default_threshold(r.n))
    kp, kd = ctx.kmer_fetch()
    assert np.array_equal(kp, g["k6_pairs"])
    assert np.array_equal(bits(kd), bits(g["k6_dists"]))
    ctx.nw_run_on_kmer(10, ca.nw_default_threshold(r.n))
    p, d = ctx.nw_fetch()
    assert int(g["b10_nfiles"]) == 1
    assert len(d) == int(g["b10_f0_count"])
    assert digest(p, d) == str(g["b10_f0_sha256"])


def run(args, cwd, timeout=300):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (args, r.stdout[-2000:], r.stderr[-2000:])
    return r


def test_cli_clean_sample(tmp_path):
    say("cli clean sample")
    gk = np.load(os.path.join(GOLD, "kmer_v2mice.npz"))
    g = np.load(os.path.join(GOLD, "gen_v2mice.npz"))
    fa = tmp_path / "V2Mice_sample1000_Clean"
    fa.write_bytes(open(CLEAN, "rb").read())
    run([os.path.join(BIN, "kmerDist"), "-i", str(fa), "-k", "6", "-d"], tmp_path)
    p = np.fromfile(str(fa) + ".kpair", np.int32).reshape(-1, 2)
    d = np.fromfile(str(fa) + ".kdist", np.float32)
    assert np.array_equal(p, gk["k6_pairs"]) and np.array_equal(bits(d), bits(gk["k6_dists"]))
    run([os.path.join(BIN, "genDist"), "-i", str(fa), "-b", "10"], tmp_path)
    assert not os.path.exists(str(fa) + ".npair_1")
    np_ = np.fromfile(str(fa) + ".npair_0", np.uint32).reshape(-1, 2)
    nd = np.fromfile(str(fa) + ".ndist_0", np.float32)
    assert np.all(np.diff(nd) >= 0)
    assert np.array_equal(canon(np_, nd), canon(g["b10_pairs"], g["b10_dists"]))


def test_cli_file_split(tmp_path):
    """719,400 input pairs: the file split follows the reference CPU path
    (one file flushed after the first 589,824-pair chunk, one for the rest)."""
    say("cli file split")
    g = np.load(os.path.join(GOLD, "synth_multifile.npz"))
    fa = str(tmp_path / "reads1200.fa")
    synth.make("reads1k", fa, n=1200)
    assert hashlib.sha256(open(fa, "rb").read()).hexdigest() == str(g["fasta_sha256"])
    all_pairs(int(g["n"])).tofile(fa + ".kpair")
    run([os.path.join(BIN, "genDist"), "-i", fa, "-b", "10", "-t", str(float(g["thr"]))], tmp_path)
    for f in range(int(g["nfiles"])):
        p = np.fromfile(f"{fa}.npair_{f}", np.uint32).reshape(-1, 2)
        d = np.fromfile(f"{fa}.ndist_{f}", np.float32)
        assert len(d) == int(g[f"f{f}_count"])
        assert np.all(np.diff(d) >= 0)
        assert digest(p, d) == str(g[f"f{f}_sha256"])
    assert not os.path.exists(f"{fa}.npair_{int(g['nfiles'])}")


def test_device_file_split_and_sorted_fetch(ctx):
    """The device output side: the file split of a 719,400-pair run read from
    the filter's scan, and each file sorted by (dist, i, j) on the device,
    equal the reference files (two files) entry for entry; the same on a
    fused reads1k run against the oracle's genDist files."""
    say("device file split")
    g = np.load(os.path.join(GOLD, "synth_multifile.npz"))
    seqs, _, _ = synth.make("reads1k", n=1200)
    r = ca.Reads.from_list(seqs)
    ctx.upload_reads(ca.NW, r)
    ctx.nw_run(all_pairs(int(g["n"])), 10, float(g["thr"]))
    files = ca.pipeline.fused_output_files(ctx)
    assert len(files) == int(g["nfiles"])
    for f, (p, d) in enumerate(files):
        assert len(d) == int(g[f"f{f}_count"])
        assert np.array_equal(canon(p, d)[:, 0], bits(d).astype(np.int64))  # already (dist, i, j)
        assert np.array_equal(canon(p, d)[:, 1:], p.astype(np.int64))
        assert digest(p, d) == str(g[f"f{f}_sha256"])
    # small chunks and flush sizes: files end after the chunks that bring a
    # file to the flush size
    ends = ctx.nw_file_split(50000, 30000)
    sel = np.array([O.passes(float(x), float(g["thr"])) for x in
                    ctx.nw_dists(all_pairs(int(g["n"])), 10)], bool)
    want, acc, c0 = [], 0, 0
    for c in range(0, len(sel), 50000):
        acc += int(sel[c:c + 50000].sum())
        if acc - c0 >= 30000:
            want.append(acc)
            c0 = acc
    if acc > c0:
        want.append(acc)
    assert len(want) > 2 and np.array_equal(ends, np.array(want, np.int64))
    # fused run against the oracle's files
    seqs, _, _ = synth.make("reads1k")
    r = ca.Reads.from_list(seqs)
    ctx.upload_reads(ca.KMER, r)
    ctx.upload_reads(ca.NW, r)
    ctx.kmer_run(6, ca.kmer_default_threshold(r.n))
    kp, _ = ctx.kmer_fetch(dists=False)
    nthr = ca.nw_default_threshold(r.n)
    ctx.nw_run_on_kmer(10, nthr)
    files = ca.pipeline.fused_output_files(ctx)
    # into reused host buffers (grown once, then views): the same files
    buf = ca.pipeline.OutputBuffers()
    buf.reserve(7)
    for _ in range(2):
        again = ca.pipeline.fused_output_files(ctx, buf)
        assert len(again) == len(files)
        for (p0, d0), (p1, d1) in zip(files, again):
            assert np.array_equal(p0, p1) and np.array_equal(bits(d0), bits(d1))
    ro = O.Reads.from_list(seqs)
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        nf, tot = O.gendist_files(ro, kp, 10, nthr, os.path.join(td, "o"))
        assert nf == len(files) and tot == sum(len(d) for _, d in files)
        for f, (p, d) in enumerate(files):
            op = np.fromfile(os.path.join(td, f"o.npair_{f}"), np.uint32).reshape(-1, 2)
            od = np.fromfile(os.path.join(td, f"o.ndist_{f}"), np.float32)
            assert np.array_equal(p.astype(np.uint32), op) and np.array_equal(bits(d), bits(od))


def test_pinned_host_arrays():
    """crispy_host_alloc-backed numpy arrays: shapes and dtypes as asked,
    writable, zero-size allowed, and a view keeps its block alive after the
    array it came from is gone."""
    say("pinned host arrays")
    a = ca.pinned_empty((1000, 2), np.int32)
    assert a.shape == (1000, 2) and a.dtype == np.int32 and a.flags.c_contiguous
    a[:] = np.arange(2000, dtype=np.int32).reshape(1000, 2)
    v = a[10:20]
    del a
    import gc
    gc.collect()
    assert np.array_equal(v, np.arange(20, 40, dtype=np.int32).reshape(10, 2))
    f = ca.pinned_empty(5, np.float32)
    f[:] = 1.5
    assert f.shape == (5,) and float(f.sum()) == 7.5
    z = ca.pinned_empty((0, 2), np.int32)
    assert z.shape == (0, 2)


def test_python_mirror_matches_cli(ctx, tmp_path):
    say("python mirror")
    fa = tmp_path / "c.fa"
    fa.write_bytes(open(CLEAN, "rb").read())
    run([os.path.join(BIN, "kmerDist"), "-i", str(fa), "-k", "6"], tmp_path)
    run([os.path.join(BIN, "genDist"), "-i", str(fa), "-b", "10"], tmp_path)
    r = ca.read_fasta(str(fa), ca.KMER)
    n = ca.compute_kmer_dist(ctx, r, 6, ca.kmer_default_threshold(r.n), str(tmp_path / "py.kpair"))
    assert open(tmp_path / "py.kpair", "rb").read() == open(str(fa) + ".kpair", "rb").read()
    rg = ca.read_fasta(str(fa), ca.NW)
    nf, tot = ca.compute_gen_dist(ctx, rg, str(fa) + ".kpair", str(tmp_path / "py.npair"),
                                  str(tmp_path / "py.ndist"), ca.nw_default_threshold(rg.n), 10)
    assert nf == 1 and tot > 0 and n > 0
    for ext in ("npair_0", "ndist_0"):
        assert open(tmp_path / f"py.{ext}", "rb").read() == open(f"{fa}.{ext}", "rb").read()


@pytest.mark.parametrize("workload,K,rows", [("reads20k", 6, 256), ("reads100k", 6, 256), ("reads100k", 8, 256),
                                           ("reads250k", 6, 256)])
def test_full_size(ctx, workload, K, rows):
    """The benchmark workloads at full size (20k, 100k and 250k reads): exact
    against the oracle on `rows` seeded random rows of the k-mer matrix spread
    over [0, n) plus its first and last four rows (edge tiles), and on a
    4,000-pair sample of NW pairs; invariants of the whole output (i < j,
    (i, j) order, thresholds, NW pairs a subset of the k-mer pairs), and the
    same output from a second run."""
    say(f"{workload} full size, k={K}")
    seqs, _, cfg = synth.make(workload)
    band = cfg["band"]
    r = ca.Reads.from_list(seqs)
    ro = O.Reads.from_list(seqs)
    kthr = ca.kmer_default_threshold(r.n)
    nthr = ca.nw_default_threshold(r.n)
    ctx.upload_reads(ca.KMER, r)
    ctx.upload_reads(ca.NW, r)
    c = ctx.kmer_run(K, kthr)
    p, d = ctx.kmer_fetch()
    say("kmer pairs", c)
    assert c == len(d) > 0
    assert np.all(p[:, 0] < p[:, 1]) and p[:, 1].max() < r.n
    key = p[:, 0].astype(np.int64) * r.n + p[:, 1]
    assert np.all(np.diff(key) > 0)
    assert all(O.passes(float(x), kthr) for x in d[:: max(1, len(d) // 5000)])
    rng = np.random.default_rng(1000 + K + r.n)
    sample_rows = np.unique(np.concatenate([rng.choice(r.n - 1, size=rows, replace=False),
                                            [0, 1, 2, 3, r.n - 5, r.n - 4, r.n - 3, r.n - 2]]))
    op, od = O.kmer_rows(ro, K, kthr, sample_rows)
    sel = np.isin(p[:, 0], sample_rows)
    say(f"{len(sample_rows)} sampled rows: {int(sel.sum())} pairs")
    assert np.array_equal(canon(p[sel], d[sel]), canon(op, od))
    nc = ctx.nw_run_on_kmer(band, nthr)
    np_, nd = ctx.nw_fetch()
    say("nw pairs", nc)
    knum = p[:, 0].astype(np.int64) * r.n + p[:, 1]
    nnum = np_[:, 0].astype(np.int64) * r.n + np_[:, 1]
    assert np.all(np.diff(nnum) > 0) and np.all(np.isin(nnum, knum))
    assert all(O.passes(float(x), nthr) for x in nd[:: max(1, len(nd) // 5000)])
    # a sample agrees with the oracle, kept exactly when it passes
    rng = np.random.default_rng(9)
    samp = rng.choice(len(p), size=min(4000, len(p)), replace=False)
    od, _, _ = O.nw_pairs(ro, p[samp], band)
    full = ctx.nw_dists(p[samp], band)
    assert np.array_equal(bits(full), bits(od))
    kept = {(int(a), int(b)): float(x) for (a, b), x in zip(np_, nd)}
    for (a, b), x in zip(p[samp], od):
        if O.passes(float(x), nthr):
            assert kept[(int(a), int(b))] == float(x)
        else:
            assert (int(a), int(b)) not in kept
    # a second run gives the same bytes
    ctx.kmer_run(K, kthr)
    p2, d2 = ctx.kmer_fetch()
    ctx.nw_run_on_kmer(band, nthr)
    np2, nd2 = ctx.nw_fetch()
    assert np.array_equal(p2, p) and np.array_equal(bits(d2), bits(d))
    assert np.array_equal(np2, np_) and np.array_equal(bits(nd2), bits(nd))


REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def _partitions(path):
    """sparsecut's .Cluster as {cutoff: sorted clusters} (its cluster order
    follows an unordered_set of tree nodes, so only the partition is defined)."""
    out = {}
    for line in open(path):
        f = line.strip().split("|")
        if f[0]:
            out[f[0].strip()] = sorted(tuple(sorted(int(x) for x in c.split())) for c in f[1:] if c.strip())
    return out


def _consumers(cwd, name, n, nfiles):
    """crispy.sh:22-29 after the two distance stages: aveclust, then sparsecut
    at fixed cutoffs (the R changepoint step that picks them is not run)."""
    run([os.path.join(REF_BIN, "aveclust"), "-i", name, "-n", str(n), "-f", str(nfiles)], cwd)
    with open(os.path.join(cwd, "cut.txt"), "w") as f:
        f.write("0.03\n0.05\n")
    run([os.path.join(REF_BIN, "sparsecut"), "-i", name, "-n", str(n), "-c", "cut.txt"], cwd)
    return open(os.path.join(cwd, name + "_Merge"), "rb").read(), _partitions(os.path.join(cwd, name + ".Cluster"))


@pytest.mark.skipif(not all(os.access(os.path.join(REF_BIN, b), os.X_OK) for b in ("aveclust", "sparsecut")),
                    reason="the reference aveclust / sparsecut (built from its sources by make -C oracle ref) "
                           "are not in oracle/_ref")
def test_dropin_consumers(tmp_path):
    """The drop-in loop of crispy.sh:18-29 on the shipped sample: bin/preprocess,
    bin/kmerDist -k 6 and bin/genDist -b 10, then the reference's own aveclust
    (k-way merge of the per-file-sorted .npair_f/.ndist_f, aveclust.cpp:747-802)
    and sparsecut.  The consumers produce the same _Merge (byte for byte) and
    the same clusters at each cutoff as on the oracle's genDist files.  Equal
    distances inside a file may come in any order from the reference's unstable
    sort: with them shuffled, the clusters at the cutoffs are still the same
    (the dendrogram's merge order, and so _Merge, can change)."""
    say("drop-in consumers")
    work = tmp_path / "gpu"
    work.mkdir()
    raw = work / "sample"
    raw.write_bytes(open(os.path.join(GOLD, "V2Mice_sample1000"), "rb").read())
    r = subprocess.run([os.path.join(BIN, "preprocess"), "-i", str(raw)], cwd=work, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 1, r.stdout[-2000:]  # the reference preprocess returns 1 on success
    name = "sample_Clean"
    run([os.path.join(BIN, "kmerDist"), "-i", name, "-k", "6"], work)
    run([os.path.join(BIN, "genDist"), "-i", name, "-b", "10"], work)
    nfiles = len([f for f in os.listdir(work) if f.startswith(name + ".ndist_")])
    n = sum(1 for line in open(work / name, "rb") if line.startswith(b">"))
    merge, parts = _consumers(str(work), name, n, nfiles)
    # the same consumers on the oracle's genDist files of the same reads
    owork = tmp_path / "oracle"
    owork.mkdir()
    (owork / name).write_bytes((work / name).read_bytes())
    ro, go = O.read_fasta(str(owork / name), "kmer"), O.read_fasta(str(owork / name), "gen")
    kp, _ = O.kmer_all(ro, 6, O.kmer_threshold(ro.n), nthreads=O.num_threads())
    kp = kp[np.lexsort((kp[:, 1], kp[:, 0]))]
    onf, _ = O.gendist_files(go, kp, 10, O.nw_threshold(go.n), str(owork / name))
    assert onf == nfiles
    omerge, oparts = _consumers(str(owork), name, n, nfiles)
    assert merge == omerge
    assert parts == oparts and len(parts) == 2
    # another tie order inside each file: the same clusters
    swork = tmp_path / "shuffled"
    swork.mkdir()
    (swork / name).write_bytes((work / name).read_bytes())
    rng = np.random.default_rng(3)
    for f in range(nfiles):
        p = np.fromfile(work / f"{name}.npair_{f}", np.uint32).reshape(-1, 2)
        d = np.fromfile(work / f"{name}.ndist_{f}", np.float32)
        key = np.lexsort((rng.random(len(d)), d.view(np.uint32)))
        p[key].tofile(swork / f"{name}.npair_{f}")
        d[key].tofile(swork / f"{name}.ndist_{f}")
    _, sparts = _consumers(str(swork), name, n, nfiles)
    assert sparts == parts
