"""GPU tests of the input side: the device merge (crispy_preprocess_order)
and primer filter (crispy_preprocess_primer_filter), and bin/preprocess,
against the reference preprocess's outputs, the oracle and a Python stable
(length, bytes) merge."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import crispy_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
PRE = os.path.join(GOLD, "pre")
BIN = os.path.join(ROOT, "bin")
CASES = [("pre_adv.fa", "pre_adv.fa_Clean", []),
         ("pre_dups.fa", "pre_dups.fa_Clean", []),
         ("pre_dups.fa", "pre_dups.fa_s1", ["-s", "1.0"]),
         ("pre_reads.fq", "pre_reads.fq_Clean", ["-f", "q"]),
         ("pre_primer.fa", "pre_primer.fa_Clean", ["-p", "pre_primer.txt"]),
         ("pre_primer.fa", "pre_primer.fa_m0", ["-p", "pre_primer.txt", "-m", "0"]),
         ("pre_primer.fa", "pre_primer.fa_m4s1", ["-p", "pre_primer.txt", "-m", "4", "-s", "1.0"])]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ca.Context(0)
    yield c
    c.close()


def py_merge(seqs, freq):
    keep = sorted((i for i in range(len(seqs)) if freq[i] != 0), key=lambda i: (len(seqs[i]), seqs[i], i))
    order, f = [], []
    for i in keep:
        if order and seqs[order[-1]] == seqs[i]:
            order[-1] = i
            f[-1] += int(freq[i])
        else:
            order.append(i)
            f.append(int(freq[i]))
    return np.array(order, np.int32), np.array(f, np.int32)


@pytest.mark.parametrize("inp,out,flags", CASES)
def test_cli_matches_reference(tmp_path, inp, out, flags):
    shutil.copyfile(os.path.join(PRE, inp), tmp_path / inp)
    if "-p" in flags:
        shutil.copyfile(os.path.join(PRE, "pre_primer.txt"), tmp_path / "pre_primer.txt")
    args = [os.path.join(BIN, "preprocess"), "-i", inp] + flags
    if not out.endswith("_Clean"):
        args += ["-o", out]
    say("preprocess", args)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, timeout=120)
    assert r.returncode == 1, r.stderr  # the reference's main returns 1
    for suffix in ("", ".frq"):
        assert (tmp_path / (out + suffix)).read_bytes() == open(os.path.join(PRE, out + suffix), "rb").read()


def test_cli_raw_sample_then_kmerdist(tmp_path):
    """crispy.sh's first two steps on the shipped raw sample: preprocess gives
    the reference's _Clean, and kmerDist on it the reference's pairs."""
    shutil.copyfile(os.path.join(GOLD, "V2Mice_sample1000"), tmp_path / "V2Mice_sample1000")
    r = subprocess.run([os.path.join(BIN, "preprocess"), "-i", "V2Mice_sample1000"], cwd=tmp_path,
                       capture_output=True, timeout=120)
    assert r.returncode == 1
    clean = tmp_path / "V2Mice_sample1000_Clean"
    assert clean.read_bytes() == open(os.path.join(GOLD, "V2Mice_sample1000_Clean"), "rb").read()
    r = subprocess.run([os.path.join(BIN, "kmerDist"), "-i", "V2Mice_sample1000_Clean", "-k", "6"],
                       cwd=tmp_path, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(str(clean) + ".kpair", np.int32).reshape(-1, 2)
    want = np.load(os.path.join(GOLD, "kmer_v2mice.npz"))["k6_pairs"].reshape(-1, 2)
    key = lambda p: sorted(map(tuple, p.tolist()))
    assert key(got) == key(want)


def test_python_mirror(ctx, tmp_path):
    out = str(tmp_path / "m_Clean")
    n = ca.preprocess(ctx, os.path.join(PRE, "pre_dups.fa"), out, sd_allowed=1.0)
    assert n == 224
    assert open(out, "rb").read() == open(os.path.join(PRE, "pre_dups.fa_s1"), "rb").read()


def test_python_mirror_primer(ctx, tmp_path):
    out = str(tmp_path / "p_Clean")
    primer = ca.read_primer(os.path.join(PRE, "pre_primer.txt"))
    n = ca.preprocess(ctx, os.path.join(PRE, "pre_primer.fa"), out, sd_allowed=1.0, primer=primer,
                      max_mismatches=4)
    assert n == 93
    assert open(out, "rb").read() == open(os.path.join(PRE, "pre_primer.fa_m4s1"), "rb").read()


@pytest.mark.parametrize("seed,plen,mm", [(11, 20, 2), (12, 8, 0), (13, 35, 5), (14, 1, 0), (15, 120, 9),
                                          (16, 0, 3)])
def test_primer_filter_random_vs_oracle(ctx, seed, plen, mm):
    """Device -p filter against the oracle read by read: reads built from the
    primer with substitutions, indels, shifts and case changes, unrelated and
    empty reads, bytes outside ACGT, zero frequencies."""
    rng = np.random.default_rng(seed)
    alph = np.frombuffer(b"ACGTNacgtM", np.uint8)
    primer = bytes(rng.choice(alph[:4], plen).tobytes())
    seqs = []
    for k in range(3000):
        p = bytearray(primer)
        for _ in range(int(rng.integers(0, 4))):
            if p:
                op = int(rng.integers(0, 3))
                q = int(rng.integers(0, len(p)))
                if op == 0:
                    p[q] = int(rng.choice(alph))
                elif op == 1:
                    p.insert(q, int(rng.choice(alph)))
                else:
                    del p[q]
        u = rng.random()
        if u < 0.1:
            p = bytearray(rng.choice(alph, len(p)).tobytes())
        elif u < 0.2:
            p = bytearray(rng.choice(alph, int(rng.integers(1, 6))).tobytes()) + p
        elif u < 0.25:
            p = bytearray(bytes(p).lower())
        seq = bytes(p) + rng.choice(alph, int(rng.integers(0, 80))).tobytes()
        seqs.append(seq[:int(rng.integers(0, len(seq) + 1))] if k % 50 == 0 else seq)
    freq = (rng.random(len(seqs)) < 0.95).astype(np.int32) * rng.integers(1, 3, len(seqs)).astype(np.int32)
    rec = ca.Records(ca.Reads.from_list(seqs), b"", np.zeros(len(seqs) + 1, np.int64), freq.copy())
    d = ctx.preprocess_primer_filter(rec, primer, mm)
    want = np.array([f if f == 0 or O.primer_keep(primer, s, mm) else 0 for s, f in zip(seqs, freq)], np.int32)
    say(f"primer {plen} bases, -m {mm}: {d} of {int((freq != 0).sum())} discarded")
    assert np.array_equal(rec.freq, want)
    assert d == int(((freq != 0) & (want == 0)).sum())


def rand_records(rng, n, maxlen, alphabet, dup_frac):
    base = [bytes(rng.choice(alphabet, int(L)).tobytes()) for L in rng.integers(0, maxlen + 1, max(1, n // 3))]
    seqs = []
    for k in range(n):
        if seqs and rng.random() < dup_frac:
            seqs.append(seqs[int(rng.integers(0, len(seqs)))])
        else:
            seqs.append(base[int(rng.integers(0, len(base)))] if rng.random() < 0.5 else
                        bytes(rng.choice(alphabet, int(rng.integers(0, maxlen + 1))).tobytes()))
    return seqs


@pytest.mark.parametrize("seed,n,maxlen,dup", [(1, 3000, 40, 0.3), (2, 2000, 600, 0.5), (3, 500, 5000, 0.2),
                                               (4, 5000, 6, 0.0), (5, 4000, 300, 0.9)])
def test_order_random_vs_python(ctx, seed, n, maxlen, dup):
    """Bytes incl. N, lowercase and >= 0x80 (strcmp compares unsigned), empty
    reads, zero frequencies, heavy duplication."""
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGTNa\x80\xff", np.uint8)
    seqs = rand_records(rng, n, maxlen, alphabet, dup)
    freq = (rng.random(n) < 0.9).astype(np.int32) * rng.integers(1, 4, n).astype(np.int32)
    reads = ca.Reads.from_list(seqs)
    order, f = ctx.preprocess_order(reads, freq)
    want_o, want_f = py_merge(seqs, freq)
    say(f"seed {seed}: {n} reads -> {len(order)}")
    assert np.array_equal(order, want_o) and np.array_equal(f, want_f)


def test_order_edges(ctx):
    empty = ca.Reads(b"", np.zeros(1, np.int64))
    o, f = ctx.preprocess_order(empty, np.zeros(0, np.int32))
    assert len(o) == 0
    same = ca.Reads.from_list([b"ACGT"] * 10000)
    o, f = ctx.preprocess_order(same, np.ones(10000, np.int32))
    assert o.tolist() == [9999] and f.tolist() == [10000]
    zero = ca.Reads.from_list([b"A", b"C"])
    o, f = ctx.preprocess_order(zero, np.zeros(2, np.int32))
    assert len(o) == 0
    distinct = [bytes([65 + (k % 4)]) * (1 + k // 4) for k in range(4000)]
    o, f = ctx.preprocess_order(ca.Reads.from_list(distinct), np.ones(4000, np.int32))
    assert np.array_equal(o, py_merge(distinct, np.ones(4000))[0]) and np.all(f == 1)
