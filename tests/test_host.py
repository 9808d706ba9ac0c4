"""CPU-only checks of the product library: it loads, exports the ABI, and its
host-side parts (FASTA reader, thresholds, predicate, band size, command
line error paths) agree with the oracle.  No device calls."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import crispy_amd as ca
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
HEADER = os.path.join(ROOT, "include", "crispy_hip.h")
BIN = os.path.join(ROOT, "bin")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(crispy_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ca.library_path())
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


FASTAS = ["V2Mice_sample1000", "V2Mice_sample1000_Clean", "adversarial.fa"]


@pytest.mark.parametrize("name", FASTAS)
@pytest.mark.parametrize("variant", [ca.KMER, ca.NW])
def test_reader_matches_oracle(name, variant):
    path = os.path.join(GOLD, name)
    mine = ca.read_fasta(path, variant)
    ref = O.read_fasta(path, "kmer" if variant == ca.KMER else "gen")
    assert mine.n == ref.n
    assert mine.bases == ref.bases
    assert np.array_equal(mine.offsets, ref.offsets)


def test_reader_edge_files(tmp_path):
    cases = {
        "crlf.fa": b">a\r\nACGT\r\nTTGG\r\n>b\r\nGGCC\r\n",
        "blank_header.fa": b">a\n>b\nACGT\n>c\nAC\n",
        "dollar.fa": b">a\nACGT\n$stop\n>b\nGGGG\n",
        "trailing_ws.fa": b">a\nACGTNN  \t\nAAA\n\n\n",
        "no_newline.fa": b">a\nAC\n>b\nACGTACGT",
        "long_line.fa": b">a\n" + b"ACGTN" * 1000 + b"\n>b\nACGT\n",
        "lower.fa": b">a\nacgtnACGTn\n",
    }
    for name, data in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        for variant in (ca.KMER, ca.NW):
            mine = ca.read_fasta(str(p), variant)
            ref = O.read_fasta(str(p), "kmer" if variant == ca.KMER else "gen")
            assert mine.bases == ref.bases, (name, variant)
            assert np.array_equal(mine.offsets, ref.offsets), (name, variant)


def test_reader_errors(tmp_path):
    with pytest.raises(ca.CrispyError):
        ca.read_fasta(str(tmp_path / "missing.fa"))
    p = tmp_path / "bad.fa"
    p.write_bytes(b"ACGT\n")
    with pytest.raises(ca.CrispyError):
        ca.read_fasta(str(p))


def test_thresholds_match_oracle():
    for n in list(range(2, 1200, 7)) + [20000, 100000, 250000, 1048576]:
        assert ca.kmer_default_threshold(n) == O.kmer_threshold(n)
        assert ca.nw_default_threshold(n) == O.nw_threshold(n)


def test_predicate_matches_oracle():
    rng = np.random.default_rng(3)
    thr = rng.random(2000).astype(np.float32)
    d = (thr + rng.normal(0, 2e-5, 2000)).astype(np.float32)
    for a, b in zip(d, thr):
        assert ca.passes(float(a), float(b)) == O.passes(float(a), float(b))


def test_band_cells_brute_force():
    for L1 in range(1, 30):
        for L2 in range(1, 30):
            for band in (1, 2, 3, 10):
                kb = (L1 + L2) // (2 * band)
                exp = sum(1 for i in range(1, L2 + 1) for j in range(1, L1 + 1)
                          if band <= 1 or abs(i - j) <= kb)
                assert ca.band_cells(L1, L2, band) == exp, (L1, L2, band)
    assert ca.band_cells(900, 900, 5) == 292320
    assert ca.band_cells(250, 250, 10) == 12100
    assert ca.band_cells(400, 900, 5) == 95885


def run_bin(args, cwd):
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_cli_error_paths(tmp_path):
    km, gd = os.path.join(BIN, "kmerDist"), os.path.join(BIN, "genDist")
    fa = tmp_path / "x.fa"
    fa.write_bytes(b">a\nACGTACGTAC\n>b\nACGTACGTAA\n")
    assert run_bin([km, "-h"], tmp_path).returncode == 255
    assert run_bin([km, "-i", str(fa), "-c"], tmp_path).returncode == 2
    assert run_bin([km, "-i", str(tmp_path / "missing")], tmp_path).returncode == 255
    assert run_bin([km, "-i", str(fa), "-k", "9"], tmp_path).returncode == 255
    short = tmp_path / "short.fa"
    short.write_bytes(b">a\nACGTA\n>b\nACGTACGT\n")
    assert run_bin([km, "-i", str(short), "-k", "6"], tmp_path).returncode == 255  # read < K
    assert run_bin([gd, "-h"], tmp_path).returncode == 0
    assert run_bin([gd, "-i", str(fa), "-c"], tmp_path).returncode == 2
    assert run_bin([gd, "-i", str(fa)], tmp_path).returncode == 255  # no .kpair


def test_server_job_estimate():
    """Admission estimate: grows with reads, pairs, K > 6 lists, and adds the
    generic-kernel scratch for full alignment or non-ACGT reads."""
    rng = np.random.default_rng(5)
    seqs = [bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), int(L)).tobytes())
            for L in rng.integers(150, 550, 3000)]
    small = ca.Reads.from_list(seqs[:300])
    big = ca.Reads.from_list(seqs)
    assert ca.job_bytes(ca.Reads(b"", np.zeros(1, np.int64)), 6, 10) == 0
    e_small, e_big = ca.job_bytes(small, 6, 10), ca.job_bytes(big, 6, 10)
    assert 0 < e_small < e_big
    # the bitmap alone: 4^6 bits per read
    assert e_big > big.n * 512 + 2 * len(big.bases)
    assert ca.job_bytes(big, 8, 10) > e_big
    full = ca.job_bytes(big, 6, 1) - e_big
    assert full == 8 * min(256 * big.n, 32768) * (max(np.diff(big.offsets)) + 1)
    dirty = ca.Reads.from_list([seqs[0][:100] + b"N" + seqs[0][101:]] + seqs[1:])
    assert ca.job_bytes(dirty, 6, 10) - ca.job_bytes(big, 6, 10) == full


PRE = os.path.join(GOLD, "pre")
PRE_CASES = [("pre_adv.fa", "pre_adv.fa_Clean", False, None),
             ("pre_dups.fa", "pre_dups.fa_Clean", False, None),
             ("pre_dups.fa", "pre_dups.fa_s1", False, 1.0),
             ("pre_reads.fq", "pre_reads.fq_Clean", True, None),
             ("../V2Mice_sample1000", "../V2Mice_sample1000_Clean", False, None)]


def py_merge(rec):
    """Stable (length, bytes) order, runs collapsed onto their last read."""
    seqs = [rec.reads.seq(i) for i in range(rec.n)]
    keep = sorted((i for i in range(rec.n) if rec.freq[i] != 0), key=lambda i: (len(seqs[i]), seqs[i], i))
    order, freq = [], []
    for k, i in enumerate(keep):
        if order and seqs[order[-1]] == seqs[i]:
            order[-1] = i
            freq[-1] += int(rec.freq[i])
        else:
            order.append(i)
            freq.append(int(rec.freq[i]))
    return np.array(order, np.int32), np.array(freq, np.int32)


@pytest.mark.parametrize("inp,out,fastq,sd", PRE_CASES)
def test_preprocess_host_parts(tmp_path, inp, out, fastq, sd):
    """The product's reader, -s filter and writer around a Python merge give
    the reference's files (the device merge is checked on the GPU)."""
    rec = ca.preprocess_read(os.path.join(PRE, inp), fastq)
    if sd is not None:
        ca.preprocess_length_filter(rec, sd)
    order, freq = py_merge(rec)
    o = str(tmp_path / "o_Clean")
    ca.preprocess_write(o, rec, order, freq, fastq)
    assert open(o, "rb").read() == open(os.path.join(PRE, out), "rb").read()
    assert open(o + ".frq", "rb").read() == open(os.path.join(PRE, out) + ".frq", "rb").read()


def test_read_primer(tmp_path):
    """The primer line as preprocess.cpp:672-683 reads it."""
    f = tmp_path / "p.txt"
    for data, want in [(b"ACGT \t\r\nGG\n", b"ACGT"), (b"ACGT", b"ACGT"), (b"", b""),
                       (b"   \nACGT\n", b"   "), (b"AC GT\n", b"AC GT")]:
        f.write_bytes(data)
        assert ca.read_primer(str(f)) == want


def test_preprocess_cli_error_paths(tmp_path):
    exe = os.path.join(BIN, "preprocess")
    assert subprocess.run([exe, "-h"], capture_output=True).returncode == 255
    # a primer file that cannot be opened ends the run before the reads are read
    assert subprocess.run([exe, "-i", "x.fa", "-p", str(tmp_path / "no_primer.txt")],
                          capture_output=True).returncode == 255
    assert subprocess.run([exe, "-i", str(tmp_path / "missing.fa")], capture_output=True).returncode == 255
    bad = tmp_path / "bad.fa"
    bad.write_bytes(b"ACGT\n")
    assert subprocess.run([exe, "-i", str(bad)], capture_output=True).returncode == 255
