"""Writes the golden fixtures in tests/golden/ from the reference's own CPU
path (kmerDist -c, preprocess), compiled with g++ from the reference sources
into oracle/_ref by `make -C oracle ref`.

The genDist fixtures (NW distances and .npair_f/.ndist_f files) come from
`make -C oracle ref-gendist-rocthrust`: genCPU.cpp compiled against rocThrust,
because the NVIDIA Thrust it includes is not in this image.  That is not the
reference's own build, so those fixtures are regression data only; the NW
oracle is pinned by hand-derived cases (tests/test_oracle_nw_rules.py).

Run from the repository root in a container that has the reference sources:
    python tests/golden/make_golden.py
The fixtures are data only (inputs and the reference binaries' outputs, as
.npz arrays without pickles, FASTA text, and sha256 digests of large outputs).
"""
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from crispy_amd import synth  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden")
REF = os.path.join(ROOT, "oracle", "_ref")
REF_DATA = "/root/reference/data/V2Mice_sample1000"


def run(args, cwd, ok=(0,)):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True)
    if r.returncode not in ok:
        raise RuntimeError(f"{args} failed: {r.stdout}\n{r.stderr}")
    return r.stdout


def kmer_ref(fa, K, cwd, thr=None):
    args = [os.path.join(REF, "kmerDist"), "-i", fa, "-k", str(K), "-c", "-d", "-n", "1"]
    if thr is not None:
        args += ["-t", str(thr)]
    run(args, cwd)
    p = np.fromfile(fa + ".kpair", np.int32).reshape(-1, 2)
    d = np.fromfile(fa + ".kdist", np.float32)
    return p, d


def gen_ref(fa, band, cwd, thr=None):
    for f in os.listdir(cwd):
        if ".npair_" in f or ".ndist_" in f:
            os.remove(os.path.join(cwd, f))
    args = [os.path.join(REF, "genDist"), "-i", fa, "-b", str(band), "-c", "-n", "8"]
    if thr is not None:
        args += ["-t", str(thr)]
    run(args, cwd)
    files = []
    f = 0
    while os.path.exists(f"{fa}.npair_{f}"):
        p = np.fromfile(f"{fa}.npair_{f}", np.uint32).reshape(-1, 2)
        d = np.fromfile(f"{fa}.ndist_{f}", np.float32)
        files.append((p, d))
        f += 1
    return files


def canon_digest(p, d):
    """sha256 of the entries sorted by (dist bits, i, j)."""
    key = np.lexsort((p[:, 1], p[:, 0], d.view(np.uint32)))
    rec = np.zeros(len(d), dtype=[("d", "<u4"), ("i", "<u4"), ("j", "<u4")])
    rec["d"] = d.view(np.uint32)[key]
    rec["i"] = p[key, 0]
    rec["j"] = p[key, 1]
    return hashlib.sha256(rec.tobytes()).hexdigest()


def all_pairs(n):
    i, j = np.triu_indices(n, 1)
    return np.stack([i, j], 1).astype(np.int32)


def write_fasta_subset(src_fa, dst_fa, nreads):
    out = []
    with open(src_fa, "rb") as f:
        lines = f.read().split(b"\n")
    k = 0
    for ln in range(0, len(lines) - 1, 2):
        if k == nreads:
            break
        out.append(lines[ln] + b"\n" + lines[ln + 1] + b"\n")
        k += 1
    with open(dst_fa, "wb") as f:
        f.write(b"".join(out))


ADVERSARIAL = (
    b">adv0 poly-A then mixed\nAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAAA\nACGTACGGT\n"
    b">adv1 lowercase and N inside\nACGTacgtNNACGTTTGCAnnACGTGGCCAATTGCGCATGCA\nTTGACCA\n"
    b">adv2 trailing N and spaces  \r\nACGTTGCAACGTGGATCCAAGGTTNN  \r\nGGCCTTAANNN\n"
    b">adv3 blank line inside the record\nACGTTGCAACGTGGATCCAAGG\n\nTTGGCCAAGGTTACGTACGATCGATCG\n"
    b">adv4 short\nACGTACGTAC\n"
    b">adv5 tandem repeat\nACGACGACGACGACGACGACGACGACGACGACGACGACGACGACGACGACGACGACGACG\n"
    b">adv6 long\n" + b"ACGTTGCATGCAAGTCCGATCGATTAGCAGGCTTACGATCGGATTACCAGT" * 4 + b"\n"
    b">adv7 near copy of adv6\n" + b"ACGTTGCATGCAAGTCCGATCGATTAGCAGGCTTACGATCGGATTACCAGT" * 3
    + b"ACGTTGCATGCTAGTCCGATCGATTAGCAGGCTTACGATCGGATTACCAGTA\n"
    b">adv8 only N\nNNNNNNNNNNNNNNNNNN\n"
    b">adv9 T/C rich for the k=8 order\nTTTTCCCCTTTTCCCCTCTCTCTCGGGGAAAATTTTCCCCTCTC\n"
    b">adv10 T/C rich sibling\nTTTTCCCCTTTTCCCCTCTCTCTCAAAAGGGGTTTTCCCCTCTCT\n"
    b">adv11 mixed case tail\nGATTACAGATTACAGATTACAgattaca\n"
    b">adv12 identical to adv4\nACGTACGTAC\n"
    b">adv13 no final newline\nTTGACCAGGTACCATTGACCAGGTACCA"
)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    os.makedirs(GOLD, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        # 1. the shipped sample and its preprocess output
        shutil.copyfile(REF_DATA, os.path.join(GOLD, "V2Mice_sample1000"))
        shutil.copyfile(REF_DATA, os.path.join(tmp, "V2Mice_sample1000"))
        # preprocess's main returns 1 on success (preprocess.cpp:725)
        run([os.path.join(REF, "preprocess"), "-i", "V2Mice_sample1000"], tmp, ok=(0, 1))
        clean = os.path.join(tmp, "V2Mice_sample1000_Clean")
        shutil.copyfile(clean, os.path.join(GOLD, "V2Mice_sample1000_Clean"))
        shutil.copyfile(clean + ".frq", os.path.join(GOLD, "V2Mice_sample1000_Clean.frq"))

        # 2. kmerDist on the cleaned sample at the default threshold
        kv = {}
        for K in range(4, 9):
            p, d = kmer_ref(clean, K, tmp)
            kv[f"k{K}_pairs"] = p
            kv[f"k{K}_dists"] = d
        np.savez_compressed(os.path.join(GOLD, "kmer_v2mice.npz"), **kv)
        k6_pairs = kv["k6_pairs"]

        # 3. all-pairs k-mer distances (-t 1) on the first 120 cleaned reads, K = 1..8
        sub = os.path.join(tmp, "sub120.fa")
        write_fasta_subset(clean, sub, 120)
        ks = {}
        for K in range(1, 9):
            p, d = kmer_ref(sub, K, tmp, thr=1)
            assert np.array_equal(p, all_pairs(120)), "expected every pair in (i, j) order"
            ks[f"k{K}_dists"] = d
        np.savez_compressed(os.path.join(GOLD, "kmer_sub120.npz"), **ks)

        # 4. adversarial FASTA: reader edge cases, non-ACGT, repeats, k=8 order
        adv = os.path.join(tmp, "adversarial.fa")
        with open(adv, "wb") as f:
            f.write(ADVERSARIAL)
        shutil.copyfile(adv, os.path.join(GOLD, "adversarial.fa"))
        ka = {}
        for K in (3, 6, 8):
            p, d = kmer_ref(adv, K, tmp, thr=1)
            ka[f"k{K}_pairs"] = p
            ka[f"k{K}_dists"] = d
        n_adv = int(np.max(ka["k6_pairs"])) + 1
        allp = all_pairs(n_adv)
        allp.tofile(adv + ".kpair")
        for band in (1, 2, 5, 10):
            files = gen_ref(adv, band, tmp, thr=1)
            assert len(files) == 1
            ka[f"b{band}_pairs"] = files[0][0]
            ka[f"b{band}_dists"] = files[0][1]
        np.savez_compressed(os.path.join(GOLD, "adversarial.npz"), **ka)

        # 5. genDist on the cleaned sample, fed the k=6 .kpair
        k6_pairs.tofile(clean + ".kpair")
        gv = {}
        for band in (10, 5, 2, 1):
            files = gen_ref(clean, band, tmp)
            assert len(files) == 1
            gv[f"b{band}_pairs"] = files[0][0]
            gv[f"b{band}_dists"] = files[0][1]
        np.savez_compressed(os.path.join(GOLD, "gen_v2mice.npz"), **gv)

        # 6. all-pairs NW distances (-t 1) on the first 60 cleaned reads
        sub60 = os.path.join(tmp, "sub60.fa")
        write_fasta_subset(clean, sub60, 60)
        all_pairs(60).tofile(sub60 + ".kpair")
        gs = {}
        for band in (1, 2, 3, 5, 10, 20):
            files = gen_ref(sub60, band, tmp, thr=1)
            assert len(files) == 1
            gs[f"b{band}_pairs"] = files[0][0]
            gs[f"b{band}_dists"] = files[0][1]
        np.savez_compressed(os.path.join(GOLD, "gen_sub60.npz"), **gs)

        # 7. synthetic 1,000-read workload (seed 1): kmer k=6 + genDist band 10
        fa = os.path.join(tmp, "reads1k.fa")
        seqs, labels, _ = synth.make("reads1k", fa)
        fa_digest = hashlib.sha256(open(fa, "rb").read()).hexdigest()
        p, d = kmer_ref(fa, 6, tmp)
        files = gen_ref(fa, 10, tmp)
        sv = {"fasta_sha256": np.array(fa_digest), "k6_pairs": p, "k6_dists": d,
              "b10_nfiles": np.array(len(files))}
        for f, (fp, fd) in enumerate(files):
            sv[f"b10_f{f}_count"] = np.array(len(fd))
            sv[f"b10_f{f}_sha256"] = np.array(canon_digest(fp, fd))
            sv[f"b10_f{f}_sorted"] = np.array(bool(np.all(np.diff(fd) >= 0)))
        np.savez_compressed(os.path.join(GOLD, "synth_reads1k.npz"), **sv)

        # 8. more than one chunk of input pairs: every pair of 1,200 synthetic
        #    reads (719,400 > 589,824) at -t 1, so the file split is exercised
        fa2 = os.path.join(tmp, "reads1200.fa")
        seqs2, labels2, _ = synth.make("reads1k", fa2, n=1200)
        n2 = len(seqs2)
        pairs2 = all_pairs(n2)
        pairs2.tofile(fa2 + ".kpair")
        # At -t 1 the reference overflows its 2*147,456-entry host vector
        # within the first chunk (genCPU.cpp:339-341,478) and crashes, so the
        # threshold is chosen to pass ~40 % of the first 589,824 pairs: the
        # first chunk then flushes one file and the rest form a second one.
        from oracle import oracle as O
        g2 = O.read_fasta(fa2, "gen")
        d_all, _, _ = O.nw_pairs(g2, pairs2[:589824], 10)
        thr2 = float(np.round(np.quantile(d_all, 0.40), 3))
        npass = int(sum(O.passes(float(x), thr2) for x in d_all))
        assert 147456 <= npass <= 2 * 147456, npass
        files = gen_ref(fa2, 10, tmp, thr=thr2)
        mv = {"fasta_sha256": np.array(hashlib.sha256(open(fa2, "rb").read()).hexdigest()),
              "n": np.array(n2), "thr": np.array(thr2), "nfiles": np.array(len(files))}
        for f, (fp, fd) in enumerate(files):
            mv[f"f{f}_count"] = np.array(len(fd))
            mv[f"f{f}_sha256"] = np.array(canon_digest(fp, fd))
            mv[f"f{f}_sorted"] = np.array(bool(np.all(np.diff(fd) >= 0)))
        np.savez_compressed(os.path.join(GOLD, "synth_multifile.npz"), **mv)
        print("files:", len(files), [len(x[1]) for x in files])
    finally:
        shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
