"""Seeded synthetic 16S-like read sets, written in the reference's
preprocess order.

The reference ships one input-only sample (data/V2Mice_sample1000) and no
benchmark inputs, so workloads are generated:

* a uniform ACGT backbone (1500 bp);
* a window of the backbone, as long as the longest read, from a seeded offset;
* `species` copies of the window, each with per-base substitution probability
  drawn from U[div_lo, div_hi];
* each read picks a species uniformly, gets per-base noise (substitution
  p_sub, indel p_indel split half insertion / half deletion, inserted bases
  uniform) and is cut, or extended with uniform bases, to a length drawn from
  the length law.

The output is what the reference's preprocess stage would write for it:
exact duplicates merged and reads sorted by (length, strcmp) (preprocess.cpp:
42-50, 244-280), headers "> label" (preprocess.cpp:562-601).  Read ids are the
positions in that order, so i < j implies len(i) <= len(j).
"""
import numpy as np

LETTERS = np.frombuffer(b"AGTC", dtype=np.uint8)

# Workloads: name -> generator parameters plus the stage parameters.
CONFIGS = {
    # 1,000 reads, 200-300 bp, 20 species at 3-10 %, k=6, band 10
    "reads1k": dict(n=1000, seed=1, species=20, div=(0.03, 0.10), law=("uniform", 200, 300),
                    p_sub=0.01, p_indel=0.005, K=6, band=10),
    # 20,000 reads, 200-300 bp, 100 species at 3-10 %, k=6, band 10
    "reads20k": dict(n=20000, seed=2, species=100, div=(0.03, 0.10), law=("uniform", 200, 300),
                     p_sub=0.01, p_indel=0.005, K=6, band=10),
    # 100,000 reads, log-normal lengths (median 300, sigma 0.35) clipped to 150-550,
    # 500 species at 1-15 %, k=6 (and 8), band 10
    "reads100k": dict(n=100000, seed=3, species=500, div=(0.01, 0.15),
                      law=("lognormal", 300, 0.35, 150, 550),
                      p_sub=0.01, p_indel=0.005, K=6, band=10),
    # 250,000 reads, 400-900 bp, 2,000 species at 1-15 %, 2 % noise, k=6, band 5
    "reads250k": dict(n=250000, seed=5, species=2000, div=(0.01, 0.15), law=("uniform", 400, 900),
                      p_sub=0.015, p_indel=0.005, K=6, band=5),
}


def _lengths(rng, n, law):
    if law[0] == "uniform":
        return rng.integers(law[1], law[2] + 1, n)
    if law[0] == "lognormal":
        med, sigma, lo, hi = law[1:]
        x = np.rint(med * np.exp(sigma * rng.standard_normal(n)))
        return np.clip(x, lo, hi).astype(np.int64)
    raise ValueError(law)


def _law_max(law):
    return law[2] if law[0] == "uniform" else law[4]


def generate(n, seed, species, div, law, p_sub, p_indel, backbone_len=1500, batch=8192,
             **_unused):
    """Returns the list of read sequences (bytes) in preprocess order and a
    parallel list of labels."""
    rng = np.random.default_rng(seed)
    backbone = rng.integers(0, 4, backbone_len).astype(np.uint8)
    win = int(_law_max(law))
    off = int(rng.integers(0, backbone_len - win + 1))
    window = backbone[off:off + win]
    dv = rng.uniform(div[0], div[1], species)
    sp = np.tile(window, (species, 1))
    m = rng.random(sp.shape) < dv[:, None]
    sp[m] = (sp[m] + rng.integers(1, 4, int(m.sum()))) % 4
    lengths = _lengths(rng, n, law)
    sid = rng.integers(0, species, n)

    seqs = []
    for b0 in range(0, n, batch):
        b1 = min(n, b0 + batch)
        base = sp[sid[b0:b1]].copy()
        u = rng.random(base.shape)
        sub = u < p_sub
        base[sub] = (base[sub] + rng.integers(1, 4, int(sub.sum()))) % 4
        ins = (u >= p_sub) & (u < p_sub + p_indel / 2)
        dele = (u >= p_sub + p_indel / 2) & (u < p_sub + p_indel)
        reps = np.ones(base.shape, np.int64)
        reps[dele] = 0
        reps[ins] = 2
        flat = np.repeat(base.ravel(), reps.ravel())
        ends = np.cumsum(reps.ravel())
        ins_pos = ends[ins.ravel()] - 1          # the second copy of an inserted base
        flat[ins_pos] = rng.integers(0, 4, len(ins_pos))
        row_len = reps.sum(axis=1)
        row_start = np.concatenate([[0], np.cumsum(row_len)[:-1]])
        tgt = lengths[b0:b1]
        ext = rng.integers(0, 4, (b1 - b0, int(tgt.max())))
        for r in range(b1 - b0):
            L = int(tgt[r])
            have = min(L, int(row_len[r]))
            s = np.empty(L, np.uint8)
            s[:have] = flat[row_start[r]:row_start[r] + have]
            s[have:] = ext[r, have:L]
            seqs.append((LETTERS[s].tobytes(), int(sid[b0 + r])))

    # preprocess: merge exact duplicates (first label kept), sort by (len, strcmp)
    seen = {}
    for idx, (s, spc) in enumerate(seqs):
        if s not in seen:
            seen[s] = f"r{idx}_s{spc}"
    order = sorted(seen.keys(), key=lambda s: (len(s), s))
    return order, [seen[s] for s in order]


def write_fasta(path, seqs, labels):
    with open(path, "wb") as f:
        for s, lab in zip(seqs, labels):
            f.write(b"> " + lab.encode() + b"\n" + s + b"\n")


def make(name, path=None, n=None):
    """Generates workload `name` (optionally only its first `n` reads before
    dedupe); writes it to `path` when given.  Returns (seqs, labels, cfg)."""
    cfg = dict(CONFIGS[name])
    if n is not None:
        cfg["n"] = n
    seqs, labels = generate(**cfg)
    if path is not None:
        write_fasta(path, seqs, labels)
    return seqs, labels, cfg
