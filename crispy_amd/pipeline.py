"""Python mirrors of the reference's two drivers, file behaviour included.

compute_kmer_dist  ~ computeKmerDist_CUDA (kmerMain.h:73, kmerCUDA.cu:14-126)
compute_gen_dist   ~ computeGenDist_CUDA  (genMain.h:77,  genCUDA.cu:83-270)

Both run on the GPU through the C ABI; the executables bin/kmerDist and
bin/genDist implement the same logic in C++ for the command line.
"""
import numpy as np

from .api import KMER, NW, Context, Reads

CHUNK_PAIRS = 576 * 256 * 4   # NUM_PAIRS, genMain.h:24-27 (pairs per .kpair chunk)
FLUSH_PAIRS = 576 * 256       # genCPU.cpp:339,484 (a file is written at >= this many)


def compute_kmer_dist(ctx: Context, reads: Reads, K: int, threshold: float, pair_path: str,
                      dist_path: str = None, block=1 << 20):
    """Writes <pair_path> (int32 i, j) and optionally <dist_path> (float32)
    for every pair i < j whose k-mer distance passes `threshold`, in (i, j)
    order.  Returns the number of pairs."""
    ctx.upload_reads(KMER, reads)
    count = ctx.kmer_run(K, threshold)
    with open(pair_path, "wb") as pf:
        df = open(dist_path, "wb") if dist_path else None
        try:
            for first in range(0, count, block):
                p, d = ctx.kmer_fetch(first, min(block, count - first), dists=df is not None)
                p.tofile(pf)
                if df is not None:
                    d.tofile(df)
        finally:
            if df is not None:
                df.close()
    return count


def compute_gen_dist(ctx: Context, reads: Reads, in_pair_path: str, pair_prefix: str,
                     dist_prefix: str, threshold: float, band: int):
    """Reads <in_pair_path> in chunks, keeps pairs whose NW distance passes
    and writes <pair_prefix>_<f> (uint32 pairs) / <dist_prefix>_<f>
    (float32), each ascending by distance (ties by (i, j)).  Returns
    (number of files, number of pairs)."""
    ctx.upload_reads(NW, reads)
    acc_p, acc_d = [], []
    nacc = 0
    nfiles = 0
    total = 0

    def flush():
        nonlocal nfiles, nacc, acc_p, acc_d
        p = np.concatenate(acc_p) if acc_p else np.zeros((0, 2), np.int32)
        d = np.concatenate(acc_d) if acc_d else np.zeros(0, np.float32)
        p, d = ctx.sort_by_dist(p, d)
        p.astype(np.uint32).tofile(f"{pair_prefix}_{nfiles}")
        d.tofile(f"{dist_prefix}_{nfiles}")
        nfiles += 1
        acc_p, acc_d, nacc = [], [], 0

    with open(in_pair_path, "rb") as f:
        while True:
            raw = np.fromfile(f, dtype=np.int32, count=2 * CHUNK_PAIRS)
            eof = len(raw) < 2 * CHUNK_PAIRS
            pairs = raw[: len(raw) // 2 * 2].reshape(-1, 2)
            c = ctx.nw_run(pairs, band, threshold)
            if c:
                p, d = ctx.nw_fetch()
                acc_p.append(p)
                acc_d.append(d)
                nacc += c
            if nacc >= FLUSH_PAIRS:
                total += nacc
                flush()
            if eof:
                break
    if nacc > 0:
        total += nacc
        flush()
    return nfiles, total


class OutputBuffers:
    """Grow-only page-locked host arrays that fused_output_files can fill
    instead of allocating new ones per call (fresh pageable pages cost a page
    fault each on the first copy, and a copy into pageable memory goes through
    a bounce buffer); the files it returns are then views into them, valid
    until the next call with the same buffers."""

    def __init__(self):
        self.pairs = np.empty((0, 2), np.int32)
        self.dists = np.empty(0, np.float32)

    def reserve(self, n):
        if len(self.dists) < n:
            from .api import pinned_empty
            self.pairs = pinned_empty((n, 2), np.int32)
            self.dists = pinned_empty(n, np.float32)


def fused_output_files(ctx: Context, buffers: OutputBuffers = None):
    """The output files of the current NW result (crispy_nw_run_on_kmer): the
    reference's file split of its input (the k-mer pairs read in
    CHUNK_PAIRS chunks, a file closed at >= FLUSH_PAIRS, genCPU.cpp:482-505)
    and each file's pairs in (dist, i, j) order, sorted on the device.
    Returns a list of (pairs (n, 2) int32, dists float32); with `buffers`,
    views into them."""
    out = []
    first = 0
    if buffers is not None:
        buffers.reserve(ctx.nw_count)
    for end in ctx.nw_file_split(CHUNK_PAIRS, FLUSH_PAIRS):
        end = int(end)
        if buffers is None:
            out.append(ctx.nw_fetch_sorted(first, end - first))
        else:
            out.append(ctx.nw_fetch_sorted(first, end - first, buffers.pairs[first:end], buffers.dists[first:end]))
        first = end
    return out


def preprocess(ctx: Context, in_path: str, out_path: str = None, fastq: bool = False,
               sd_allowed: float = None, primer: bytes = None, max_mismatches: int = 2):
    """The reference preprocess main (preprocess.cpp:686-726): read, optional
    -s length filter, optional -p primer filter (primer as bytes, see
    api.read_primer), merge exact duplicates on the device, write <out>
    (default <in>_Clean) and <out>.frq.  Returns the number of reads
    written."""
    from .api import preprocess_length_filter, preprocess_read, preprocess_write
    rec = preprocess_read(in_path, fastq)
    if sd_allowed is not None:
        preprocess_length_filter(rec, sd_allowed)
    if primer is not None:
        ctx.preprocess_primer_filter(rec, primer, max_mismatches)
    order, freq = ctx.preprocess_order(rec.reads, rec.freq)
    preprocess_write(out_path or (in_path + "_Clean"), rec, order, freq, fastq)
    return len(order)
