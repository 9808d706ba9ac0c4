"""crispy_amd: MI355X (gfx950) kmerDist + genDist behind the reference's
interface.

The compute path is libcrispy_hip.so (hand-written HIP kernels behind the C
ABI in include/crispy_hip.h), loaded with ctypes.  There is no CPU fallback:
if the library is missing or no GPU is present, calls raise.

    import crispy_amd as ca
    reads = ca.read_fasta("in_Clean", ca.KMER)
    with ca.Context() as ctx:
        ctx.upload_reads(ca.KMER, reads)
        n = ctx.kmer_run(K=6, threshold=ca.kmer_default_threshold(reads.n))
        pairs, dists = ctx.kmer_fetch()
"""
from .api import (  # noqa: F401
    KMER, NW, MAX_NUM_READS, MAX_READ_LEN, CrispyError, Context, Reads, Records, Server, band_cells,
    job_bytes, kmer_default_threshold, library_path, load_library, nw_default_threshold, passes, pinned_empty,
    preprocess_length_filter, preprocess_read, preprocess_write, read_fasta, read_primer,
)
from .pipeline import compute_gen_dist, compute_kmer_dist, preprocess  # noqa: F401
