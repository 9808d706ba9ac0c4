"""ctypes wrapper around oracle/liboracle.so (the CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package crispy_amd never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.POINTER(ctypes.c_float)
_u16p = ctypes.POINTER(ctypes.c_uint16)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_read_fasta.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_void_p), _i32p]
        L.oracle_read_fasta.restype = ctypes.c_int
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_kmer_threshold.argtypes = [ctypes.c_int]
        L.oracle_kmer_threshold.restype = ctypes.c_float
        L.oracle_nw_threshold.argtypes = [ctypes.c_int]
        L.oracle_nw_threshold.restype = ctypes.c_float
        L.oracle_kmer_array_dim.argtypes = [ctypes.c_int]
        L.oracle_pass.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_kmer_tuples.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int, _u16p]
        L.oracle_kmer_pair.argtypes = [_u16p, ctypes.c_int, _u16p, ctypes.c_int, _i32p]
        L.oracle_kmer_pair.restype = ctypes.c_float
        L.oracle_kmer_all.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int32, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_int, ctypes.c_int32,
                                      ctypes.c_int32, _i32p, _f32p, ctypes.c_int64]
        L.oracle_kmer_all.restype = ctypes.c_int64
        L.oracle_kmer_rowlist.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int32, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_int, _i32p, ctypes.c_int32,
                                          _i32p, _f32p, ctypes.c_int64]
        L.oracle_kmer_rowlist.restype = ctypes.c_int64
        L.oracle_kmer_pairs.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int32, ctypes.c_int,
                                        _i32p, ctypes.c_int64, _f32p, _i32p]
        L.oracle_nw_pair.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                     ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p]
        L.oracle_nw_pair.restype = ctypes.c_float
        L.oracle_nw_pairs.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int32, _i32p,
                                      ctypes.c_int64, ctypes.c_int, ctypes.c_int, _f32p,
                                      _i32p, _i32p]
        L.oracle_nw_pairs_as_shipped.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int32, _i32p,
                                                 ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, _f32p]
        L.oracle_gendist_files.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int32, _i32p,
                                           ctypes.c_int64, ctypes.c_int, ctypes.c_float,
                                           ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                           _i64p]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_preprocess_file.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_char_p, ctypes.c_int]
        L.oracle_preprocess_file.restype = ctypes.c_int
        L.oracle_primer_keep.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.c_int]
        L.oracle_primer_keep.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Reads:
    """Reads as concatenated bytes plus int64 offsets (n+1)."""

    def __init__(self, bases: bytes, offsets: np.ndarray, maxlen: int = 0):
        self.bases = bases
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.maxlen = maxlen

    @property
    def n(self):
        return len(self.offsets) - 1

    def seq(self, i):
        return self.bases[self.offsets[i]:self.offsets[i + 1]]

    def lengths(self):
        return np.diff(self.offsets).astype(np.int32)

    @classmethod
    def from_list(cls, seqs):
        seqs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
        offs = np.zeros(len(seqs) + 1, np.int64)
        offs[1:] = np.cumsum([len(s) for s in seqs])
        return cls(b"".join(seqs), offs, max([len(s) for s in seqs] or [0]))


def read_fasta(path, variant="kmer"):
    """kmerMain.cpp:298-397 (variant 'kmer') / genMain.cpp:184-277 ('gen')."""
    L = lib()
    b = ctypes.c_void_p()
    o = ctypes.c_void_p()
    ml = ctypes.c_int32()
    n = L.oracle_read_fasta(path.encode(), 0 if variant == "kmer" else 1,
                            ctypes.byref(b), ctypes.byref(o), ctypes.byref(ml))
    if n < 0:
        raise IOError(f"oracle_read_fasta({path}) failed: {n}")
    offs = np.ctypeslib.as_array(ctypes.cast(o, _i64p), shape=(n + 1,)).copy()
    bases = ctypes.string_at(b, int(offs[-1])) if offs[-1] > 0 else b""
    L.oracle_free(b)
    L.oracle_free(o)
    return Reads(bases, offs, ml.value)


def kmer_threshold(n):
    return lib().oracle_kmer_threshold(n)


def nw_threshold(n):
    return lib().oracle_nw_threshold(n)


def passes(dist, thr):
    return bool(lib().oracle_pass(dist, thr))


def kmer_tuples(seq: bytes, K: int) -> np.ndarray:
    out = np.zeros(max(len(seq) - K + 1, 1), np.uint16)
    T = lib().oracle_kmer_tuples(seq, len(seq), K, _p(out, _u16p))
    if T < 0:
        raise ValueError("read shorter than K")
    return out[:T]


def kmer_all(reads: Reads, K: int, thr: float, nthreads=1, rows=None, cap=None):
    """All pairs i<j passing the k-mer filter (kmerCPU.cpp:11-76)."""
    r0, r1 = (0, reads.n) if rows is None else rows
    if cap is None:
        cap = reads.n * (reads.n - 1) // 2 if rows is None else (r1 - r0) * reads.n
        cap = max(int(cap), 1)
    pairs = np.zeros((cap, 2), np.int32)
    dists = np.zeros(cap, np.float32)
    cnt = lib().oracle_kmer_all(reads.bases, _p(reads.offsets, _i64p), reads.n, K, thr,
                                nthreads, r0, r1, _p(pairs, _i32p), _p(dists, _f32p), cap)
    if cnt < 0:
        raise ValueError("read shorter than K")
    cnt = min(cnt, cap)
    return pairs[:cnt], dists[:cnt]


def kmer_rows(reads: Reads, K: int, thr: float, rows, nthreads=None, cap=None):
    """kmer_all restricted to the listed rows i (each against every j > i)."""
    rows = np.ascontiguousarray(np.unique(np.asarray(rows, np.int32)))
    if nthreads is None:
        nthreads = lib().oracle_num_threads()
    if cap is None:
        cap = min(int(sum(reads.n - 1 - int(i) for i in rows)), 1 << 22)
    while True:
        out = np.zeros((max(cap, 1), 2), np.int32)
        d = np.zeros(max(cap, 1), np.float32)
        c = lib().oracle_kmer_rowlist(reads.bases, _p(reads.offsets, _i64p), reads.n, K, thr, nthreads,
                                      _p(rows, _i32p), len(rows), _p(out, _i32p), _p(d, _f32p), cap)
        if c < 0:
            raise ValueError("read shorter than K")
        if c <= cap:
            return out[:c], d[:c]
        cap = int(c)


def kmer_pairs(reads: Reads, K: int, pairs: np.ndarray):
    pairs = np.ascontiguousarray(pairs, dtype=np.int32)
    d = np.zeros(len(pairs), np.float32)
    m = np.zeros(len(pairs), np.int32)
    rc = lib().oracle_kmer_pairs(reads.bases, _p(reads.offsets, _i64p), reads.n, K,
                                 _p(pairs, _i32p), len(pairs), _p(d, _f32p), _p(m, _i32p))
    if rc:
        raise ValueError("read shorter than K")
    return d, m


def nw_pair(s1: bytes, s2: bytes, band: int):
    ml, al, sc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    d = lib().oracle_nw_pair(s1, len(s1), s2, len(s2), band, ctypes.byref(ml),
                             ctypes.byref(al), ctypes.byref(sc))
    return d, ml.value, al.value, sc.value


def nw_pairs(reads: Reads, pairs: np.ndarray, band: int, nthreads=None):
    pairs = np.ascontiguousarray(pairs, dtype=np.int32)
    if nthreads is None:
        nthreads = lib().oracle_num_threads()
    d = np.zeros(len(pairs), np.float32)
    ml = np.zeros(len(pairs), np.int32)
    al = np.zeros(len(pairs), np.int32)
    rc = lib().oracle_nw_pairs(reads.bases, _p(reads.offsets, _i64p), reads.n,
                               _p(pairs, _i32p), len(pairs), band, nthreads,
                               _p(d, _f32p), _p(ml, _i32p), _p(al, _i32p))
    if rc:
        raise ValueError("pair index out of range")
    return d, ml, al


def nw_pairs_as_shipped(reads: Reads, pairs: np.ndarray, band: int, nthreads: int, ncpus: int):
    """NW distances under genDist's shipped CPU settings (nthreads threads
    confined to ncpus CPUs, static,1 schedule, a critical section per pair)."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int32)
    d = np.zeros(len(pairs), np.float32)
    rc = lib().oracle_nw_pairs_as_shipped(reads.bases, _p(reads.offsets, _i64p), reads.n,
                                          _p(pairs, _i32p), len(pairs), band, nthreads, ncpus,
                                          _p(d, _f32p))
    if rc:
        raise ValueError("pair index out of range or CPU set unavailable")
    return d


def gendist_files(reads: Reads, pairs: np.ndarray, band: int, thr: float, prefix: str,
                  nthreads=None):
    """genCPU.cpp:290-507 file policy; writes <prefix>.npair_f / .ndist_f."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int32)
    if nthreads is None:
        nthreads = lib().oracle_num_threads()
    tot = ctypes.c_int64()
    nf = lib().oracle_gendist_files(reads.bases, _p(reads.offsets, _i64p), reads.n,
                                    _p(pairs, _i32p), len(pairs), band, thr, nthreads,
                                    (prefix + ".npair").encode(), (prefix + ".ndist").encode(),
                                    ctypes.byref(tot))
    if nf < 0:
        raise IOError("oracle_gendist_files failed")
    return nf, tot.value


def preprocess_file(in_path, out_path, fastq=False, sd=None, primer=None, max_mismatches=2):
    """The reference preprocess path (preprocess.cpp:686-726): -s filter when
    sd is given, -p filter when primer (bytes, already read from its file) is
    given; writes out_path and out_path.frq, returns the number of reads
    written."""
    n = lib().oracle_preprocess_file(os.fsencode(in_path), os.fsencode(out_path), int(bool(fastq)),
                                     -1.0 if sd is None else float(sd), primer, int(max_mismatches))
    if n < 0:
        raise IOError(f"oracle_preprocess_file({in_path}) failed: {n}")
    return n


def primer_keep(primer, read, max_mismatches=2):
    """processPrimer's verdict for one read (preprocess.cpp:394-560): True = kept."""
    return bool(lib().oracle_primer_keep(primer, len(primer), read, len(read), int(max_mismatches)))


def num_threads():
    return lib().oracle_num_threads()


def ref_available():
    return all(os.access(os.path.join(REF_DIR, b), os.X_OK)
               for b in ("kmerDist", "preprocess", "aveclust", "sparsecut"))
