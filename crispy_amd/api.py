"""ctypes binding of include/crispy_hip.h."""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(PKG, "lib", "libcrispy_hip.so")

KMER = 0  # CRISPY_READS_KMER
NW = 1    # CRISPY_READS_NW
MAX_READ_LEN = 2048
MAX_NUM_READS = 1024 * 1024

_ERRNAMES = {-1: "ARG", -2: "HIP", -3: "NOMEM", -4: "STATE", -5: "IO"}

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

_lib = None


class CrispyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"crispy error {code} ({_ERRNAMES.get(code, '?')}): {msg}")
        self.code = code


def library_path():
    return _LIB_PATH


def load_library():
    """Loads libcrispy_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"{_LIB_PATH} is missing: build it with `python -m crispy_amd.build`")
    L = ctypes.CDLL(_LIB_PATH)
    sig = {
        "crispy_ctx_create": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
        "crispy_ctx_destroy": ([_vp], None),
        "crispy_ctx_error": ([_vp], ctypes.c_char_p),
        "crispy_ctx_sync": ([_vp], ctypes.c_int),
        "crispy_ctx_device_bytes": ([_vp], ctypes.c_int64),
        "crispy_ctx_trim": ([_vp], ctypes.c_int),
        "crispy_read_fasta": ([ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_vp),
                               ctypes.POINTER(_vp), _i32p], ctypes.c_int32),
        "crispy_free": ([_vp], None),
        "crispy_kmer_default_threshold": ([ctypes.c_int32], ctypes.c_float),
        "crispy_nw_default_threshold": ([ctypes.c_int32], ctypes.c_float),
        "crispy_pass": ([ctypes.c_float, ctypes.c_float], ctypes.c_int),
        "crispy_band_cells": ([ctypes.c_int32, ctypes.c_int32, ctypes.c_int32], ctypes.c_int64),
        "crispy_reads_upload": ([_vp, ctypes.c_int, ctypes.c_char_p, _i64p, ctypes.c_int32],
                                ctypes.c_int),
        "crispy_kmer_run": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_int64, ctypes.c_int64,
                             _i64p], ctypes.c_int),
        "crispy_kmer_num_tiles": ([ctypes.c_int32], ctypes.c_int64),
        "crispy_kmer_fetch": ([_vp, ctypes.c_int64, ctypes.c_int64, _i32p, _f32p], ctypes.c_int),
        "crispy_nw_run": ([_vp, _i32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, _i64p],
                          ctypes.c_int),
        "crispy_nw_run_on_kmer": ([_vp, ctypes.c_int, ctypes.c_float, _i64p], ctypes.c_int),
        "crispy_nw_fetch": ([_vp, ctypes.c_int64, ctypes.c_int64, _i32p, _f32p], ctypes.c_int),
        "crispy_nw_dists": ([_vp, _i32p, ctypes.c_int64, ctypes.c_int, _f32p], ctypes.c_int),
        "crispy_nw_csr": ([_vp, ctypes.c_int32, _i64p, _i32p, _f32p], ctypes.c_int),
        "crispy_nw_set_slice_width": ([_vp, ctypes.c_int], ctypes.c_int),
        "crispy_set_graphs": ([_vp, ctypes.c_int], ctypes.c_int),
        "crispy_kmer_set_exact_pass": ([_vp, ctypes.c_int], ctypes.c_int),
        "crispy_kmer_set_hash_buckets": ([_vp, ctypes.c_int], ctypes.c_int),
        "crispy_host_alloc": ([ctypes.c_int64, ctypes.POINTER(_vp)], ctypes.c_int),
        "crispy_host_free": ([_vp], ctypes.c_int),
        "crispy_kmer_tile_bits": ([_vp], ctypes.c_int),
        "crispy_nw_autotune": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "crispy_kmer_autotune": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                                 ctypes.c_int),
        "crispy_sort_by_dist": ([_vp, _i32p, _f32p, ctypes.c_int64], ctypes.c_int),
        "crispy_nw_dists_noalloc": ([_vp, _vp, ctypes.c_int64, ctypes.c_int, _vp], ctypes.c_int),
        "crispy_nw_file_split": ([_vp, ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "crispy_nw_fetch_sorted": ([_vp, ctypes.c_int64, ctypes.c_int64, _i32p, _f32p], ctypes.c_int),
        "crispy_timing_enable": ([_vp, ctypes.c_int], ctypes.c_int),
        "crispy_timing_reset": ([_vp], ctypes.c_int),
        "crispy_timing_query": ([_vp, ctypes.c_char_p, _f64p, _i64p, _f64p], ctypes.c_int),
        "crispy_selftest_division": ([_vp, ctypes.c_int32, _f32p, _f32p], ctypes.c_int),
        "crispy_preprocess_read": ([ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp)], ctypes.c_int32),
        "crispy_preprocess_primer_filter": ([_vp, ctypes.c_char_p, _i64p, ctypes.c_int32, ctypes.c_char_p,
                                             ctypes.c_int32, ctypes.c_int32, _i32p], ctypes.c_int32),
        "crispy_preprocess_length_filter": ([_i64p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_double,
                                             _i32p], ctypes.c_int32),
        "crispy_preprocess_order": ([_vp, ctypes.c_char_p, _i64p, _i32p, ctypes.c_int32, _i32p, _i32p,
                                     _i32p], ctypes.c_int),
        "crispy_preprocess_write": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, _i64p,
                                     ctypes.c_char_p, _i64p, _i32p, _i32p, ctypes.c_int32],
                                    ctypes.c_int),
        "crispy_server_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(_vp)],
                                 ctypes.c_int),
        "crispy_server_destroy": ([_vp], None),
        "crispy_server_job_bytes": ([ctypes.c_char_p, _i64p, ctypes.c_int32, ctypes.c_int,
                                     ctypes.c_int], ctypes.c_int64),
        "crispy_server_submit": ([_vp, ctypes.c_char_p, _i64p, ctypes.c_char_p, _i64p,
                                  ctypes.c_int32, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_float], ctypes.c_int64),
        "crispy_server_wait": ([_vp, ctypes.c_int64, _i64p, _i64p, _f64p, _f64p], ctypes.c_int),
        "crispy_server_job_times": ([_vp, ctypes.c_int64, _f64p, _f64p, _f64p, _i64p, _i64p],
                                    ctypes.c_int),
        "crispy_server_job_error": ([_vp, ctypes.c_int64], ctypes.c_char_p),
        "crispy_server_fetch": ([_vp, ctypes.c_int64, _i32p, _f32p], ctypes.c_int),
        "crispy_server_fetch_csr": ([_vp, ctypes.c_int64, ctypes.c_int32, _i64p, _i32p, _f32p], ctypes.c_int),
        "crispy_server_status": ([_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  _i64p, _i64p], ctypes.c_int),
        "crispy_server_job_solo": ([_vp, ctypes.c_int64], ctypes.c_int),
        "crispy_ctx_set_mem_limit": ([_vp, ctypes.c_int64], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


class Reads:
    """Reads as concatenated bytes plus int64 offsets (n + 1)."""

    def __init__(self, bases: bytes, offsets, maxlen=None):
        self.bases = bytes(bases)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        lens = np.diff(self.offsets)
        self.maxlen = int(lens.max()) if maxlen is None and len(lens) else (maxlen or 0)

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
        return cls(b"".join(seqs), offs)


def read_fasta(path, variant=KMER):
    """The reference's FASTA reader (kmerMain.cpp:298-397 / genMain.cpp:184-277)."""
    L = load_library()
    b, o, ml = _vp(), _vp(), ctypes.c_int32()
    n = L.crispy_read_fasta(os.fsencode(path), variant, ctypes.byref(b), ctypes.byref(o),
                            ctypes.byref(ml))
    if n < 0:
        raise CrispyError(n, f"cannot read FASTA {path}")
    try:
        offs = np.ctypeslib.as_array(ctypes.cast(o, _i64p), shape=(n + 1,)).copy()
        bases = ctypes.string_at(b, int(offs[-1])) if offs[-1] > 0 else b""
    finally:
        L.crispy_free(b)
        L.crispy_free(o)
    return Reads(bases, offs, ml.value)


class Records:
    """Records of the preprocess reader: sequences (Reads), labels (bytes +
    offsets, whitespace removed) and a frequency per record."""

    def __init__(self, reads, labels, label_offsets, freq):
        self.reads = reads
        self.labels = bytes(labels)
        self.label_offsets = np.ascontiguousarray(label_offsets, dtype=np.int64)
        self.freq = np.ascontiguousarray(freq, dtype=np.int32)

    @property
    def n(self):
        return self.reads.n

    def label(self, i):
        return self.labels[self.label_offsets[i]:self.label_offsets[i + 1]]


def preprocess_read(path, fastq=False):
    """The reference preprocess reader (preprocess.cpp:74-240)."""
    L = load_library()
    b, o, lb, lo, fr = _vp(), _vp(), _vp(), _vp(), _vp()
    n = L.crispy_preprocess_read(os.fsencode(path), int(bool(fastq)), ctypes.byref(b), ctypes.byref(o),
                                 ctypes.byref(lb), ctypes.byref(lo), ctypes.byref(fr))
    if n < 0:
        raise CrispyError(n, f"cannot read {path}")
    try:
        offs = np.ctypeslib.as_array(ctypes.cast(o, _i64p), shape=(n + 1,)).copy()
        loffs = np.ctypeslib.as_array(ctypes.cast(lo, _i64p), shape=(n + 1,)).copy()
        freq = np.ctypeslib.as_array(ctypes.cast(fr, _i32p), shape=(max(n, 1),)).copy()[:n]
        bases = ctypes.string_at(b, int(offs[-1])) if offs[-1] > 0 else b""
        labels = ctypes.string_at(lb, int(loffs[-1])) if loffs[-1] > 0 else b""
    finally:
        for p in (b, o, lb, lo, fr):
            L.crispy_free(p)
    return Records(Reads(bases, offs), labels, loffs, freq)


def preprocess_length_filter(rec: Records, sd_allowed):
    """-s filter (preprocess.cpp:322-392): zeroes the frequency of outliers and
    reads with N in place; returns the number discarded."""
    return int(load_library().crispy_preprocess_length_filter(
        _p(rec.reads.offsets, _i64p), rec.reads.bases, rec.n, float(sd_allowed), _p(rec.freq, _i32p)))


def read_primer(path):
    """The primer as the reference reads it (preprocess.cpp:672-683): the
    file's first line with trailing whitespace removed (a line of only
    whitespace stays as it is)."""
    with open(path, "rb") as f:
        line = f.readline().split(b"\n")[0]
    t = line.rstrip(b"\n\r \t\v\f\b")
    return t if t else line


def preprocess_write(path, rec: Records, order, freq, fastq=False):
    """writeToFile (preprocess.cpp:562-601): <path> and <path>.frq."""
    order = np.ascontiguousarray(order, dtype=np.int32)
    freq = np.ascontiguousarray(freq, dtype=np.int32)
    rc = load_library().crispy_preprocess_write(
        os.fsencode(path), int(bool(fastq)), rec.reads.bases, _p(rec.reads.offsets, _i64p), rec.labels,
        _p(rec.label_offsets, _i64p), _p(order, _i32p), _p(freq, _i32p), len(order))
    if rc != 0:
        raise CrispyError(rc, f"cannot write {path}")


class _PinnedBlock:
    """Owns one crispy_host_alloc block; frees it when the last array viewing
    it is gone."""

    def __init__(self, nbytes):
        self._L = load_library()
        p = _vp()
        rc = self._L.crispy_host_alloc(nbytes, ctypes.byref(p))
        if rc != 0:
            raise CrispyError(rc, f"cannot allocate {nbytes} bytes of page-locked host memory")
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            self._L.crispy_host_free(self.ptr)
            self.ptr = None


def pinned_empty(shape, dtype):
    """An uninitialised numpy array in page-locked host memory (copies from
    the device into it run at the bus's rate); freed with its last view."""
    dtype = np.dtype(dtype)
    count = int(np.prod(shape)) if np.ndim(shape) else int(shape)
    nbytes = max(1, count * dtype.itemsize)
    block = _PinnedBlock(nbytes)
    raw = (ctypes.c_byte * nbytes).from_address(block.ptr)
    raw._owner = block  # the ctypes array, and through it every numpy view, keeps the block alive
    return np.ctypeslib.as_array(raw)[:count * dtype.itemsize].view(dtype).reshape(shape)


def kmer_default_threshold(n):
    return float(load_library().crispy_kmer_default_threshold(n))


def nw_default_threshold(n):
    return float(load_library().crispy_nw_default_threshold(n))


def passes(dist, thr):
    return bool(load_library().crispy_pass(dist, thr))


def band_cells(len1, len2, band):
    return int(load_library().crispy_band_cells(len1, len2, band))


class Context:
    """One device context (stream + resident buffers)."""

    def __init__(self, device=0):
        self._L = load_library()
        h = _vp()
        rc = self._L.crispy_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise CrispyError(rc, f"cannot open HIP device {device}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.crispy_ctx_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise CrispyError(rc, self._L.crispy_ctx_error(self._h).decode(errors="replace"))

    # -- reads
    def upload_reads(self, slot, reads: Reads):
        self._check(self._L.crispy_reads_upload(self._h, slot, reads.bases,
                                                _p(reads.offsets, _i64p), reads.n))

    # -- kmerDist
    def kmer_run(self, K, threshold, tiles=None):
        t0, t1 = (0, -1) if tiles is None else tiles
        c = ctypes.c_int64()
        self._check(self._L.crispy_kmer_run(self._h, K, threshold, t0, t1, ctypes.byref(c)))
        self.kmer_count = c.value
        return c.value

    def kmer_fetch(self, first=0, count=None, dists=True):
        if count is None:
            count = self.kmer_count - first
        p = np.zeros((count, 2), np.int32)
        d = np.zeros(count, np.float32) if dists else None
        self._check(self._L.crispy_kmer_fetch(self._h, first, count, _p(p, _i32p),
                                              _p(d, _f32p) if dists else None))
        return p, d

    # -- genDist
    def nw_run(self, pairs, band, threshold):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        c = ctypes.c_int64()
        self._check(self._L.crispy_nw_run(self._h, _p(pairs, _i32p), len(pairs), band, threshold,
                                          ctypes.byref(c)))
        self.nw_count = c.value
        return c.value

    def nw_run_on_kmer(self, band, threshold):
        c = ctypes.c_int64()
        self._check(self._L.crispy_nw_run_on_kmer(self._h, band, threshold, ctypes.byref(c)))
        self.nw_count = c.value
        return c.value

    def nw_fetch(self, first=0, count=None):
        if count is None:
            count = self.nw_count - first
        p = np.zeros((count, 2), np.int32)
        d = np.zeros(count, np.float32)
        self._check(self._L.crispy_nw_fetch(self._h, first, count, _p(p, _i32p), _p(d, _f32p)))
        return p, d

    def nw_slice_width(self, cells):
        """Band cells per lane of the lane-sliced NW kernel: 0 = cost model (default), 24 or 32."""
        self._check(self._L.crispy_nw_set_slice_width(self._h, cells))

    def kmer_tile_bits(self):
        """Presence bits per read the last k-mer tile pass multiplied (4^min(K,6)
        padded to >= 512, or the bucket count of the hashed k = 6 candidate
        filter)."""
        return int(self._L.crispy_kmer_tile_bits(self._h))

    def kmer_exact_pass(self, mode):
        """K <= 6: exact counts of the pairs the tile's bound keeps in the tile
        kernel (0 or False), in a second kernel over the tile's candidates
        (1 or True; for K = 6 the candidate tile runs on the hashed bitmaps),
        the same without hashing (2), or chosen by the pair count (-1, the
        default); K = 7, 8 take candidates, hashed unless the mode is 2.  The
        results are the same."""
        self._check(self._L.crispy_kmer_set_exact_pass(self._h, int(mode)))

    def kmer_hash_buckets(self, buckets):
        """Buckets of the hashed k = 6 candidate filter: 1280, 1536, 2048 or 0 (the
        default, 1280).  The results are the same."""
        self._check(self._L.crispy_kmer_set_hash_buckets(self._h, int(buckets)))

    def kmer_autotune(self, K, threshold, repeats=2):
        """Times the k-mer run under each exact-count form that applies to K
        (and each hash bucket count for the hashed form), keeps the fastest
        (and its result); returns the chosen mode (kmer_tile_bits() tells the
        bucket count)."""
        c = ctypes.c_int()
        self._check(self._L.crispy_kmer_autotune(self._h, K, threshold, repeats, ctypes.byref(c)))
        self.kmer_run(K, threshold)  # the chosen form's result and its count
        return c.value

    def graphs(self, enable):
        """Kernel chains (k-mer profile and tile passes, NW DP phase, NW
        filter) replayed from captured HIP graphs (True, the default) or
        launched one by one (False); the results are the same."""
        self._check(self._L.crispy_set_graphs(self._h, 1 if enable else 0))

    def nw_autotune(self, band, repeats=3):
        """Times the NW DP on the current k-mer result under the slice-width
        cost model and each forced width, keeps the fastest; returns it."""
        c = ctypes.c_int()
        self._check(self._L.crispy_nw_autotune(self._h, band, repeats, ctypes.byref(c)))
        return c.value

    def nw_csr(self, n_rows):
        """(row_ptr, cols, vals) of the last NW result, built on the device."""
        rp = np.zeros(n_rows + 1, np.int64)
        c = self.nw_count
        cols = np.zeros(max(c, 1), np.int32)
        vals = np.zeros(max(c, 1), np.float32)
        self._check(self._L.crispy_nw_csr(self._h, n_rows, _p(rp, _i64p), _p(cols, _i32p), _p(vals, _f32p)))
        return rp, cols[:c], vals[:c]

    def nw_dists(self, pairs, band):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        d = np.zeros(len(pairs), np.float32)
        self._check(self._L.crispy_nw_dists(self._h, _p(pairs, _i32p), len(pairs), band,
                                            _p(d, _f32p)))
        return d

    def nw_dists_noalloc(self, d_pairs, npairs, band, d_dists):
        """crispy_nw_dists_noalloc: distances of `npairs` device pairs (int32
        x2 at address d_pairs) into the float32 device array at d_dists, with
        no device allocation (the library's fixed per-device scratch)."""
        self._check(self._L.crispy_nw_dists_noalloc(self._h, ctypes.c_void_p(d_pairs), int(npairs), band,
                                                    ctypes.c_void_p(d_dists)))

    def sort_by_dist(self, pairs, dists):
        pairs = np.array(pairs, dtype=np.int32).reshape(-1, 2)
        dists = np.array(dists, dtype=np.float32)
        self._check(self._L.crispy_sort_by_dist(self._h, _p(pairs, _i32p), _p(dists, _f32p),
                                                len(dists)))
        return pairs, dists

    def nw_file_split(self, chunk_pairs=576 * 256 * 4, flush_pairs=576 * 256, max_files=4096):
        """Result indices where the reference's output files end (its input
        read in `chunk_pairs` chunks, a file closed at >= `flush_pairs`)."""
        ends = np.zeros(max_files, np.int64)
        nf = ctypes.c_int32()
        self._check(self._L.crispy_nw_file_split(self._h, chunk_pairs, flush_pairs, _p(ends, _i64p),
                                                 max_files, ctypes.byref(nf)))
        return ends[:nf.value]

    def nw_fetch_sorted(self, first=0, count=None, pairs=None, dists=None):
        """Result entries [first, first + count) ordered by (dist, i, j), sorted
        on the device: one output file's contents.  `pairs` ((count, 2) int32)
        and `dists` ((count,) float32), C-contiguous, receive them when given
        (so a caller can reuse its host buffers); returns the two arrays."""
        if count is None:
            count = self.nw_count - first
        p = np.empty((count, 2), np.int32) if pairs is None else pairs
        d = np.empty(count, np.float32) if dists is None else dists
        if p.shape != (count, 2) or p.dtype != np.int32 or not p.flags.c_contiguous or \
                d.shape != (count,) or d.dtype != np.float32 or not d.flags.c_contiguous:
            raise ValueError("pairs must be C-contiguous (count, 2) int32 and dists (count,) float32")
        self._check(self._L.crispy_nw_fetch_sorted(self._h, first, count, _p(p, _i32p), _p(d, _f32p)))
        return p, d

    # -- input side
    def preprocess_primer_filter(self, rec, primer: bytes, max_mismatches=2):
        """-p filter (preprocess.cpp:394-560) on the device: zeroes the
        frequency of reads without the primer at their start, in place in
        rec.freq; returns the number discarded."""
        r = self._L.crispy_preprocess_primer_filter(self._h, rec.reads.bases, _p(rec.reads.offsets, _i64p), rec.n,
                                                    primer, len(primer), int(max_mismatches), _p(rec.freq, _i32p))
        if r < 0:
            self._check(r)
        return int(r)

    def preprocess_order(self, reads: Reads, freq):
        """mergeReads on the device: (order, frequencies) of the output reads."""
        freq = np.ascontiguousarray(freq, dtype=np.int32)
        n = reads.n
        order = np.zeros(max(n, 1), np.int32)
        ofreq = np.zeros(max(n, 1), np.int32)
        c = ctypes.c_int32()
        self._check(self._L.crispy_preprocess_order(self._h, reads.bases, _p(reads.offsets, _i64p),
                                                    _p(freq, _i32p), n, _p(order, _i32p),
                                                    _p(ofreq, _i32p), ctypes.byref(c)))
        return order[:c.value].copy(), ofreq[:c.value].copy()

    # -- timing
    def timing(self, enable=True):
        self._check(self._L.crispy_timing_enable(self._h, int(enable)))

    def timing_reset(self):
        self._check(self._L.crispy_timing_reset(self._h))

    def timing_query(self, kernel):
        ms, n, u = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        self._check(self._L.crispy_timing_query(self._h, kernel.encode(), ctypes.byref(ms),
                                                ctypes.byref(n), ctypes.byref(u)))
        return ms.value, n.value, u.value

    def sync(self):
        self._check(self._L.crispy_ctx_sync(self._h))

    def mem_limit(self, nbytes):
        """Caps the device bytes the context's buffers may grow to (0: no cap)."""
        self._check(self._L.crispy_ctx_set_mem_limit(self._h, int(nbytes)))

    def device_bytes(self):
        return int(self._L.crispy_ctx_device_bytes(self._h))

    def trim(self):
        self._check(self._L.crispy_ctx_trim(self._h))

    def selftest_division(self, maxv):
        a = np.zeros((maxv + 1) * (maxv + 1), np.float32)
        b = np.zeros_like(a)
        self._check(self._L.crispy_selftest_division(self._h, maxv, _p(a, _f32p), _p(b, _f32p)))
        return a.reshape(maxv + 1, maxv + 1), b.reshape(maxv + 1, maxv + 1)


def job_bytes(reads: Reads, K, band):
    """Device bytes the server charges a job at admission."""
    return int(load_library().crispy_server_job_bytes(reads.bases, _p(reads.offsets, _i64p),
                                                      reads.n, K, band))


class Server:
    """Job server: `slots` contexts behind a FIFO queue with HBM-budget
    admission; one job = k-mer filter + banded NW on the passing pairs."""

    def __init__(self, device=0, slots=4, hbm_budget=64 << 30):
        self._L = load_library()
        h = _vp()
        rc = self._L.crispy_server_create(device, slots, int(hbm_budget), ctypes.byref(h))
        if rc != 0:
            raise CrispyError(rc, f"cannot start a server on HIP device {device}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.crispy_server_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, reads: Reads, K, kmer_threshold, band, nw_threshold, nw_reads=None):
        nb = nw_reads.bases if nw_reads is not None else None
        no = _p(nw_reads.offsets, _i64p) if nw_reads is not None else None
        jid = self._L.crispy_server_submit(self._h, reads.bases, _p(reads.offsets, _i64p), nb, no,
                                           reads.n, K, kmer_threshold, band, nw_threshold)
        if jid < 0:
            raise CrispyError(int(jid), "job rejected")
        return int(jid)

    def _job_check(self, jid, rc):
        if rc != 0:
            msg = self._L.crispy_server_job_error(self._h, jid)
            raise CrispyError(rc, (msg or b"").decode(errors="replace") or f"job {jid} failed")

    def wait(self, jid):
        """Blocks until job `jid` is done; returns a dict of its counts and times (ms)."""
        kp, npairs = ctypes.c_int64(), ctypes.c_int64()
        q, r = ctypes.c_double(), ctypes.c_double()
        rc = self._L.crispy_server_wait(self._h, jid, ctypes.byref(kp), ctypes.byref(npairs),
                                        ctypes.byref(q), ctypes.byref(r))
        self._job_check(jid, rc)
        s, b, e = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        est, held = ctypes.c_int64(), ctypes.c_int64()
        self._job_check(jid, self._L.crispy_server_job_times(
            self._h, jid, ctypes.byref(s), ctypes.byref(b), ctypes.byref(e), ctypes.byref(est),
            ctypes.byref(held)))
        solo = self._L.crispy_server_job_solo(self._h, jid)
        return dict(kmer_pairs=kp.value, nw_pairs=npairs.value, queue_ms=q.value, run_ms=r.value,
                    submit_ms=s.value, start_ms=b.value, end_ms=e.value, est_bytes=est.value,
                    held_bytes=held.value, ran_alone_again=solo == 1)

    def fetch(self, jid):
        """Waits for job `jid`, returns (pairs (n, 2) int32, dists float32) and forgets it."""
        info = self.wait(jid)
        n = info["nw_pairs"]
        p = np.zeros((n, 2), np.int32)
        d = np.zeros(n, np.float32)
        self._job_check(jid, self._L.crispy_server_fetch(self._h, jid, _p(p, _i32p), _p(d, _f32p)))
        return p, d

    def fetch_csr(self, jid, n_rows):
        """Waits for job `jid`; returns its result as (row_ptr, cols, vals) and forgets it."""
        info = self.wait(jid)
        c = info["nw_pairs"]
        rp = np.zeros(n_rows + 1, np.int64)
        cols = np.zeros(max(c, 1), np.int32)
        vals = np.zeros(max(c, 1), np.float32)
        self._job_check(jid, self._L.crispy_server_fetch_csr(self._h, jid, int(n_rows), _p(rp, _i64p),
                                                             _p(cols, _i32p), _p(vals, _f32p)))
        return rp, cols[:c], vals[:c]

    def status(self):
        r, q, b, pk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        rc = self._L.crispy_server_status(self._h, ctypes.byref(r), ctypes.byref(q), ctypes.byref(b),
                                          ctypes.byref(pk))
        if rc != 0:
            raise CrispyError(rc, "status")
        return dict(running=r.value, queued=q.value, reserved_bytes=b.value,
                    peak_reserved_bytes=pk.value)
