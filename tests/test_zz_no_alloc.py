"""crispy_nw_dists_noalloc (the entry point that allocates nothing) against
crispy_nw_dists on the same pairs, bit for bit, on two calls in a row (the
second smaller, so the scratch holds the first call's contents), plus its
refusals.  Device buffers for the inputs and outputs come from the HIP
runtime directly (hipMalloc through ctypes), as a caller's would."""
import ctypes
import os
import sys

import numpy as np
import pytest

import crispy_amd as ca
from crispy_amd import synth

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DevArray:
    """A device copy of a numpy array (hipMalloc'd, freed on close)."""
    _hip = None

    def __init__(self, nbytes):
        if DevArray._hip is None:
            rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
            DevArray._hip = ctypes.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
        self.ptr = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 4)
        assert DevArray._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)) == 0

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        assert cls._hip.hipMemcpy(d.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), H2D) == 0
        return d

    def to_host(self, dtype, count):
        out = np.zeros(count, dtype)
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes),
                                   D2H) == 0
        return out

    def close(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = None


def test_noalloc_matches_nw_dists():
    seqs, _, _ = synth.make("reads1k", n=1500)
    r = ca.Reads.from_list(seqs)
    rng = np.random.default_rng(11)
    n = r.n
    with ca.Context(0) as ctx:
        ctx.upload_reads(ca.NW, r)
        for band, npairs in ((10, 300000), (5, 70001), (3, 1000)):
            i = rng.integers(0, n, npairs)
            j = rng.integers(0, n, npairs)
            pairs = np.stack([i, j], 1).astype(np.int32)
            say("noalloc band", band, "pairs", npairs)
            want = ctx.nw_dists(pairs, band)
            dp, dd = DevArray.of(pairs), DevArray(4 * npairs)
            try:
                ctx.nw_dists_noalloc(dp.ptr.value, npairs, band, dd.ptr.value)
                got = dd.to_host(np.float32, npairs)
            finally:
                dp.close()
                dd.close()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (band, npairs)
        # refusals: too many pairs, a read id out of range, full alignment
        bad = np.array([[0, 1], [2, n]], np.int32)
        dp, dd = DevArray.of(bad), DevArray(8)
        try:
            with pytest.raises(ca.CrispyError):
                ctx.nw_dists_noalloc(dp.ptr.value, (1 << 20) + 1, 10, dd.ptr.value)
            with pytest.raises(ca.CrispyError):
                ctx.nw_dists_noalloc(dp.ptr.value, 2, 10, dd.ptr.value)
            with pytest.raises(ca.CrispyError):
                ctx.nw_dists_noalloc(dp.ptr.value, 1, 1, dd.ptr.value)
        finally:
            dp.close()
            dd.close()
