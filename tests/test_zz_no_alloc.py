"""bt_direct_fp64_noalloc, the entry point that allocates nothing (it has no
counterpart in the reference; it is for callers that cannot allocate device
memory), against bt_direct_fp64 through the C ABI.

Two calls in a row on one context, the second with fewer targets, for a
Treecode<float> and a Treecode<double>.  Bars: with as many source slices as
bt_direct_fp64 uses (ns <= 256 here) the rows are bit-equal; with fewer slices
(the 16 MiB scratch bounds ns * slices) only the FP64 summation order differs,
so rows agree to 1e-12 relative.  The device buffers of the test itself are
allocated with the HIP runtime through ctypes.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BT_ERR_ARG = -1


def _hip():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    h = ctypes.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _rows_close(got, ref, tol):
    da = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
    dp = np.abs(got[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])
    return float(max(da.max(), dp.max())), tol


@pytest.mark.parametrize("fp64", [False, True], ids=["float", "double"])
def test_noalloc_direct_matches_direct(bt, fp64):
    print("case noalloc fp64=%s" % fp64, flush=True)
    n = 20000
    pos, vel = bt.plummer(n, seed=4242)
    if fp64:
        pos, vel = pos.astype(np.float64), vel.astype(np.float64)
    t = bt.Treecode(0.01, 0.5, 64, fp64=fp64)
    t.alloc(n)
    t.ptcl_h2d(pos, vel)
    t.buildTree(16)
    t.computeMultipoles()
    t.makeGroups(5, 64)
    hip = _hip()
    d_idx = DevBuf(hip, 4 * 4096)
    d_out = DevBuf(hip, 32 * 4096)
    rng = np.random.default_rng(11)
    try:
        for ns in (3000, 200):       # the second call has fewer targets
            idx = np.sort(rng.choice(n, ns, replace=False)).astype(np.int32)
            ref = t.direct_fp64(idx)
            d_idx.put(idx)
            d_out.put(np.full((4096, 4), 7.0))   # stale values the call must overwrite
            assert hip.hipDeviceSynchronize() == 0
            assert t.direct_fp64_noalloc(ns, d_idx.ptr.value, d_out.ptr.value) == 0
            assert hip.hipDeviceSynchronize() == 0
            got = d_out.get((ns, 4), np.float64)
            assert np.all(np.isfinite(got))
            if ns <= 256:
                assert np.array_equal(got, ref), "same slices: rows must be bit-equal"
            else:
                err, tol = _rows_close(got, ref, 1e-12)
                assert err <= tol, "ns=%d: max row rel diff %.3g" % (ns, err)
            # the rows past ns are left alone
            rest = d_out.get((4096, 4), np.float64)[ns:]
            assert np.all(rest == 7.0)
        # out-of-range target indices: NaN rows, the others unchanged
        idx = np.array([5, -1, n, 17], np.int32)
        ref = t.direct_fp64(np.array([5, 17], np.int32))
        d_idx.put(idx)
        assert t.direct_fp64_noalloc(4, d_idx.ptr.value, d_out.ptr.value) == 0
        assert hip.hipDeviceSynchronize() == 0
        got = d_out.get((4, 4), np.float64)
        assert np.all(np.isnan(got[1])) and np.all(np.isnan(got[2]))
        assert np.array_equal(got[[0, 3]], ref)
        # sizes beyond the scratch are refused before anything is launched
        assert t.direct_fp64_noalloc(4097, d_idx.ptr.value, d_out.ptr.value) == BT_ERR_ARG
        assert t.direct_fp64_noalloc(0, d_idx.ptr.value, d_out.ptr.value) == BT_ERR_ARG
        assert t.direct_fp64_noalloc(4, 0, d_out.ptr.value) == BT_ERR_ARG
    finally:
        d_idx.free()
        d_out.free()
        t.close()
