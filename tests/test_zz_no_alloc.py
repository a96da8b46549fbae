"""The allocation-free pair finder (prx_lbvh_pairs_noalloc) against the
allocating one (prx_lbvh_pairs): same pairs in the same order, on two calls
in a row (the second with fewer proxies, so stale scratch contents from the
first call must not leak into it), through the C ABI on device buffers.

The test allocates its device buffers with the HIP runtime the library
itself links (libamdhip64.so.7), so one runtime owns the device."""
import ctypes

import numpy as np
import pytest

import playrho_amd as pa

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the library's HIP runtime."""

    def __init__(self):
        pa.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.live = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(nbytes, 8)) == 0
        self.live.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _scene(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 10 + np.sqrt(n) * 2, size=(n, 2)).astype(np.float32)
    k = n // 10
    if k:
        c[:k] = c[n - k:]  # repeated centres: equal Morton codes
    h = rng.uniform(0.05, 1.5, size=(n, 2)).astype(np.float32)
    aabbs = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    bodies = (np.arange(n) // 2).astype(np.int32)
    moved = (rng.uniform(size=n) < 0.5).astype(np.uint8)
    return aabbs, bodies, moved


def _noalloc(aabbs, bodies, moved, cap):
    hip = _Hip()
    try:
        n = aabbs.shape[0]
        da, db, dm = hip.put(aabbs.reshape(-1)), hip.put(bodies), hip.put(moved)
        dp = hip.put(np.full((cap, 2), -7, dtype=np.int32))
        dc = hip.put(np.full(1, -9, dtype=np.int32))
        st = pa.lib().prx_lbvh_pairs_noalloc(da, db, dm, n, dp, cap, dc, None)
        assert st == 0, pa.lib().prx_last_error()
        count = int(hip.get(dc, np.int32, 1)[0])
        pairs = hip.get(dp, np.int32, 2 * cap).reshape(cap, 2)
        return count, pairs
    finally:
        hip.free()


def test_noalloc_pairs_equal_allocating_pairs_two_calls():
    for n, seed in ((6000, 11), (900, 12)):
        print("noalloc lbvh n", n, "seed", seed, flush=True)
        aabbs, bodies, moved = _scene(n, seed)
        cap = 16 * n
        want = pa.lbvh_pairs(aabbs, bodies, moved, cap=cap)
        count, pairs = _noalloc(aabbs, bodies, moved, cap)
        assert count == want.shape[0]
        assert np.array_equal(pairs[:count], want)
        assert np.all(pairs[count:] == -7)  # nothing written past the count


def test_noalloc_rejects_oversize_and_reports_truncation():
    n = (1 << 16) + 1
    hip = _Hip()
    try:
        buf = hip.alloc(16 * n)
        st = pa.lib().prx_lbvh_pairs_noalloc(buf, buf, buf, n, buf, 4, buf, None)
    finally:
        hip.free()
    assert st == 2  # PRX_E_LENGTH, before any launch
    # Capacity below the pair count: the count says how many there were.
    aabbs, bodies, moved = _scene(500, 13)
    want = pa.lbvh_pairs(aabbs, bodies, moved, cap=16 * 500)
    assert want.shape[0] > 10
    count, pairs = _noalloc(aabbs, bodies, moved, 10)
    assert count == want.shape[0]
    assert np.array_equal(pairs[:10], want[:10])
