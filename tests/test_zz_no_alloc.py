"""The allocation-free extraction (dct_map_extract64_noalloc) against the
engine's dct_map_extract64 on the same map, through the C ABI: two calls in
a row on one GPU, the second on a smaller map (the static scratch keeps the
first call's contents), plus the host-side rejections (size above the
scratch's 2^30 words, misaligned map) that queue no device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class _Hip:
    """Device buffers for the test (the entry point itself allocates none)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 256))) == 0
        self.bufs.append(p)
        return p

    def put(self, dst, a):
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_size_t(a.nbytes), 1) == 0   # H2D

    def get(self, src, a):
        assert self.lib.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), src,
                                  ctypes.c_size_t(a.nbytes), 2) == 0   # D2H

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _engine_extract(g, thr, cap):
    """The existing entry point on the live cell's map."""
    idx = np.empty(max(cap, 1), dtype=np.int64)
    val = np.empty(max(cap, 1), dtype=np.int32)
    found = g._lib.dct_map_extract64(int(thr), idx.ctypes.data_as(_i64p),
                                     val.ctypes.data_as(_i32p), int(cap))
    assert found >= 0
    m = min(found, cap)
    return idx[:m], val[:m], found


def _noalloc_extract(lib, hip, dmap, n, thr, cap, didx, dval, dfound):
    fn = lib.dct_map_extract64_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int32,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                   ctypes.c_void_p, ctypes.c_void_p]
    rc = fn(dmap, n, thr, didx, dval, cap, dfound, None)
    assert rc == 0
    hip.sync()
    found = np.zeros(1, dtype=np.int64)
    hip.get(dfound, found)
    m = int(min(found[0], cap))
    idx = np.empty(max(m, 1), dtype=np.int64)
    val = np.empty(max(m, 1), dtype=np.int32)
    if m:
        hip.get(didx, idx)
        hip.get(dval, val)
    return idx[:m], val[:m], int(found[0])


def test_noalloc_extract_equals_engine_extract():
    from coldcore_amd.gpu import GpuBackend
    hip = _Hip()
    try:
        cap_max = 1 << 18
        dmap = hip.alloc(4 << 20)                 # room for 2^20 words
        didx = hip.alloc(8 * cap_max)
        dval = hip.alloc(4 * cap_max)
        dfound = hip.alloc(8)
        for axes, R, p in (([2] * 20, 3, 3e-4), ([3] * 8, 2, 2e-2)):
            size = int(np.prod(axes))
            print(f"no-alloc extraction: axes={axes} R={R} words={size}", flush=True)
            code = np.flatnonzero(np.random.default_rng(size).random(size) < p)
            with GpuBackend(axes, R) as g:
                g.recount(code.astype(np.int64))
                g.gain_map()
                mp = g.map_all()
                mx = int(mp.max())
                for thr, capf in ((max(1, mx // 2), 0.5), (1, 2.0), (mx + 1, 1.0)):
                    full = int((mp >= thr).sum())
                    cap = min(cap_max, max(1, int(full * capf)))
                    want = _engine_extract(g, thr, cap)
                    hip.put(dmap, mp)
                    got = _noalloc_extract(g._lib, hip, dmap, size, thr, cap,
                                           didx, dval, dfound)
                    assert got[2] == want[2] == full, (thr, cap)
                    assert (got[0] == want[0]).all() and (got[1] == want[1]).all()
                    assert (np.diff(got[0]) > 0).all()
    finally:
        hip.free()


def test_noalloc_extract_rejects_without_device_work():
    from coldcore_amd.gpu import load_plugin
    lib = load_plugin()
    fn = lib.dct_map_extract64_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int32,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                   ctypes.c_void_p, ctypes.c_void_p]
    hip = _Hip()
    try:
        d = hip.alloc(1 << 12)
        # above the scratch's 2^30 words: refused on the host
        assert fn(d, (1 << 30) + 1, 1, d, d, 1, d, None) == -1
        # a map that is not 16-byte aligned, a negative cap, no found slot
        assert fn(ctypes.c_void_p(d.value + 4), 16, 1, d, d, 1, d, None) == -1
        assert fn(d, 16, 1, d, d, -1, d, None) == -1
        assert fn(d, 16, 1, d, d, 1, None, None) == -1
    finally:
        hip.free()
