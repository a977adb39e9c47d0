"""ghet_bfs_noalloc through the C ABI against ghet_bfs on the same graphs.

Device buffers for the inputs and outputs come from the HIP runtime that
libghet.so itself links (hipMalloc / hipMemcpy through ctypes): the entry
point under test allocates nothing.  Two calls in a row, the second on a
smaller graph (the shared scratch keeps the first call's contents), then a
second source on the first graph, then the size and argument checks.
"""
import ctypes as C

import numpy as np
import pytest

import graphhet_amd as G

pytestmark = pytest.mark.gpu

MAX_VERTICES = 1 << 22


def _hip():
    G.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


H2D, D2H = 1, 2


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 4)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc_bfs(hip, o, src):
    rp, col = DevArray(hip, 4 * (o.n + 1)), DevArray(hip, 4 * o.m)
    rp.put(o.rowptr.astype(np.int32))
    col.put(o.col.astype(np.int32))
    out = DevArray(hip, 4 * o.n)
    out.put(np.full(o.n, 12345, np.int32))  # stale contents must not leak into the result
    try:
        st = G.lib().ghet_bfs_noalloc(o.n, o.m, rp.p, col.p, src, out.p, None)
        assert st == G.OK, G.lib().ghet_last_error()
        return out.get(np.int32, o.n)
    finally:
        for a in (rp, col, out):
            a.free()


def test_noalloc_bfs_matches_ghet_bfs_twice_then_smaller():
    hip = _hip()
    cases = [G.HostGraph.rmat(16, 16, seed=1), G.HostGraph.rmat(11, 8, seed=4), G.HostGraph.er(3000, 9000, seed=2)]
    for k, hg in enumerate(cases):
        o = hg.view(G.VIEW_OUT)
        dg = G.DeviceGraph.from_host(hg)
        srcs = [int(np.argmax(o.degrees())), 0] if k == 0 else [int(np.argmax(o.degrees()))]
        for src in srcs:
            ref, _ = dg.bfs(src)
            got = _noalloc_bfs(hip, o, src)
            np.testing.assert_array_equal(got, ref, err_msg=f"case {k} source {src}")
        dg.close()


def test_noalloc_bfs_rejects_what_it_cannot_hold():
    L = G.lib()
    dummy = C.c_void_p(0x1000)  # rejected before any device access
    assert L.ghet_bfs_noalloc(MAX_VERTICES + 1, 10, dummy, dummy, 0, dummy, None) == G.ERR_UNSUPPORTED
    assert b"scratch" in L.ghet_last_error()
    assert L.ghet_bfs_noalloc(10, 10, dummy, dummy, 10, dummy, None) == G.ERR_INVALID
    assert L.ghet_bfs_noalloc(10, 10, None, dummy, 0, dummy, None) == G.ERR_INVALID
    assert L.ghet_bfs_noalloc(0, 0, None, None, 0, None, None) == G.OK
