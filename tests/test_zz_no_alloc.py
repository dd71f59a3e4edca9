"""smdp_sssp_device_noalloc (no device allocation; working buffers in the
library's fixed per-GPU array) against the existing entry point smdp_sssp on
the same graphs, two calls in a row with the second graph smaller, for each
solution; plus its size limits and its check of a malformed device CSR.

The test allocates the caller's device buffers itself (through the HIP
runtime): only the entry point must not allocate.
"""
import ctypes as C

import numpy as np
import pytest

import orc
import smdp
from conftest import random_graph

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArrays:
    """The caller's device copies of a CSR and an output array."""

    def __init__(self, g):
        self.bufs = []
        self.n, self.m = g.n, g.m
        self.row = self._up(np.ascontiguousarray(g.row, np.int32))
        self.col = self._up(np.ascontiguousarray(g.col, np.int32))
        self.w = self._up(np.ascontiguousarray(g.w, np.int32))
        self.dist = self._alloc(4 * g.n)

    def _alloc(self, nbytes):
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 4))) == 0
        self.bufs.append(p)
        return p

    def _up(self, a):
        p = self._alloc(a.nbytes)
        if a.nbytes:
            assert hip().hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def read_dist(self):
        out = np.zeros(self.n, np.int32)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.dist, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.bufs:
            hip().hipFree(p)
        self.bufs = []


def noalloc_sssp(ctx, d, src, sol, n=None, m=None):
    st = smdp.Stats()
    rc = smdp.lib().smdp_sssp_device_noalloc(ctx.h, d.n if n is None else n, d.m if m is None else m, d.row, d.col,
                                             d.w, src, sol, d.dist, C.byref(st))
    return rc, st


@pytest.fixture(scope="module")
def nctx():
    c = smdp.Context(0)
    yield c
    c.close()


def _graphs():
    big = smdp.gen_rmat(16, 16, wlo=1, whi=255, seed=3)
    row, col, w = random_graph(3000, 9000, 41)
    small = smdp.HostGraph(row, col, w, 0)
    return big, small


@pytest.mark.parametrize("sol", [smdp.SOL_FLAT, smdp.SOL_BLOCK, smdp.SOL_GRID, smdp.SOL_TUNED])
def test_noalloc_matches_dropin(nctx, sol):
    big, small = _graphs()
    for g in (big, small):  # two calls in a row, the second on a smaller graph
        print(f"noalloc sol={sol} n={g.n} m={g.m}", flush=True)
        ref, _ = nctx.sssp(g, g.source, sol)  # the existing entry point
        d = DeviceArrays(g)
        try:
            rc, st = noalloc_sssp(nctx, d, g.source, sol)
            assert rc == smdp.OK, smdp.strerror(rc)
            out = d.read_dist()
        finally:
            d.close()
        assert np.array_equal(out, ref)
        assert np.array_equal(out, orc.dijkstra(g.row, g.col, g.w, g.source))
        reached = ref != orc.INF
        assert st.m_reached == int(np.diff(g.row)[reached].sum())


def test_noalloc_limits_and_malformed(nctx):
    _, small = _graphs()
    d = DeviceArrays(small)
    try:
        # sizes beyond the fixed array are rejected before any device work
        rc, _ = noalloc_sssp(nctx, d, 0, smdp.SOL_TUNED, n=smdp_max_n() + 1)
        assert rc == smdp.ERR_ARG
        rc, _ = noalloc_sssp(nctx, d, 0, smdp.SOL_TUNED, m=smdp_max_m() + 1)
        assert rc == smdp.ERR_ARG
        rc, _ = noalloc_sssp(nctx, d, small.n, smdp.SOL_TUNED)  # source out of range
        assert rc == smdp.ERR_ARG
        rc, _ = noalloc_sssp(nctx, d, 0, smdp.SOL_HIER)
        assert rc == smdp.ERR_SOLUTION
    finally:
        d.close()
    # a column out of range is caught on the device, and the context still works
    bad = smdp.HostGraph(small.row, small.col.copy(), small.w, 0)
    bad.col[len(bad.col) // 2] = small.n
    d = DeviceArrays(bad)
    try:
        rc, _ = noalloc_sssp(nctx, d, 0, smdp.SOL_TUNED)
        assert rc == smdp.ERR_ARG
    finally:
        d.close()
    d = DeviceArrays(small)
    try:
        rc, _ = noalloc_sssp(nctx, d, 0, smdp.SOL_TUNED)
        assert rc == smdp.OK
        assert np.array_equal(d.read_dist(), orc.dijkstra(small.row, small.col, small.w, 0))
    finally:
        d.close()


def smdp_max_n():
    return 1 << 19


def smdp_max_m():
    return 1 << 23
