"""GPU: wtc_count_noalloc (the count of a caller-owned device CSR with every
workspace buffer in the library's fixed per-device scratch array) against
wtc_dgraph_count (the same count with library-allocated workspace) and the
oracle, through the C ABI.  The caller's device buffers are allocated here
with the HIP runtime, exactly as large as the CSR (no tail slack)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import wtc

pytestmark = pytest.mark.gpu

UINT64_MAX = 2 ** 64 - 1


def _hip():
    """The HIP runtime that libwtc.so already loaded into this process."""
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 not mapped"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevCSR:
    """A caller-owned device copy of an oriented CSR (rowPtr in the given
    width, colInd) plus an 8-byte device output word."""

    def __init__(self, hip, g, ob):
        self.hip, self.n, self.m, self.ob = hip, g.n, g.m, ob
        rp = np.ascontiguousarray(g.row_ptr, dtype=np.uint32 if ob == 4 else np.uint64)
        ci = np.ascontiguousarray(g.col_ind, dtype=np.uint32)
        self.rp, self.ci, self.out = C.c_void_p(), C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(self.rp), rp.nbytes) == 0
        assert hip.hipMalloc(C.byref(self.ci), max(ci.nbytes, 4)) == 0
        assert hip.hipMalloc(C.byref(self.out), 8) == 0
        assert hip.hipMemcpy(self.rp, rp.ctypes.data, rp.nbytes, H2D) == 0
        if ci.nbytes:
            assert hip.hipMemcpy(self.ci, ci.ctypes.data, ci.nbytes, H2D) == 0

    def count(self, spread, a, w_begin=0, w_end=UINT64_MAX, n=None):
        rc = wtc._lib.wtc_count_noalloc(self.rp, self.ob, self.ci, self.n if n is None else n, self.m, spread, a,
                                        w_begin, w_end, self.out, None)
        if rc != 0:
            return rc, None
        assert self.hip.hipDeviceSynchronize() == 0
        res = np.zeros(1, np.uint64)
        assert self.hip.hipMemcpy(res.ctypes.data, self.out, 8, D2H) == 0
        return 0, int(res[0])

    def free(self):
        for p in (self.rp, self.ci, self.out):
            self.hip.hipFree(p)


def _rmat(scale, seed):
    return wtc.preprocess(wtc.gen_rmat(scale, 16, seed=seed))


def test_noalloc_matches_allocating_count(gpu_ctx):
    hip = _hip()
    big, small = _rmat(14, 9), _rmat(10, 3)
    for g in (big, small):  # two calls in a row, the second on a smaller graph
        expect = O.tc_merge(g)
        for ob in (4, 8):
            dc = DevCSR(hip, g, ob)
            dg = gpu_ctx.upload(g, offset_bytes=ob)
            try:
                for s, a in ((5, 64), (12, 4096), (1, 0), (17, 32)):
                    ref = dg.count(s, a)
                    assert ref == expect, (g.n, ob, s, a)
                    assert dc.count(s, a) == (0, ref), (g.n, ob, s, a)
                W = g.wedges()
                cuts = [0, W // 3, (2 * W) // 3, W]
                parts = []
                for b, e in zip(cuts[:-1], cuts[1:]):
                    rc, got = dc.count(3, 64, b, e)
                    assert rc == 0 and got == dg.count(3, 64, b, e), (b, e)
                    parts.append(got)
                assert sum(parts) == expect
            finally:
                dg.free()
                dc.free()


def test_noalloc_rejects_and_recovers(gpu_ctx):
    hip = _hip()
    g = _rmat(11, 5)
    expect = O.tc_merge(g)
    # too large for the scratch array: rejected on the host, nothing enqueued
    dc = DevCSR(hip, g, 4)
    try:
        assert dc.count(5, 64, n=40_000_000)[0] == -7
        assert dc.count(5, 1 << 16)[0] == -7  # a bitmap over 2^16 ranks alone needs 512 MiB
        assert dc.count(0, 64)[0] == -1
        assert dc.count(5, 33)[0] == -1
        assert dc.count(5, 64) == (0, expect)
    finally:
        dc.free()
    # a CSR that is not oriented: checked on the device, not counted
    ci = g.col_ind.copy()
    ci[-1] = g.n  # an id out of range
    bad = wtc.CSR(g.n, g.m, g.row_ptr.copy(), ci)
    ci2 = g.col_ind.copy()
    r = int(np.argmax(np.diff(g.row_ptr.astype(np.int64)) >= 2))
    r0 = int(g.row_ptr[r])
    ci2[r0], ci2[r0 + 1] = ci2[r0 + 1], ci2[r0]  # a row out of order
    bad2 = wtc.CSR(g.n, g.m, g.row_ptr.copy(), ci2)
    for b in (bad, bad2):
        for ob in (4, 8):
            dc = DevCSR(hip, b, ob)
            try:
                assert dc.count(5, 64) == (0, UINT64_MAX)
            finally:
                dc.free()
    # the scratch array keeps the last call's contents: a valid call after
    # invalid ones is still exact
    dc = DevCSR(hip, g, 8)
    try:
        assert dc.count(5, 64) == (0, expect)
    finally:
        dc.free()
