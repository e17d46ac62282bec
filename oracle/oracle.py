"""ctypes wrapper of liboracle.so, the CPU restatement of WeTriC's counting path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product never loads
it.  Function-level citations of the reference are in oracle.h / oracle.c.

Parity status of each function (how it is pinned) is listed in
tests/test_oracle.py; every function has at least one independent check.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
_lib = C.CDLL(LIB_PATH)

TREE_SIZE = 4095


class _G(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("rowPtr", C.POINTER(C.c_uint64)),
                ("colInd", C.POINTER(C.c_uint32))]


_P = C.c_void_p
_u64p = C.POINTER(C.c_uint64)
_sig = {
    "orc_read_graph": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(_G)]),
    "orc_graph_from_pairs": (C.c_int, [_P, C.c_uint64, C.POINTER(_G)]),
    "orc_preprocess": (C.c_int, [C.POINTER(_G), C.POINTER(_G)]),
    "orc_free": (None, [C.POINTER(_G)]),
    "orc_wedge_prefix": (C.c_uint64, [C.POINTER(_G), _P]),
    "orc_search_tree": (None, [_P, C.c_uint32, _P]),
    "orc_search_closest": (C.c_uint32, [_P, C.c_uint32, C.c_uint32, C.c_uint64]),
    "orc_search_closest_tree": (C.c_uint32, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint64]),
    "orc_decode": (None, [C.c_uint64, C.c_uint64, _u64p, _u64p]),
    "orc_decode_ref_fp": (None, [C.c_uint64, C.c_uint64, _u64p, _u64p]),
    "orc_bitmap_bits": (C.c_uint64, [C.c_uint32]),
    "orc_bitmap_words": (C.c_uint64, [C.c_uint32]),
    "orc_bitmap_index": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "orc_bitmap_build": (None, [C.POINTER(_G), C.c_uint32, _P]),
    "orc_tc_wedge": (C.c_int, [C.POINTER(_G), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, _u64p]),
    "orc_arrow_prefix": (C.c_uint64, [C.POINTER(_G), _P]),
    "orc_tc_arrow": (C.c_int, [C.POINTER(_G), C.c_uint32, _u64p, _u64p]),
    "orc_tc_style": (C.c_int, [C.POINTER(_G), C.c_int, C.c_uint32, _u64p]),
    "orc_tc_merge": (C.c_uint64, [C.POINTER(_G), C.c_int]),
    "orc_tc_merge_range": (C.c_uint64, [C.POINTER(_G), C.c_uint32, C.c_uint32, C.c_int]),
    "orc_decode_check": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_int, _u64p]),
    "orc_alg_bytes": (None, [C.POINTER(_G), C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
}
for _n, (_r, _a) in _sig.items():
    _f = getattr(_lib, _n)
    _f.restype = _r
    _f.argtypes = _a


@dataclass
class OGraph:
    n: int
    m: int
    row_ptr: np.ndarray
    col_ind: np.ndarray


def _struct(g) -> _G:
    rp = np.ascontiguousarray(g.row_ptr, dtype=np.uint64)
    ci = np.ascontiguousarray(g.col_ind, dtype=np.uint32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.uint32)
    s = _G(g.n, g.m, rp.ctypes.data_as(_u64p), ci.ctypes.data_as(C.POINTER(C.c_uint32)))
    s._keep = (rp, ci)
    return s


def _take(s: _G) -> OGraph:
    n, m = int(s.n), int(s.m)
    rp = np.ctypeslib.as_array(s.rowPtr, shape=(n + 1,)).copy()
    ci = np.ctypeslib.as_array(s.colInd, shape=(m,)).copy() if m else np.zeros(0, np.uint32)
    _lib.orc_free(C.byref(s))
    return OGraph(n, m, rp, ci)


def _rc(rc, where):
    if rc != 0:
        raise RuntimeError(f"oracle {where} failed: {rc}")


def read_graph(path, matrix_market=False, zero_indexed=False) -> OGraph:
    s = _G()
    rc = _lib.orc_read_graph(path.encode(), int(matrix_market), int(zero_indexed), C.byref(s))
    if rc != 0:
        return rc
    return _take(s)


def graph_from_pairs(pairs) -> OGraph:
    p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    s = _G()
    _rc(_lib.orc_graph_from_pairs(p.ctypes.data if p.size else None, p.shape[0], C.byref(s)), "from_pairs")
    return _take(s)


def preprocess(g) -> OGraph:
    s = _struct(g)
    o = _G()
    _rc(_lib.orc_preprocess(C.byref(s), C.byref(o)), "preprocess")
    return _take(o)


def wedge_prefix(g) -> np.ndarray:
    s = _struct(g)
    ws = np.zeros(g.n + 1, dtype=np.uint64)
    _lib.orc_wedge_prefix(C.byref(s), ws.ctypes.data)
    return ws


def search_tree(ws: np.ndarray, n: int) -> np.ndarray:
    tree = np.zeros(TREE_SIZE, dtype=np.uint64)
    ws = np.ascontiguousarray(ws, dtype=np.uint64)
    _lib.orc_search_tree(ws.ctypes.data, n, tree.ctypes.data)
    return tree


def search_closest(lst, start, end, target) -> int:
    lst = np.ascontiguousarray(lst, dtype=np.uint64)
    return int(_lib.orc_search_closest(lst.ctypes.data, start, end, target))


def search_closest_tree(tree, lst, start, end, target) -> int:
    lst = np.ascontiguousarray(lst, dtype=np.uint64)
    tree = np.ascontiguousarray(tree, dtype=np.uint64)
    return int(_lib.orc_search_closest_tree(tree.ctypes.data, lst.ctypes.data, start, end, target))


def decode(d, iv):
    w, u = C.c_uint64(), C.c_uint64()
    _lib.orc_decode(d, iv, C.byref(w), C.byref(u))
    return int(w.value), int(u.value)


def decode_ref_fp(d, iv):
    w, u = C.c_uint64(), C.c_uint64()
    _lib.orc_decode_ref_fp(d, iv, C.byref(w), C.byref(u))
    return int(w.value), int(u.value)


def bitmap_words(a) -> int:
    return int(_lib.orc_bitmap_words(a))


def bitmap_index(n, a, w, u) -> int:
    return int(_lib.orc_bitmap_index(n, a, w, u))


def bitmap_build(g, a) -> np.ndarray:
    s = _struct(g)
    words = np.zeros(max(bitmap_words(a), 1), dtype=np.uint32)
    _lib.orc_bitmap_build(C.byref(s), a, words.ctypes.data)
    return words[:bitmap_words(a)]


def tc_wedge(g, spread, a, w_begin=0, w_end=2 ** 64 - 1) -> int:
    s = _struct(g)
    out = C.c_uint64()
    _rc(_lib.orc_tc_wedge(C.byref(s), spread, a, w_begin, w_end, C.byref(out)), "tc_wedge")
    return int(out.value)


def arrow_prefix(g) -> np.ndarray:
    s = _struct(g)
    vi = np.zeros(g.m + 1, dtype=np.uint64)
    _lib.orc_arrow_prefix(C.byref(s), vi.ctypes.data)
    return vi


def tc_arrow(g, a=0):
    """(triangles, arrow wedges) of the arrow-wedge count."""
    s = _struct(g)
    A, out = C.c_uint64(), C.c_uint64()
    _rc(_lib.orc_tc_arrow(C.byref(s), a, C.byref(A), C.byref(out)), "tc_arrow")
    return int(out.value), int(A.value)


def tc_style(g, style, a=0) -> int:
    """Count of the vertex-/edge-parallel baselines in wedge style 0/1/2."""
    s = _struct(g)
    out = C.c_uint64()
    _rc(_lib.orc_tc_style(C.byref(s), style, a, C.byref(out)), "tc_style")
    return int(out.value)


def tc_merge(g, threads=0) -> int:
    s = _struct(g)
    return int(_lib.orc_tc_merge(C.byref(s), threads))


def tc_merge_range(g, v0, v1, threads=0) -> int:
    s = _struct(g)
    return int(_lib.orc_tc_merge_range(C.byref(s), v0, v1, threads))


def decode_check(d_lo, d_hi, fp=True):
    """0 if orc_decode (and the reference formula when fp) match enumeration."""
    bad = C.c_uint64()
    d = int(_lib.orc_decode_check(d_lo, d_hi, int(fp), C.byref(bad)))
    return d, int(bad.value)


def alg_bytes(g, offset_bytes=4):
    s = _struct(g)
    a, b = C.c_double(), C.c_double()
    _lib.orc_alg_bytes(C.byref(s), offset_bytes, C.byref(a), C.byref(b))
    return float(a.value), float(b.value)
