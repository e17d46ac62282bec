"""Python binding (ctypes) of libwtc, the MI355X wedge-parallel triangle counter.

The library's C ABI is declared in include/wtc.h.  This module mirrors the
reference's operator seams -- read_graph / preprocess / tc_GPU of tc.cu
(tc.cu:561, tc.cu:687, tc.cu:315) -- plus the device-resident, batched and
served paths.  There is no CPU fallback: every GPU entry point runs the HIP
kernels in libwtc.so, and importing this module fails loudly when the library
has not been built.

Graphs cross the boundary as CSR objects: n vertices, m edges, a uint64
row_ptr of n+1 offsets and a uint32 col_ind of m ids (the BIG_GRAPH_TYPE
layout of experiments/tc_big_graphs.cu:58-63).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwtc.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")

_lib = C.CDLL(LIB_PATH)

WTC_OK = 0
ERRORS = {
    -1: "WTC_ERR_ARG", -2: "WTC_ERR_IO", -3: "WTC_ERR_FORMAT", -4: "WTC_ERR_ZERO_ID",
    -5: "WTC_ERR_NOMEM", -6: "WTC_ERR_HIP", -7: "WTC_ERR_RANGE", -8: "WTC_ERR_NODEV",
    -9: "WTC_ERR_ADMIT", -10: "WTC_ERR_CSR", -11: "WTC_ERR_BUSY",
}


class WtcError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        detail = _lib.wtc_last_error_detail().decode()
        msg = f"{where}: {ERRORS.get(code, code)} ({_lib.wtc_strerror(code).decode()})"
        super().__init__(msg + (f" [{detail}]" if detail else ""))


class _BigGraph(C.Structure):
    _fields_ = [("numVertices", C.c_uint32), ("numEdges", C.c_uint64),
                ("rowPtr", C.POINTER(C.c_uint64)), ("colInd", C.POINTER(C.c_uint32))]


class _Graph32(C.Structure):
    _fields_ = [("numVertices", C.c_uint32), ("numEdges", C.c_uint32),
                ("rowPtr", C.POINTER(C.c_uint32)), ("colInd", C.POINTER(C.c_uint32))]


class GpuTime(C.Structure):
    """GPU_time of tc.cu:76-79 (milliseconds, accumulated)."""
    _fields_ = [("copy", C.c_double), ("exec", C.c_double)]


class CountStats(C.Structure):
    _fields_ = [("prefix_ms", C.c_double), ("bitmap_ms", C.c_double), ("count_ms", C.c_double),
                ("total_ms", C.c_double), ("wedges", C.c_uint64), ("blocks", C.c_uint32),
                ("launches", C.c_uint32), ("small_graphs", C.c_uint32), ("split", C.c_uint32),
                ("rows_ms", C.c_double), ("top_wedges", C.c_uint64), ("in_edges", C.c_uint64),
                ("list_ms", C.c_double), ("list_in_edges", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch", C.c_uint32), ("slots", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("probe_ways", C.c_int32), ("lds_chunk", C.c_uint32),
                ("waves_per_block", C.c_int32), ("small_graph_words", C.c_int32)]


class RequestStats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("service_ms", C.c_double), ("latency_ms", C.c_double),
                ("batch_graphs", C.c_uint32), ("batch_id", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("graphs", "submitted", "completed", "rejected", "batches",
                                          "max_batch_graphs", "uploads", "evictions", "resident_bytes",
                                          "resident_graphs", "queued")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_P = C.c_void_p
_BGP = C.POINTER(_BigGraph)
_sig = {
    "wtc_strerror": (C.c_char_p, [C.c_int]),
    "wtc_last_error_detail": (C.c_char_p, []),
    "wtc_read_graph": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(_BGP)]),
    "wtc_graph_from_pairs": (C.c_int, [_P, C.c_uint64, C.POINTER(_BGP)]),
    "wtc_write_csr": (C.c_int, [C.c_char_p, _BGP, C.c_uint64]),
    "wtc_read_csr": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(_BGP)]),
    "wtc_preprocess": (C.c_int, [_BGP, C.c_int, C.POINTER(_BGP)]),
    "wtc_free_graph": (None, [_BGP]),
    "wtc_narrow_graph": (C.c_int, [_BGP, C.POINTER(C.POINTER(_Graph32))]),
    "wtc_free_graph32": (None, [C.POINTER(_Graph32)]),
    "wtc_validate_oriented": (C.c_int, [_BGP]),
    "wtc_wedge_total": (C.c_uint64, [_BGP]),
    "wtc_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                              C.POINTER(_BGP)]),
    "wtc_gen_rmat_pairs": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                    C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint64)]),
    "wtc_gen_chung_lu": (C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
                                  C.POINTER(_BGP)]),
    "wtc_gen_erdos_renyi": (C.c_int, [C.c_uint32, C.c_double, C.c_uint64, C.POINTER(_BGP)]),
    "wtc_gen_barabasi_albert": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(_BGP)]),
    "wtc_free_pairs": (None, [C.POINTER(C.c_uint32)]),
    "wtc_tc_gpu": (C.c_int, [C.POINTER(_Graph32), C.c_uint32, C.c_uint32, C.POINTER(GpuTime),
                            C.POINTER(C.c_uint64)]),
    "wtc_tc_gpu_big": (C.c_int, [_BGP, C.c_uint32, C.c_uint32, C.POINTER(GpuTime), C.POINTER(C.c_uint64)]),
    "wtc_tc_arrow_gpu": (C.c_int, [C.POINTER(_Graph32), C.c_uint32, C.c_uint32, C.POINTER(GpuTime),
                                  C.POINTER(C.c_uint64)]),
    "wtc_tc_arrow_gpu_big": (C.c_int, [_BGP, C.c_uint32, C.c_uint32, C.POINTER(GpuTime),
                                      C.POINTER(C.c_uint64)]),
    "wtc_tc_vertex_gpu": (C.c_int, [C.POINTER(_Graph32), C.c_uint32, C.c_int, C.POINTER(GpuTime),
                                   C.POINTER(C.c_uint64)]),
    "wtc_tc_vertex_gpu_big": (C.c_int, [_BGP, C.c_uint32, C.c_int, C.POINTER(GpuTime), C.POINTER(C.c_uint64)]),
    "wtc_tc_edge_gpu": (C.c_int, [C.POINTER(_Graph32), C.c_uint32, C.c_int, C.POINTER(GpuTime),
                                 C.POINTER(C.c_uint64)]),
    "wtc_tc_edge_gpu_big": (C.c_int, [_BGP, C.c_uint32, C.c_int, C.POINTER(GpuTime), C.POINTER(C.c_uint64)]),
    "wtc_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "wtc_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "wtc_ctx_destroy": (None, [_P]),
    "wtc_dgraph_upload": (C.c_int, [_P, _BGP, C.c_uint32, C.POINTER(_P)]),
    "wtc_dgraph_free": (None, [_P]),
    "wtc_dgraph_bytes": (C.c_uint64, [_P]),
    "wtc_dgraph_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "wtc_dgraph_from_pairs": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.POINTER(_P)]),
    "wtc_dgraph_gen_rmat": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                     C.c_uint32, C.POINTER(_P)]),
    "wtc_dgraph_download": (C.c_int, [_P, C.POINTER(_BGP)]),
    "wtc_dgraph_count": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                                  C.POINTER(C.c_uint64), C.POINTER(CountStats)]),
    "wtc_dgraph_count_arrow": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64), C.POINTER(CountStats)]),
    "wtc_dgraph_count_baseline": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_uint32, C.POINTER(C.c_uint64),
                                           C.POINTER(CountStats)]),
    "wtc_autotune_shape": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, _P]),
    "wtc_tuning_source": (C.c_char_p, []),
    "wtc_ctx_use_tuned": (C.c_int, [_P, _P]),
    "wtc_count_noalloc": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                   C.c_uint64, C.c_uint64, _P, _P]),
    "wtc_ctx_debug_prefix": (C.c_int, [_P, C.c_uint64, _P, _P]),
    "wtc_ctx_debug_bitmap": (C.c_int, [_P, C.c_uint64, _P]),
    "wtc_debug_decode": (C.c_int, [_P, _P, _P, C.c_uint64, _P, _P]),
    "wtc_debug_copy": (C.c_int, [_P, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "wtc_ctx_set_tuning": (C.c_int, [_P, C.c_int, C.c_uint32, C.c_int]),
    "wtc_ctx_set_small_graphs": (C.c_int, [_P, C.c_uint32]),
    "wtc_ctx_set_count_kernel": (C.c_int, [_P, C.c_int]),
    "wtc_ctx_set_work_split": (C.c_int, [_P, C.c_uint32, C.c_int]),
    "wtc_ctx_get_small_graphs": (C.c_int, [_P, _P, _P]),
    "wtc_alg_bytes": (None, [_BGP, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "wtc_batch_upload": (C.c_int, [_P, C.POINTER(_BGP), C.c_size_t, C.POINTER(_P)]),
    "wtc_batch_free": (None, [_P]),
    "wtc_batch_count": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(CountStats)]),
    "wtc_batch_count_host": (C.c_int, [_P, C.POINTER(_BGP), C.c_size_t, _P, _P, _P, C.POINTER(CountStats)]),
    "wtc_autotune": (None, [C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "wtc_autotune_batch": (None, [C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32)]),
    "wtc_server_create": (C.c_int, [C.c_int, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "wtc_server_register": (C.c_int, [_P, _BGP, C.POINTER(C.c_uint64)]),
    "wtc_server_submit": (C.c_int, [_P, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    "wtc_server_wait": (C.c_int, [_P, C.c_uint64, C.c_int64, C.POINTER(C.c_uint64), C.POINTER(RequestStats)]),
    "wtc_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "wtc_server_destroy": (None, [_P]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_sig)


def _check(rc: int, where: str):
    if rc != WTC_OK:
        raise WtcError(rc, where)


# ------------------------------------------------------------------ graphs

class _Owner:
    """Frees a library-allocated wtc_big_graph when the last view dies."""

    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        if self.ptr:
            _lib.wtc_free_graph(self.ptr)
            self.ptr = None


def _view(addr: int, count: int, ctype, owner) -> np.ndarray:
    if count == 0:
        return np.zeros(0, dtype=np.dtype(ctype))
    buf = (ctype * count).from_address(addr)
    buf._owner = owner
    return np.frombuffer(buf, dtype=np.dtype(ctype))


@dataclass
class CSR:
    """A CSR graph: row_ptr (uint64, n+1), col_ind (uint32, m)."""
    n: int
    m: int
    row_ptr: np.ndarray
    col_ind: np.ndarray

    def _struct(self) -> _BigGraph:
        rp = np.ascontiguousarray(self.row_ptr, dtype=np.uint64)
        ci = np.ascontiguousarray(self.col_ind, dtype=np.uint32)
        self.row_ptr, self.col_ind = rp, ci
        if ci.size == 0:
            ci = self.col_ind = np.zeros(1, dtype=np.uint32)[:0]
        s = _BigGraph(self.n, self.m, rp.ctypes.data_as(C.POINTER(C.c_uint64)),
                      C.cast(C.c_void_p(ci.ctypes.data), C.POINTER(C.c_uint32)))
        return s

    @property
    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr.astype(np.int64))

    def wedges(self) -> int:
        s = self._struct()
        return int(_lib.wtc_wedge_total(C.byref(s)))

    def validate_oriented(self) -> bool:
        s = self._struct()
        return _lib.wtc_validate_oriented(C.byref(s)) == WTC_OK

    def alg_bytes(self, offset_bytes: int = 4):
        """(B_alg, B_min) of one count over this oriented CSR."""
        s = self._struct()
        a, b = C.c_double(), C.c_double()
        _lib.wtc_alg_bytes(C.byref(s), offset_bytes, C.byref(a), C.byref(b))
        return float(a.value), float(b.value)


def _adopt(ptr) -> CSR:
    g = ptr.contents
    owner = _Owner(ptr)
    n, m = int(g.numVertices), int(g.numEdges)
    rp = _view(C.cast(g.rowPtr, C.c_void_p).value, n + 1, C.c_uint64, owner)
    ci = _view(C.cast(g.colInd, C.c_void_p).value or 0, m, C.c_uint32, owner)
    return CSR(n, m, rp, ci)


def read_graph(path: str, matrix_market: bool = False, zero_indexed: bool = False,
               preprocess_style: int = 0) -> CSR:
    """read_graph of tc.cu:561-685.  Style 0 builds on the host; 1 and 2 sort
    on the GPU (same CSR).  The CLI's default is 2, like the reference's."""
    out = _BGP()
    _check(_lib.wtc_read_graph(path.encode(), int(matrix_market), int(zero_indexed), preprocess_style,
                               C.byref(out)), "read_graph")
    return _adopt(out)


def write_csr(path: str, g: CSR, tag: int = 0):
    """Binary CSR file with a caller tag and a checksum."""
    s = g._struct()
    _check(_lib.wtc_write_csr(path.encode(), C.byref(s), tag), "write_csr")


def read_csr(path: str):
    """(CSR, tag) from a binary CSR file; WtcError on any size, checksum or
    structure mismatch."""
    out = _BGP()
    tag = C.c_uint64(0)
    _check(_lib.wtc_read_csr(path.encode(), C.byref(tag), C.byref(out)), "read_csr")
    return _adopt(out), int(tag.value)


def graph_from_pairs(pairs: np.ndarray) -> CSR:
    """Symmetric CSR of (src,dst) pairs, as read_graph -z would build it."""
    p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    out = _BGP()
    _check(_lib.wtc_graph_from_pairs(p.ctypes.data if p.size else None, p.shape[0], C.byref(out)),
           "graph_from_pairs")
    return _adopt(out)


def preprocess(g: CSR, preprocess_style: int = 0) -> CSR:
    """preprocess of tc.cu:687-787 (degree order, orientation, sorted rows).
    Style 0 on the host; 1 and 2 on the GPU (same CSR)."""
    s = g._struct()
    out = _BGP()
    _check(_lib.wtc_preprocess(C.byref(s), preprocess_style, C.byref(out)), "preprocess")
    return _adopt(out)


def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1) -> CSR:
    out = _BGP()
    _check(_lib.wtc_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(out)), "gen_rmat")
    return _adopt(out)


def gen_rmat_pairs(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1) -> np.ndarray:
    p = C.POINTER(C.c_uint32)()
    n = C.c_uint64()
    _check(_lib.wtc_gen_rmat_pairs(scale, edge_factor, a, b, c, seed, C.byref(p), C.byref(n)), "gen_rmat_pairs")
    arr = np.ctypeslib.as_array(p, shape=(int(n.value) * 2,)).copy().reshape(-1, 2)
    _lib.wtc_free_pairs(p)
    return arr


def gen_chung_lu(n: int = 262144, exponent: float = 2.7, wmin: float = 1.0, wmax: float = 420.0,
                 target_edges: int = 900000, seed: int = 1) -> CSR:
    out = _BGP()
    _check(_lib.wtc_gen_chung_lu(n, exponent, wmin, wmax, target_edges, seed, C.byref(out)), "gen_chung_lu")
    return _adopt(out)


def gen_erdos_renyi(n: int, p: float, seed: int = 1) -> CSR:
    out = _BGP()
    _check(_lib.wtc_gen_erdos_renyi(n, p, seed, C.byref(out)), "gen_erdos_renyi")
    return _adopt(out)


def gen_barabasi_albert(n: int, m0: int, seed: int = 1) -> CSR:
    out = _BGP()
    _check(_lib.wtc_gen_barabasi_albert(n, m0, seed, C.byref(out)), "gen_barabasi_albert")
    return _adopt(out)


def autotune(n: int, m: int = 0, wedges: int = 0, batch: bool = False):
    """(spread, a) for a graph counted alone, or as a batch member."""
    s, a = C.c_uint32(), C.c_uint32()
    fn = _lib.wtc_autotune_batch if batch else _lib.wtc_autotune
    fn(n, m, wedges, C.byref(s), C.byref(a))
    return int(s.value), int(a.value)


class Tuned(C.Structure):
    """wtc_tuned: the autotuner's whole choice for one graph."""
    _fields_ = [("spread", C.c_uint32), ("adjacency_matrix_len", C.c_uint32), ("kernel", C.c_int32),
                ("waves_per_block", C.c_int32), ("probe_ways", C.c_int32), ("split_waves", C.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def autotune_shape(n: int, m: int = 0, wedges: int = 0, batch: bool = False) -> Tuned:
    """The tuning table's row for a graph: spread, a, kernel and block shape."""
    t = Tuned()
    _check(_lib.wtc_autotune_shape(n, m, wedges, int(batch), C.byref(t)), "autotune_shape")
    return t


def tuning_source() -> str:
    """Path of the tuning table in use, or 'built-in'."""
    return _lib.wtc_tuning_source().decode()


# ---------------------------------------------------------- reference-style

def tc_gpu(g: CSR, spread: int, a: int, t: Optional[GpuTime] = None) -> int:
    """tc_GPU of tc.cu:315 (32-bit offsets): per-call alloc/copy/count/free."""
    s = g._struct()
    g32 = C.POINTER(_Graph32)()
    _check(_lib.wtc_narrow_graph(C.byref(s), C.byref(g32)), "narrow_graph")
    try:
        cnt = C.c_uint64()
        _check(_lib.wtc_tc_gpu(g32, spread, a, C.byref(t) if t is not None else None, C.byref(cnt)), "tc_gpu")
        return int(cnt.value)
    finally:
        _lib.wtc_free_graph32(g32)


def tc_gpu_big(g: CSR, spread: int, a: int, t: Optional[GpuTime] = None) -> int:
    """tc_GPU of experiments/tc_big_graphs.cu:315 (64-bit offsets)."""
    s = g._struct()
    cnt = C.c_uint64()
    _check(_lib.wtc_tc_gpu_big(C.byref(s), spread, a, C.byref(t) if t is not None else None, C.byref(cnt)),
           "tc_gpu_big")
    return int(cnt.value)


def tc_arrow_gpu(g: CSR, spread: int, a: int, t: Optional[GpuTime] = None) -> int:
    """tc_arrow_GPU of experiments/tc_arrow.cu:311 (32-bit offsets; arrow wedges v->w->u)."""
    s = g._struct()
    g32 = C.POINTER(_Graph32)()
    _check(_lib.wtc_narrow_graph(C.byref(s), C.byref(g32)), "narrow_graph")
    try:
        cnt = C.c_uint64()
        _check(_lib.wtc_tc_arrow_gpu(g32, spread, a, C.byref(t) if t is not None else None, C.byref(cnt)),
               "tc_arrow_gpu")
        return int(cnt.value)
    finally:
        _lib.wtc_free_graph32(g32)


def tc_arrow_gpu_big(g: CSR, spread: int, a: int, t: Optional[GpuTime] = None) -> int:
    """tc_arrow_GPU of experiments/tc_big_graphs_arrow.cu:334 (64-bit offsets)."""
    s = g._struct()
    cnt = C.c_uint64()
    _check(_lib.wtc_tc_arrow_gpu_big(C.byref(s), spread, a, C.byref(t) if t is not None else None,
                                     C.byref(cnt)), "tc_arrow_gpu_big")
    return int(cnt.value)


WEDGE_OUTGOING, WEDGE_ARROW, WEDGE_MIXED = 0, 1, 2
PARALLEL_VERTEX, PARALLEL_EDGE = 0, 1
KERNEL_WALK, KERNEL_STAGED, KERNEL_SPLIT = 0, 1, 2


def _baseline_call(fn32, fn64, g: CSR, a: int, style: int, t: Optional[GpuTime], big: bool) -> int:
    s = g._struct()
    cnt = C.c_uint64()
    tp = C.byref(t) if t is not None else None
    if big:
        _check(fn64(C.byref(s), a, style, tp, C.byref(cnt)), fn64.__name__)
        return int(cnt.value)
    g32 = C.POINTER(_Graph32)()
    _check(_lib.wtc_narrow_graph(C.byref(s), C.byref(g32)), "narrow_graph")
    try:
        _check(fn32(g32, a, style, tp, C.byref(cnt)), fn32.__name__)
        return int(cnt.value)
    finally:
        _lib.wtc_free_graph32(g32)


def tc_vertex_gpu(g: CSR, a: int, style: int = WEDGE_OUTGOING, t: Optional[GpuTime] = None,
                  big: bool = False) -> int:
    """tc_vertex_GPU of experiments/tc_vertex.cu:287 (big: tc_big_graphs_vertex.cu)."""
    return _baseline_call(_lib.wtc_tc_vertex_gpu, _lib.wtc_tc_vertex_gpu_big, g, a, style, t, big)


def tc_edge_gpu(g: CSR, a: int, style: int = WEDGE_OUTGOING, t: Optional[GpuTime] = None,
                big: bool = False) -> int:
    """tc_edge_GPU of experiments/tc_edge.cu:374 (big: tc_big_graphs_edge.cu)."""
    return _baseline_call(_lib.wtc_tc_edge_gpu, _lib.wtc_tc_edge_gpu_big, g, a, style, t, big)


def device_count() -> int:
    c = C.c_int()
    _lib.wtc_device_count(C.byref(c))
    return int(c.value)


# -------------------------------------------------------- device-resident

class Context:
    """A HIP stream plus reusable device workspace on one GPU."""

    def __init__(self, device: int = 0):
        self._p = C.c_void_p()
        _check(_lib.wtc_ctx_create(device, C.byref(self._p)), "ctx_create")

    def close(self):
        if self._p:
            _lib.wtc_ctx_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, g: CSR, offset_bytes: int = 0) -> "DeviceGraph":
        return DeviceGraph(self, g, offset_bytes)

    def debug_prefix(self, n: int):
        ws = np.zeros(n + 1, dtype=np.uint64)
        tree = np.zeros(4095, dtype=np.uint64)
        _check(_lib.wtc_ctx_debug_prefix(self._p, n, ws.ctypes.data, tree.ctypes.data), "debug_prefix")
        return ws, tree

    def set_tuning(self, probe_ways: int = 4, lds_chunk: int = 16, waves_per_block: int = 4):
        """Counting-kernel knobs: probes advanced together per lane, wedges
        per lane staged in LDS per phase, waves per block (1, 2, 4 or 8)."""
        _check(_lib.wtc_ctx_set_tuning(self._p, probe_ways, lds_chunk, waves_per_block), "set_tuning")

    def set_count_kernel(self, kernel: int):
        """KERNEL_WALK (0), KERNEL_STAGED (1) or KERNEL_SPLIT (2: a single graph
        with a > 0 counts its top-a wedges by middle vertex against an LDS
        bitmap row, the rest by the walk); counts do not depend on it."""
        _check(_lib.wtc_ctx_set_count_kernel(self._p, kernel), "set_count_kernel")

    def use_tuned(self, t: "Tuned"):
        """Applies an autotuned shape (kernel, block waves, probe ways, row-kernel waves)."""
        _check(_lib.wtc_ctx_use_tuned(self._p, C.byref(t)), "use_tuned")

    def set_work_split(self, tiles_per_grab: int = 0, deal_to_xcds: bool = False):
        """Most tiles per grab (0: the kernel's default, 4 for the walk and 32
        for the staged kernel) and XCD dealing of the counting kernels
        (default: the kernel's grab, one stream of grabs)."""
        _check(_lib.wtc_ctx_set_work_split(self._p, tiles_per_grab, int(deal_to_xcds)), "set_work_split")

    def set_small_graphs(self, max_words: int):
        """In counts over two or more graphs, graphs with 2(n+1) + m <=
        max_words are counted from LDS by the small-graph kernel (0: off)."""
        _check(_lib.wtc_ctx_set_small_graphs(self._p, max_words), "set_small_graphs")

    def small_graphs(self):
        """(current small-graph limit, device maximum), in LDS words."""
        cur, dev = C.c_uint32(), C.c_uint32()
        _check(_lib.wtc_ctx_get_small_graphs(self._p, C.byref(cur), C.byref(dev)), "get_small_graphs")
        return int(cur.value), int(dev.value)

    def debug_copy(self, nbytes: int, reps: int = 1) -> float:
        ms = C.c_double()
        _check(_lib.wtc_debug_copy(self._p, nbytes, reps, C.byref(ms)), "debug_copy")
        return float(ms.value)

    def debug_decode(self, d, iv):
        d = np.ascontiguousarray(d, dtype=np.uint32)
        iv = np.ascontiguousarray(iv, dtype=np.uint64)
        w = np.zeros(d.size, dtype=np.uint32)
        u = np.zeros(d.size, dtype=np.uint32)
        _check(_lib.wtc_debug_decode(self._p, d.ctypes.data, iv.ctypes.data, d.size, w.ctypes.data, u.ctypes.data),
               "debug_decode")
        return w, u

    def debug_bitmap(self, words: int) -> np.ndarray:
        out = np.zeros(max(words, 1), dtype=np.uint32)
        _check(_lib.wtc_ctx_debug_bitmap(self._p, words, out.ctypes.data), "debug_bitmap")
        return out[:words]


class DeviceGraph:
    """An oriented CSR resident in HBM."""

    def __init__(self, ctx: Context, g: Optional[CSR], offset_bytes: int = 0, _handle=None):
        self.ctx = ctx
        self._p = C.c_void_p()
        if _handle is not None:
            self._p = _handle
        else:
            s = g._struct()
            _check(_lib.wtc_dgraph_upload(ctx._p, C.byref(s), offset_bytes, C.byref(self._p)), "dgraph_upload")
        n, m, ob = C.c_uint32(), C.c_uint64(), C.c_uint32()
        _check(_lib.wtc_dgraph_info(self._p, C.byref(n), C.byref(m), C.byref(ob)), "dgraph_info")
        self.n, self.m, self.offset_bytes = int(n.value), int(m.value), int(ob.value)

    @classmethod
    def from_pairs(cls, ctx: Context, pairs: np.ndarray, offset_bytes: int = 0) -> "DeviceGraph":
        """Oriented CSR built on the GPU from (src,dst) pairs (as read_graph -z + preprocess)."""
        p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        h = C.c_void_p()
        _check(_lib.wtc_dgraph_from_pairs(ctx._p, p.ctypes.data if p.size else None, p.shape[0], offset_bytes,
                                          C.byref(h)), "dgraph_from_pairs")
        return cls(ctx, None, _handle=h)

    @classmethod
    def gen_rmat(cls, ctx: Context, scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1,
                 offset_bytes: int = 0) -> "DeviceGraph":
        """The graph of gen_rmat + preprocess, generated and oriented on the GPU."""
        h = C.c_void_p()
        _check(_lib.wtc_dgraph_gen_rmat(ctx._p, scale, edge_factor, a, b, c, seed, offset_bytes, C.byref(h)),
               "dgraph_gen_rmat")
        return cls(ctx, None, _handle=h)

    def download(self) -> CSR:
        out = _BGP()
        _check(_lib.wtc_dgraph_download(self._p, C.byref(out)), "dgraph_download")
        return _adopt(out)

    @property
    def nbytes(self) -> int:
        return int(_lib.wtc_dgraph_bytes(self._p))

    def count(self, spread: int, a: int, w_begin: int = 0, w_end: int = 2 ** 64 - 1, stats: bool = False):
        cnt = C.c_uint64()
        st = CountStats()
        _check(_lib.wtc_dgraph_count(self.ctx._p, self._p, spread, a, w_begin, w_end, C.byref(cnt), C.byref(st)),
               "dgraph_count")
        return (int(cnt.value), st) if stats else int(cnt.value)

    def count_baseline(self, parallelism: int, style: int, a: int, stats: bool = False):
        """Vertex- (0) or edge-parallel (1) count in wedge style 0/1/2."""
        cnt = C.c_uint64()
        st = CountStats()
        _check(_lib.wtc_dgraph_count_baseline(self.ctx._p, self._p, parallelism, style, a, C.byref(cnt),
                                              C.byref(st)), "dgraph_count_baseline")
        return (int(cnt.value), st) if stats else int(cnt.value)

    def count_arrow(self, spread: int, a: int, stats: bool = False):
        """Arrow-wedge count: (triangles, arrow wedges[, stats])."""
        cnt, aw = C.c_uint64(), C.c_uint64()
        st = CountStats()
        _check(_lib.wtc_dgraph_count_arrow(self.ctx._p, self._p, spread, a, C.byref(cnt), C.byref(aw),
                                           C.byref(st)), "dgraph_count_arrow")
        return (int(cnt.value), int(aw.value), st) if stats else (int(cnt.value), int(aw.value))

    def free(self):
        if self._p:
            _lib.wtc_dgraph_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _graph_array(graphs: Sequence[CSR]):
    structs = [g._struct() for g in graphs]
    arr = (_BGP * len(structs))(*[C.pointer(s) for s in structs])
    return arr, structs


def _u32_array(vals, k):
    if vals is None:
        return None, None
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(vals, dtype=np.uint32), (k,)))
    return a.ctypes.data, a


class Batch:
    """Many graphs resident in HBM as one disjoint union, counted by one launch."""

    def __init__(self, ctx: Context, graphs: Sequence[CSR]):
        self.ctx = ctx
        self.k = len(graphs)
        self._p = C.c_void_p()
        arr, keep = _graph_array(graphs)
        _check(_lib.wtc_batch_upload(ctx._p, arr, self.k, C.byref(self._p)), "batch_upload")

    def count(self, spread=None, a=None, stats: bool = False):
        out = np.zeros(self.k, dtype=np.uint64)
        sp, ks = _u32_array(spread, self.k)
        ap, ka = _u32_array(a, self.k)
        st = CountStats()
        _check(_lib.wtc_batch_count(self.ctx._p, self._p, sp, ap, out.ctypes.data, C.byref(st)), "batch_count")
        return (out, st) if stats else out

    def free(self):
        if self._p:
            _lib.wtc_batch_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def batch_count_host(ctx: Context, graphs: Sequence[CSR], spread=None, a=None, stats: bool = False):
    """Host CSRs in, per-graph counts out, in one batched launch."""
    k = len(graphs)
    out = np.zeros(k, dtype=np.uint64)
    arr, keep = _graph_array(graphs)
    sp, ks = _u32_array(spread, k)
    ap, ka = _u32_array(a, k)
    st = CountStats()
    _check(_lib.wtc_batch_count_host(ctx._p, arr, k, sp, ap, out.ctypes.data, C.byref(st)), "batch_count_host")
    return (out, st) if stats else out


# ------------------------------------------------------------------ serving

class Server:
    """Request loop with continuous batching over device-resident CSRs."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 8 << 30, max_batch: int = 256, slots: int = 2,
                 queue_capacity: int = 65536, probe_ways: int = 0, lds_chunk: int = 0, waves_per_block: int = 0,
                 small_graph_words: int = 0):
        cfg = ServerConfig(hbm_budget_bytes, max_batch, slots, queue_capacity, probe_ways, lds_chunk, waves_per_block,
                           small_graph_words)
        self._p = C.c_void_p()
        _check(_lib.wtc_server_create(device, C.byref(cfg), C.byref(self._p)), "server_create")

    def register(self, g: CSR) -> int:
        s = g._struct()
        gid = C.c_uint64()
        _check(_lib.wtc_server_register(self._p, C.byref(s), C.byref(gid)), "server_register")
        return int(gid.value)

    def submit(self, graph_id: int, spread: int, a: int) -> int:
        """Returns a request id; raises WtcError (code -11 when the queue is full)."""
        rid = C.c_uint64()
        _check(_lib.wtc_server_submit(self._p, graph_id, spread, a, C.byref(rid)), "server_submit")
        return int(rid.value)

    def wait(self, request_id: int, timeout_us: int = -1):
        """(triangles, RequestStats) once the request has finished."""
        cnt = C.c_uint64()
        st = RequestStats()
        _check(_lib.wtc_server_wait(self._p, request_id, timeout_us, C.byref(cnt), C.byref(st)), "server_wait")
        return int(cnt.value), st

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib.wtc_server_get_stats(self._p, C.byref(st)), "server_stats")
        return st.as_dict()

    def close(self):
        if self._p:
            _lib.wtc_server_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
