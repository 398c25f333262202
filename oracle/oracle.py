"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, only to check results against (and to time the
CPU baseline), never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
EXE = os.path.join(HERE, "build", "clast_oracle")

FINAL, RAW, BOUNDARY, SORTED, ISOLATE, DUPLICATE, ALIGNED, FILTERED = range(8)
FIELDS = ("tID", "tIdx", "qID", "qIdx", "tHL", "qHL", "match", "score")


class OParams(C.Structure):
    _fields_ = [("lmer", C.c_int32), ("stride", C.c_int32), ("cut_repeat", C.c_int32), ("width", C.c_int32),
                ("gap", C.c_int32), ("local", C.c_int32), ("cut_off", C.c_double), ("total_db_size", C.c_double),
                ("K", C.c_double), ("lambda_", C.c_double)]


class OList(C.Structure):
    _fields_ = [("base", C.c_void_p), ("n_slots", C.c_int64), ("length", C.c_void_p), ("gateway", C.c_void_p),
                ("n_seq", C.c_int32)]


_lib = None


def use_build(dirname: str) -> str:
    """Load the oracle from oracle/<dirname>/ (e.g. build_native, made by
    `make -C oracle native`) instead of oracle/build/; returns the path."""
    global LIB, EXE, _lib
    LIB = os.path.join(HERE, dirname, "liboracle.so")
    EXE = os.path.join(HERE, dirname, "clast_oracle")
    _lib = None
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.oracle_kmer_code.restype = C.c_int64
        L.oracle_kmer_code.argtypes = [C.c_char_p, C.c_int64, C.c_int32]
        L.oracle_search.restype = vp
        L.oracle_search.argtypes = [C.POINTER(OParams), C.POINTER(OList), vp, C.c_int32, C.POINTER(OList), vp,
                                    C.c_int32, C.c_int32]
        L.oracle_count.restype = C.c_int64
        L.oracle_count.argtypes = [vp, C.c_int32]
        L.oracle_copy.restype = C.c_int32
        L.oracle_copy.argtypes = [vp, C.c_int32] + [vp] * 10
        L.oracle_free.argtypes = [vp]
        L.oracle_times.argtypes = [vp, vp]
        L.oracle_index_build.restype = vp
        L.oracle_index_build.argtypes = [C.POINTER(OParams), C.POINTER(OList), C.c_int32, C.c_int32]
        L.oracle_index_cells.restype = C.c_int64
        L.oracle_index_cells.argtypes = [vp]
        L.oracle_index_size.restype = C.c_int64
        L.oracle_index_size.argtypes = [vp]
        L.oracle_index_copy.argtypes = [vp, vp, vp, vp, vp]
        L.oracle_index_free.argtypes = [vp]
        L.oracle_align_one.argtypes = [C.POINTER(OParams), C.POINTER(OList), C.c_int32, C.c_int32, C.POINTER(OList),
                                       C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp]
        L.oracle_total_db_size.restype = C.c_int64
        L.oracle_total_db_size.argtypes = [C.POINTER(C.c_char_p), C.c_int32]
        _lib = L
    return _lib


def _params(p) -> OParams:
    return OParams(p.lmer, p.stride, p.cut_repeat, p.width, p.gap, p.local, p.cut_off, p.total_db_size, p.K, p.lam)


def _olist(lst) -> OList:
    return OList(lst.base.ctypes.data, len(lst.base), lst.length.ctypes.data, lst.gateway.ctypes.data, lst.n_seq)


def kmer_code(s: bytes, lmer: int, avail: int | None = None) -> int:
    return int(lib().oracle_kmer_code(s, len(s) if avail is None else avail, lmer))


def _pairs(wins) -> np.ndarray:
    return np.ascontiguousarray(np.array(wins, dtype=np.int32).reshape(-1))


def search(params, tlist, twins, qlist, qwins, dump: bool = False) -> dict:
    """Every (target window, query window) unit, appended in order and put in
    fixResult order; with dump (one unit) also the intermediate stages."""
    L = lib()
    tw, qw = _pairs(twins), _pairs(qwins)
    tl, ql = _olist(tlist), _olist(qlist)
    h = L.oracle_search(C.byref(_params(params)), C.byref(tl), tw.ctypes.data, len(twins), C.byref(ql),
                        qw.ctypes.data, len(qwins), int(dump))
    try:
        out = {"final": _copy(L, h, FINAL, hits=True)}
        tm = np.zeros(2, np.float64)
        L.oracle_times(h, tm.ctypes.data)
        out["index_s"], out["search_s"] = float(tm[0]), float(tm[1])
        if dump:
            for name, st in (("raw", RAW), ("boundary", BOUNDARY), ("sorted", SORTED), ("isolate", ISOLATE),
                             ("duplicate", DUPLICATE)):
                out[name] = _copy(L, h, st, hits=False)
            out["aligned"] = _copy(L, h, ALIGNED, hits=True)
            out["filtered"] = _copy(L, h, FILTERED, hits=True)
    finally:
        L.oracle_free(h)
    return out


def _copy(L, h, stage: int, hits: bool) -> dict:
    n = int(L.oracle_count(h, stage))
    names = FIELDS if hits else FIELDS[:4]
    out = {k: np.empty(n, np.int32) for k in names}
    ptrs = [out[k].ctypes.data if k in out else None for k in FIELDS]
    ev = None
    tw = None
    if hits:
        out["evalue"] = np.empty(n, np.float64)
        ev = out["evalue"].ctypes.data
        if stage == FINAL:
            out["twin"] = np.empty(n, np.int32)
            tw = out["twin"].ctypes.data
    if n:
        L.oracle_copy(h, stage, *ptrs, ev, tw)
    return out


def index(params, tlist, t_begin: int, t_end: int) -> dict:
    """The reference index of one target window: gateway keys, cell sizes
    (after the repeat cut), gateway index (offsets), sorted positions."""
    L = lib()
    tl = _olist(tlist)
    h = L.oracle_index_build(C.byref(_params(params)), C.byref(tl), t_begin, t_end)
    try:
        nc, n = int(L.oracle_index_cells(h)), int(L.oracle_index_size(h))
        out = {"key": np.empty(nc, np.int64), "cell_size": np.empty(nc, np.int32),
               "gateway_index": np.empty(nc, np.int32), "positions": np.empty(n, np.int32)}
        L.oracle_index_copy(h, out["key"].ctypes.data, out["cell_size"].ctypes.data, out["gateway_index"].ctypes.data,
                            out["positions"].ctypes.data)
    finally:
        L.oracle_index_free(h)
    return out


def align_one(params, tlist, twin, qlist, qwin, tID, tIdx, qID, qIdx) -> tuple:
    """(tIdx, qIdx, tHL, qHL, match, score) of one seed's two DP extensions."""
    L = lib()
    out = np.zeros(6, np.int32)
    tl, ql = _olist(tlist), _olist(qlist)
    L.oracle_align_one(C.byref(_params(params)), C.byref(tl), twin[0], twin[1], C.byref(ql), qwin[0], qwin[1], tID,
                       tIdx, qID, qIdx, out.ctypes.data)
    return tuple(int(x) for x in out)


def total_db_size(files) -> int:
    arr = (C.c_char_p * len(files))(*[f.encode() for f in files])
    return int(lib().oracle_total_db_size(arr, len(files)))
