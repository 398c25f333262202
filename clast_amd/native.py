"""ctypes bindings of libclast_hip (include/clast_hip.h) and libclast_host.

The product path: every search here runs the gfx950 kernels of
``libclast_hip.so``; there is no CPU fallback.  If the library is missing the
loaders raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")

CLAST_KEEP_STAGES = 1
STAGE_SEEDS, STAGE_SORTED, STAGE_HITS, STAGE_ALIGNED = 1, 2, 3, 4

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


class CParams(C.Structure):
    _fields_ = [("lmer", C.c_int32), ("stride", C.c_int32), ("cut_repeat", C.c_int32), ("width", C.c_int32),
                ("gap", C.c_int32), ("local", C.c_int32), ("cut_off", C.c_double), ("total_db_size", C.c_double),
                ("K", C.c_double), ("lambda_", C.c_double)]


class CWindow(C.Structure):
    _fields_ = [("bases", C.c_void_p), ("n_slots", C.c_int64), ("lengths", C.c_void_p), ("gateway", C.c_void_p),
                ("n_seq", C.c_int32), ("id_begin", C.c_int32)]


@dataclass
class Params:
    """CHostSetting's search parameters with the reference defaults
    (common.hpp:12-43)."""
    lmer: int = 15
    stride: int = 5
    cut_repeat: int = 20
    width: int = 100
    gap: int = 8
    local: int = 0
    cut_off: float = 10.0
    total_db_size: float = 0.0
    K: float = 0.71
    lam: float = 1.37

    def c(self) -> CParams:
        return CParams(self.lmer, self.stride, self.cut_repeat, self.width, self.gap, self.local, self.cut_off,
                       self.total_db_size, self.K, self.lam)


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


_hip = None
_host = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        L = _load("libclast_hip.so")
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        pp = C.POINTER(CParams)
        wp = C.POINTER(CWindow)
        sig = {
            "clast_init": (C.c_int, [i32]),
            "clast_last_error": (C.c_char_p, []),
            "clast_version": (C.c_char_p, []),
            "clast_index_create": (C.c_int, [pp, wp, C.POINTER(vp)]),
            "clast_index_create_multi": (C.c_int, [pp, wp, i32, C.POINTER(vp)]),
            "clast_index_entry_count": (i64, [vp]),
            "clast_index_entries": (C.c_int, [vp, vp, vp, vp, vp]),
            "clast_index_destroy": (C.c_int, [vp]),
            "clast_index_save": (C.c_int, [vp, C.c_char_p]),
            "clast_index_load": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
            "clast_query_create": (C.c_int, [pp, wp, C.POINTER(vp)]),
            "clast_query_create_multi": (C.c_int, [pp, wp, i32, C.POINTER(vp)]),
            "clast_query_create_reads": (C.c_int, [pp, vp, vp, vp, vp, i32, C.POINTER(vp)]),
            "clast_query_destroy": (C.c_int, [vp]),
            "clast_search": (C.c_int, [pp, vp, vp, i32, i32, C.POINTER(vp)]),
            "clast_search_batch": (C.c_int, [pp, vp, vp, i32, C.POINTER(vp)]),
            "clast_search_batch_noalloc": (C.c_int, [pp, vp, vp, i64] + [vp] * 12),
            "clast_hits_count": (i64, [vp]),
            "clast_hits_wait": (C.c_int, [vp]),
            "clast_hits_counts": (C.c_int, [vp, vp]),
            "clast_hits_copy": (C.c_int, [vp] + [vp] * 9),
            "clast_hits_copy_windows": (C.c_int, [vp, vp, vp]),
            "clast_hits_stage_count": (i64, [vp, i32]),
            "clast_hits_stage_copy": (C.c_int, [vp, i32] + [vp] * 8),
            "clast_hits_stage_times": (C.c_int, [vp, vp]),
            "clast_hits_kernel_times": (C.c_int, [vp, vp]),
            "clast_hits_dp_kernels": (C.c_int, [vp, vp]),
            "clast_hits_dp_config": (C.c_int, [vp, vp]),
            "clast_tuner_report": (i64, [C.c_char_p, i64]),
            "clast_mem_info": (C.c_int, [vp, vp]),
            "clast_workspace_bytes": (C.c_int, [vp]),
            "clast_hits_destroy": (C.c_int, [vp]),
            "clast_result_sort": (C.c_int, [i64] + [vp] * 9),
            "clast_result_finalize": (C.c_int, [i64] + [vp] * 11 + [i32, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _hip = L
    return _hip


def host() -> C.CDLL:
    global _host
    if _host is None:
        hip()
        L = _load("libclast_host.so")
        L.clast_host_list_slots.restype = C.c_int64
        L.clast_host_list_slots.argtypes = [_i64p, C.c_int32, C.c_int32, C.c_int32]
        L.clast_host_build_list.restype = C.c_int32
        L.clast_host_build_list.argtypes = [_u8p, _i64p, C.c_int32, C.c_int32, C.c_int32, _u8p, _i32p, _i32p]
        L.clast_host_windows.restype = C.c_int32
        L.clast_host_windows.argtypes = [_i32p, C.c_int32, C.c_double, C.c_int32, _i32p]
        L.clast_host_chunks.restype = C.c_int32
        L.clast_host_chunks.argtypes = [_i64p, C.c_int32, C.c_int64, _i32p, _i8p]
        _host = L
    return _host


class ClastError(RuntimeError):
    pass


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ClastError(f"{what} failed ({rc}): {hip().clast_last_error().decode()}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def init(device: int = 0) -> None:
    _check(hip().clast_init(device), "clast_init")


def mem_info() -> tuple[int, int]:
    """(free, total) device bytes of the calling thread's device."""
    f, t = C.c_int64(0), C.c_int64(0)
    _check(hip().clast_mem_info(C.byref(f), C.byref(t)), "clast_mem_info")
    return int(f.value), int(t.value)


def tuner_report() -> str:
    """The DP autotuner's table: one line per kind of search tuned."""
    n = int(hip().clast_tuner_report(None, 0))
    buf = C.create_string_buffer(n + 1)
    hip().clast_tuner_report(buf, n + 1)
    return buf.value.decode()


def workspace_bytes() -> int:
    """Device bytes of the calling thread's search workspace."""
    b = C.c_int64(0)
    _check(hip().clast_workspace_bytes(C.byref(b)), "clast_workspace_bytes")
    return int(b.value)


# ---------------------------------------------------------------- host lists

@dataclass
class HostList:
    """CHostSeqList layout: slot array (sequence + lMer-1 joint each), lengths
    and the n_seq+1 gateway prefix sums."""
    base: np.ndarray
    length: np.ndarray
    gateway: np.ndarray
    lmer: int
    query: bool = False
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_seq(self) -> int:
        return int(self.length.shape[0])

    def window(self, a: int, b: int) -> CWindow:
        """CHostSeqList::get*Array(a, b) as a clast_window (sequence IDs)."""
        lo = int(self.gateway[a])
        gw = np.ascontiguousarray(self.gateway[a:b] - lo, dtype=np.int32)
        ln = np.ascontiguousarray(self.length[a:b], dtype=np.int32)
        self._keep.extend([gw, ln])
        w = CWindow()
        w.bases = _ptr(self.base) + lo
        w.n_slots = int(self.gateway[b]) - lo
        w.lengths = _ptr(ln)
        w.gateway = _ptr(gw)
        w.n_seq = b - a
        w.id_begin = a
        return w

    def release_windows(self) -> None:
        self._keep.clear()


def build_list(seqs: np.ndarray, offsets: np.ndarray, lmer: int, query: bool) -> HostList:
    """Builds the CHostSeqList layout of concatenated sequences (queries add
    the reverse-complement strand after each one)."""
    L = host()
    seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    n = int(offsets.shape[0]) - 1
    slots = int(L.clast_host_list_slots(offsets, n, lmer, int(query)))
    if slots >= 2 ** 31:
        raise ClastError("a host list is limited to 2^31 slots (int32 indexes); chunk the input")
    nseq = 2 * n if query else n
    base = np.empty(max(slots, 1), dtype=np.uint8)
    length = np.empty(max(nseq, 1), dtype=np.int32)
    gateway = np.empty(nseq + 1, dtype=np.int32)
    L.clast_host_build_list(seqs, offsets, n, lmer, int(query), base, length, gateway)
    return HostList(base[:slots] if slots else base[:0], length[:nseq], gateway, lmer, query)


def windows(gateway: np.ndarray, vram_slots: float, query: bool) -> list[tuple[int, int]]:
    """moveTargetWindow / moveQueryWindow (query windows in pair units)."""
    g = np.ascontiguousarray(gateway, dtype=np.int32)
    out = np.empty(2 * max(len(g), 1), dtype=np.int32)
    k = host().clast_host_windows(g, len(g), float(vram_slots), int(query), out)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(k)]


def read_gateway(lengths: np.ndarray, lmer: int) -> np.ndarray:
    """Slot gateway of the '+'/'-' strands of reads of the given lengths
    (each strand followed by lMer-1 joint slots), as a query host list
    would hold it: 2 n + 1 entries."""
    span = np.repeat(np.asarray(lengths, dtype=np.int64) + max(lmer - 1, 0), 2)
    g = np.zeros(len(span) + 1, dtype=np.int64)
    np.cumsum(span, out=g[1:])
    if g[-1] >= 2**31:
        raise ValueError("query set over 2^31 slots")
    return g.astype(np.int32)


def chunks(lengths: np.ndarray, avail: int) -> tuple[list[tuple[int, int]], np.ndarray]:
    """The FASTA loader's chunk boundaries for records of these lengths."""
    ln = np.ascontiguousarray(lengths, dtype=np.int64)
    out = np.empty(2 * (len(ln) + 2), dtype=np.int32)
    keep = np.zeros(max(len(ln), 1), dtype=np.int8)
    k = host().clast_host_chunks(ln, len(ln), int(avail), out, keep)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(k)], keep[:len(ln)].astype(bool)


# ---------------------------------------------------------------- handles

class Index:
    """A resident target index (CDeviceHashTable) over one or more windows."""

    def __init__(self, params: Params, lst: HostList, wins: list[tuple[int, int]]):
        self.params = params
        self.wins = list(wins)
        arr = (CWindow * len(wins))(*[lst.window(a, b) for a, b in wins])
        h = C.c_void_p()
        _check(hip().clast_index_create_multi(C.byref(params.c()), arr, len(wins), C.byref(h)),
               "clast_index_create_multi")
        lst.release_windows()
        self.h = h

    @classmethod
    def merged(cls, params: Params, lists_and_windows) -> "Index":
        """One index over the windows of several host lists (e.g. every
        target RAM chunk), numbered in the given (list, window) order."""
        self = cls.__new__(cls)
        self.params = params
        self.wins = []
        ws = []
        for lst, wins in lists_and_windows:
            ws.extend(lst.window(a, b) for a, b in wins)
            self.wins.extend(wins)
        arr = (CWindow * len(ws))(*ws)
        h = C.c_void_p()
        _check(hip().clast_index_create_multi(C.byref(params.c()), arr, len(ws), C.byref(h)),
               "clast_index_create_multi")
        for lst, _ in lists_and_windows:
            lst.release_windows()
        self.h = h
        return self

    def save(self, path: str) -> None:
        """Writes the index to a file (clast_index_save)."""
        _check(hip().clast_index_save(self.h, os.fsencode(path)), "clast_index_save")

    @classmethod
    def load(cls, params: Params, path: str) -> "Index":
        """An index read back from a file written by save (clast_index_load);
        params must be the ones it was built with."""
        self = cls.__new__(cls)
        self.params = params
        self.wins = []
        h = C.c_void_p()
        _check(hip().clast_index_load(os.fsencode(path), C.byref(h)), "clast_index_load")
        self.h = h
        return self

    def entries(self) -> dict[str, np.ndarray]:
        n = int(hip().clast_index_entry_count(self.h))
        out = {"key": np.empty(n, np.int64), "window": np.empty(n, np.int32), "tID": np.empty(n, np.int32),
               "tIdx": np.empty(n, np.int32)}
        if n:
            _check(hip().clast_index_entries(self.h, _ptr(out["key"]), _ptr(out["window"]), _ptr(out["tID"]),
                                             _ptr(out["tIdx"])), "clast_index_entries")
        return out

    def close(self) -> None:
        if self.h:
            hip().clast_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Query:
    """Uploaded query windows (CDeviceSeqList_query); wins are pair ranges."""

    def __init__(self, params: Params, lst: HostList, wins: list[tuple[int, int]]):
        self.wins = list(wins)
        arr = (CWindow * len(wins))(*[lst.window(2 * a, 2 * b) for a, b in wins])
        h = C.c_void_p()
        _check(hip().clast_query_create_multi(C.byref(params.c()), arr, len(wins), C.byref(h)),
               "clast_query_create_multi")
        lst.release_windows()
        self.h = h

    @classmethod
    def merged(cls, params: Params, lists_and_windows) -> "Query":
        """One query set over the pair windows of several host lists (e.g.
        several requests), numbered in the given (list, window) order."""
        self = cls.__new__(cls)
        self.wins = []
        ws = []
        for lst, wins in lists_and_windows:
            ws.extend(lst.window(2 * a, 2 * b) for a, b in wins)
            self.wins.extend(wins)
        arr = (CWindow * len(ws))(*ws)
        h = C.c_void_p()
        _check(hip().clast_query_create_multi(C.byref(params.c()), arr, len(ws), C.byref(h)),
               "clast_query_create_multi")
        for lst, _ in lists_and_windows:
            lst.release_windows()
        self.h = h
        return self

    @classmethod
    def from_reads(cls, params: Params, bases: np.ndarray, offsets: np.ndarray, wins: list[tuple[int, int]],
                   id_begin: list[int] | None = None) -> "Query":
        """Query windows built on the device from the reads as given
        (clast_query_create_reads): wins are consecutive read (pair) ranges;
        window w's first strand has list ID id_begin[w] (default 2 x its
        first read)."""
        self = cls.__new__(cls)
        self.wins = list(wins)
        b = np.ascontiguousarray(bases, dtype=np.uint8)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        w = np.ascontiguousarray(np.asarray(wins, dtype=np.int32).reshape(-1))
        ids = np.ascontiguousarray(np.asarray(id_begin if id_begin is not None else [2 * a for a, _ in wins],
                                              dtype=np.int32))
        h = C.c_void_p()
        _check(hip().clast_query_create_reads(C.byref(params.c()), _ptr(b), _ptr(off), _ptr(w), _ptr(ids), len(wins),
                                              C.byref(h)), "clast_query_create_reads")
        self.h = h
        return self

    def close(self) -> None:
        if self.h:
            hip().clast_query_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HIT_FIELDS = ("tID", "tIdx", "qID", "qIdx", "tHL", "qHL", "match", "score")


class Hits:
    """Result rows of a search (CDeviceHitList::getResult)."""

    def __init__(self, h: C.c_void_p):
        self.h = h

    def __len__(self) -> int:
        return int(hip().clast_hits_count(self.h))

    def wait(self) -> None:
        """Blocks until the rows are in host memory (they arrive after the
        search returns, overlapping whatever the caller launches next)."""
        _check(hip().clast_hits_wait(self.h), "clast_hits_wait")

    def arrays(self) -> dict[str, np.ndarray]:
        n = len(self)
        out = {k: np.empty(n, np.int32) for k in HIT_FIELDS}
        out["evalue"] = np.empty(n, np.float64)
        out["twin"] = np.empty(n, np.int32)
        out["qwin"] = np.empty(n, np.int32)
        if n:
            _check(hip().clast_hits_copy(self.h, *[_ptr(out[k]) for k in HIT_FIELDS], _ptr(out["evalue"])),
                   "clast_hits_copy")
            _check(hip().clast_hits_copy_windows(self.h, _ptr(out["twin"]), _ptr(out["qwin"])),
                   "clast_hits_copy_windows")
        return out

    def stage(self, stage: int) -> dict[str, np.ndarray]:
        n = int(hip().clast_hits_stage_count(self.h, stage))
        out = {k: np.empty(max(n, 0), np.int32) for k in HIT_FIELDS}
        if n > 0:
            _check(hip().clast_hits_stage_copy(self.h, stage, *[_ptr(out[k]) for k in HIT_FIELDS]),
                   "clast_hits_stage_copy")
        if stage != STAGE_ALIGNED:
            for k in HIT_FIELDS[4:]:
                del out[k]
        return out

    def counts(self) -> np.ndarray:
        """raw seeds, hits entering the DP, DP cells, result rows, DP tasks
        run in packed 16-bit lanes, their cells, DP tasks run one lane per
        band, their cells."""
        c = np.zeros(8, np.int64)
        _check(hip().clast_hits_counts(self.h, _ptr(c)), "clast_hits_counts")
        return c

    def times(self) -> np.ndarray:
        """seed count, seed write, seed sort, filters, DP, results, total (ms), seeds."""
        t = np.zeros(8, np.float32)
        _check(hip().clast_hits_stage_times(self.h, _ptr(t)), "clast_hits_stage_times")
        return t

    def kernel_times(self) -> np.ndarray:
        """k_seed_count, k_dp16, k_dp, k_dp_band, DP stage, seed generation,
        results, whole search (ms)."""
        t = np.zeros(8, np.float32)
        _check(hip().clast_hits_kernel_times(self.h, _ptr(t)), "clast_hits_kernel_times")
        return t

    def dp_kernels(self) -> np.ndarray:
        """DP cells of k_dp16, of its past-the-end launch, of k_dp and of
        k_dp_band, the same four kernels' device times (ms), then the second
        DP pass's tasks, cells and time (ms), and 1.0 for a two-pass stage
        (whose first-pass 16-bit kernels computed scores only)."""
        d = np.zeros(12, np.float64)
        _check(hip().clast_hits_dp_kernels(self.h, _ptr(d)), "clast_hits_dp_kernels")
        return d

    def dp_config(self) -> tuple[int, int, int]:
        """(16-bit schedule, lane-per-band mode, radix digit bits) the
        search's DP ran with."""
        c = np.zeros(3, np.int32)
        _check(hip().clast_hits_dp_config(self.h, _ptr(c)), "clast_hits_dp_config")
        return int(c[0]), int(c[1]), int(c[2])

    def close(self) -> None:
        if self.h:
            hip().clast_hits_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def search(params: Params, index: Index, query: Query, t_begin: int, q_begin: int) -> Hits:
    """CDeviceHitList for one (target window, query window); q_begin is the
    strand ID of the window's first sequence (2 * pair)."""
    h = C.c_void_p()
    _check(hip().clast_search(C.byref(params.c()), index.h, query.h, t_begin, q_begin, C.byref(h)), "clast_search")
    return Hits(h)


def search_batch(params: Params, index: Index, query: Query, keep_stages: bool = False) -> Hits:
    h = C.c_void_p()
    _check(hip().clast_search_batch(C.byref(params.c()), index.h, query.h, CLAST_KEEP_STAGES if keep_stages else 0,
                                    C.byref(h)), "clast_search_batch")
    return Hits(h)


NOALLOC_SCRATCH_BYTES = 256 << 20  # CLAST_NOALLOC_SCRATCH_BYTES


def search_batch_noalloc(params: Params, index: Index, query: Query, capacity: int) -> tuple[int, dict]:
    """clast_search_batch_noalloc: the batched search with its device buffers
    in the library's fixed per-device scratch and the rows written into
    host arrays of `capacity` rows.  Returns (status, rows); on
    CLAST_ERR_LIMIT rows["needed"] holds the rows the search produced."""
    out = {k: np.zeros(max(capacity, 1), np.int32) for k in HIT_FIELDS + ("twin", "qwin")}
    out["evalue"] = np.zeros(max(capacity, 1), np.float64)
    n = C.c_int64(0)
    cols = [_ptr(out[k]) for k in HIT_FIELDS]
    rc = hip().clast_search_batch_noalloc(C.byref(params.c()), index.h, query.h, capacity, *cols,
                                         _ptr(out["evalue"]), _ptr(out["twin"]), _ptr(out["qwin"]), C.byref(n))
    rows = {k: v[:n.value] for k, v in out.items()} if rc == 0 else {}
    rows["needed"] = int(n.value)
    return rc, rows


def last_error() -> str:
    return hip().clast_last_error().decode()


def result_sort(rows: dict[str, np.ndarray]) -> None:
    """resultSorting on the device, in place (stable (qID/2, score desc))."""
    n = len(rows["qID"])
    cols = [np.ascontiguousarray(rows[k], dtype=np.int32) for k in HIT_FIELDS]
    ev = np.ascontiguousarray(rows["evalue"], dtype=np.float64)
    _check(hip().clast_result_sort(n, *[_ptr(c) for c in cols], _ptr(ev)), "clast_result_sort")
    for k, c in zip(HIT_FIELDS, cols):
        rows[k] = c
    rows["evalue"] = ev


def result_finalize(rows: dict[str, np.ndarray], qlabel: np.ndarray, tlabel: np.ndarray, tstart: np.ndarray,
                    num_out: int) -> tuple[np.ndarray, np.ndarray, int]:
    """fixResult's order and printResult's row selection on the device
    (clast_result_finalize): (order, printed flags per ordered position,
    rows printed).  qlabel / tlabel / tstart are per-row label ids and
    target start offsets."""
    n = len(rows["qID"])
    c = {k: np.ascontiguousarray(rows[k], dtype=np.int32) for k in HIT_FIELDS}
    ev = np.ascontiguousarray(rows["evalue"], dtype=np.float64)
    ql, tl, ts = (np.ascontiguousarray(a, dtype=np.int32) for a in (qlabel, tlabel, tstart))
    order = np.zeros(max(n, 1), np.int32)
    printed = np.zeros(max(n, 1), np.uint8)
    npr = C.c_int64(0)
    _check(hip().clast_result_finalize(n, *[_ptr(c[k]) for k in HIT_FIELDS[1:]], _ptr(ev), _ptr(ql), _ptr(tl),
                                       _ptr(ts), num_out, _ptr(order), _ptr(printed), C.byref(npr)),
           "clast_result_finalize")
    return order[:n], printed[:n], int(npr.value)
