"""Serving loop: a resident target index in HBM and continuous batching of
read requests.

Each request is a set of reads (a query RAM chunk of its own).  Its query
windows are the ones the driver would build for that request alone (qVRAM
windows over '+'/'-' pairs, computed from the read lengths); the strands
themselves are built on the device from one upload of the batch's reads
(clast_query_create_reads).  Requests that have arrived while the device was
busy are admitted in arrival order and coalesced into one query set (one
window list, request after request), searched in one launch chain, and the
rows are split back per request.  Because the only coupling between queries
is per (target window, query window) unit and every request keeps its own
windows, each request's rows are identical alone and in the mix.

A request is cut at query-window boundaries into parts of a bounded number
of query k-mer positions, so that the device threads share a large request
instead of one of them running it alone; its rows are its parts' rows in
window order, equal to the request's rows alone.

Admission is by device memory.  At start the server reads the free device
memory (hipMemGetInfo, with the index already resident) and keeps a budget of
90% of it less a reserve; each of the `workers` device threads may hold one
batch, so a batch is closed before its estimated device bytes pass the
budget's share of one worker, or the free memory left at that moment
(parts are admitted in order, whole).  The
estimate is bytes per query k-mer position (the query set's own tables plus
the search workspace), learned from the workers: after each search a worker
reads its workspace size (it grows to the largest batch it has run and is
kept), and the estimate is the largest ratio of workspace to positions seen.
A part larger than a whole share is admitted alone.

Host preparation of the next batch (list building, windows) overlaps the
device search of the current one: a preparation thread feeds a queue and
`workers` device threads drain it, each with its own library streams, so
that many batches are in flight on the GPU at once.  ctypes releases the GIL
during native calls.
"""
from __future__ import annotations

import collections
import dataclasses
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import native


@dataclass
class Request:
    rid: int
    bases: np.ndarray        # concatenated reads (uint8)
    offsets: np.ndarray      # n_reads + 1
    arrival: float = 0.0     # seconds since the start of the run
    # filled by the server
    done: float = 0.0
    rows: dict = field(default_factory=dict)
    _parts: dict = field(default_factory=dict, repr=False)
    _nparts: int = 0

    @property
    def n_bases(self) -> int:
        return int(self.offsets[-1] - self.offsets[0])


@dataclass
class Part:
    """Consecutive query windows of one request (read pair ranges in the
    request's own numbering), searched within one batch."""
    req: Request
    k: int
    wins: list
    positions: int
    slots: int


class Server:
    """Continuous-batching search over a resident index, admitted by a device
    memory budget.

    A request is cut at query-window boundaries into parts of at most
    part_positions query k-mer positions (a window is never cut), so that the
    device threads can share a large request; a request's rows are its parts'
    rows in window order, which is the order the request alone gets (each
    part holds whole read pairs and every target window, and the rows of a
    batch come pair by pair)."""

    # starting estimate of device bytes per query k-mer position, before any
    # search has run: 16 B cell record, 4 B strand, and room for seeds and hits
    START_BYTES_PER_POSITION = 160.0
    # device bytes per query slot of the query set itself (class bytes, 4-bit
    # copy) on top of the positions' share
    BYTES_PER_SLOT = 2.0

    def __init__(self, params: native.Params, index: native.Index, q_vram: float = 2e6,
                 workers: int = 2, mem_fraction: float = 0.9, reserve_bytes: int = 2 << 30,
                 max_positions: int | None = None, part_positions: int | None = None):
        self.p = params
        self.index = index
        self.q_vram = q_vram
        self.workers = max(1, workers)
        self.max_positions = max_positions
        self.part_positions = part_positions
        free, total = native.mem_info()
        self.reserve = reserve_bytes
        self.mem_budget = max(0, int((free - reserve_bytes) * mem_fraction))
        self.share = self.mem_budget // self.workers
        self.bytes_per_position = self.START_BYTES_PER_POSITION
        self._ws = {}        # worker -> (largest positions run, workspace bytes)
        self._lock = threading.Lock()
        self.records = []    # per batch: (parts, positions, kernel times, counts, DP kernels)

    def estimate(self, positions: int, slots: int) -> float:
        """Device bytes of a batch: the workspace estimate with a quarter of
        headroom (seed counts per position vary between batches), plus the
        query set: 4 B strand index per position and its class bytes."""
        return (1.25 * self.bytes_per_position + 4.0) * positions + self.BYTES_PER_SLOT * slots

    def _learn(self, worker: int, positions: int) -> None:
        ws = native.workspace_bytes()
        with self._lock:
            top, _ = self._ws.get(worker, (0, 0))
            self._ws[worker] = (max(top, positions), ws)
            ratios = [w / t for t, w in self._ws.values() if t > 0]
            if ratios:
                self.bytes_per_position = max(ratios)

    def _cap(self) -> float:
        """Device bytes the next batch may take: one worker's share of the
        budget, or what is free now (less the reserve) plus the smallest
        workspace a worker already holds, whichever is less."""
        free, _ = native.mem_info()
        with self._lock:
            held = min((w for _, w in self._ws.values()), default=0) if len(self._ws) == self.workers else 0
        return min(self.share, max(free - self.reserve, 0) + held)

    def positions(self, r: Request) -> int:
        ln = np.diff(r.offsets)
        return int(2 * np.clip(ln - self.p.lmer + 1, 0, None).sum())

    def split(self, r: Request, whole: bool = False) -> list:
        """The request's parts (one part when whole or when no part size is set)."""
        ln = np.diff(r.offsets)
        g = native.read_gateway(ln, self.p.lmer)
        wins = native.windows(g, self.q_vram, True)
        pos = np.concatenate([[0], np.cumsum(2 * np.clip(ln - self.p.lmer + 1, 0, None))])
        slot = np.concatenate([[0], np.cumsum(2 * (ln + max(self.p.lmer - 1, 0)))])
        parts, cur = [], []
        cap = None if whole else self.part_positions

        def close():
            a, b = cur[0][0], cur[-1][1]
            parts.append(Part(r, len(parts), list(cur), int(pos[b] - pos[a]), int(slot[b] - slot[a])))
            cur.clear()

        for a, b in wins:
            if cur and cap is not None and pos[b] - pos[cur[0][0]] > cap:
                close()
            cur.append((a, b))
        if cur:
            close()
        r._nparts = len(parts)
        r._parts = {}
        if not parts:  # a request with no reads
            r.rows = {}
        return parts

    def search(self, parts: list, worker: int = 0, record: bool = False) -> list:
        """One launch chain for the batch; returns each part's rows."""
        owner = [k for k, pt in enumerate(parts) for _ in pt.wins]
        if not owner:
            return [{} for _ in parts]
        # one upload of the batch's reads; each part's windows shifted to its
        # reads' place in the batch, list IDs numbered per request
        chunks, offs, bwins, ids, first = [], [np.zeros(1, np.int64)], [], [], 0
        for pt in parts:
            r = pt.req
            a, b = pt.wins[0][0], pt.wins[-1][1]
            lo, hi = int(r.offsets[a]), int(r.offsets[b])
            chunks.append(r.bases[lo:hi])
            offs.append((r.offsets[a + 1:b + 1] - lo).astype(np.int64) + offs[-1][-1])
            bwins += [(first + wa - a, first + wb - a) for wa, wb in pt.wins]
            ids += [2 * wa for wa, _ in pt.wins]
            first += b - a
        q = native.Query.from_reads(self.p, np.concatenate(chunks), np.concatenate(offs), bwins, ids)
        npos = sum(pt.positions for pt in parts)
        try:
            hits = native.search_batch(self.p, self.index, q)
            rows = hits.arrays()
            if record:
                with self._lock:
                    self.records.append((len(parts), npos, hits.kernel_times(), hits.counts(), hits.dp_kernels()))
            hits.close()
        finally:
            q.close()
        self._learn(worker, npos)
        part_of_row = np.asarray(owner, dtype=np.int64)[rows["qwin"]] if len(rows["qwin"]) else np.zeros(0, np.int64)
        out = []
        for k in range(len(parts)):
            sel = np.nonzero(part_of_row == k)[0]
            out.append({key: v[sel] for key, v in rows.items() if key != "qwin"})
        return out

    def _deliver(self, parts: list, rows: list, t: float) -> int:
        """Stores each part's rows; a request whose parts are all in gets its
        rows (parts in order) and its completion time.  Returns the number of
        requests completed."""
        done = 0
        with self._lock:
            for pt, rw in zip(parts, rows):
                r = pt.req
                r._parts[pt.k] = rw
                if len(r._parts) == r._nparts:
                    ps = [r._parts[k] for k in range(r._nparts)]
                    r.rows = {key: np.concatenate([x[key] for x in ps]) for key in ps[0]} if ps[0] else {}
                    r._parts = {}
                    r.done = t
                    done += 1
        return done

    def search_alone(self, r: Request) -> dict:
        """The request as one batch of its own, unsplit."""
        parts = self.split(r, whole=True)
        if parts:
            self._deliver(parts, self.search(parts), 0.0)
        return r.rows

    def run(self, requests: list, realtime: bool = True, record: bool = False) -> dict:
        """Serves `requests` (sorted by arrival).  realtime: requests become
        visible at their arrival time on the wall clock; otherwise all are
        queued at once.  record: keep each batch's kernel times and counts
        in self.records.  Returns latency, throughput and admission figures."""
        requests = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        ready: queue.Queue = queue.Queue(maxsize=self.workers + 1)
        admitted, caps, nparts = [0], [], [0]

        def producer():
            i = 0
            pend: collections.deque = collections.deque()
            while i < len(requests) or pend:
                if realtime and not pend:  # wait for the next arrival
                    wait = requests[i].arrival - now()
                    if wait > 0:
                        time.sleep(wait)
                while i < len(requests) and (not realtime or requests[i].arrival <= now()):
                    ps = self.split(requests[i])
                    if not ps:
                        requests[i].done = now()
                    pend.extend(ps)
                    nparts[0] += len(ps)
                    i += 1
                if not pend:
                    continue
                # admit pending parts in order while the batch's estimated
                # device bytes fit
                cap = self._cap()
                batch, npos, nslot = [], 0, 0
                while pend:
                    pt = pend[0]
                    over = self.estimate(npos + pt.positions, nslot + pt.slots) > cap or (
                        self.max_positions is not None and npos + pt.positions > self.max_positions)
                    if batch and over:
                        break
                    batch.append(pend.popleft())
                    npos += pt.positions
                    nslot += pt.slots
                with self._lock:
                    admitted[0] = max(admitted[0], npos)
                    caps.append(cap)
                ready.put(batch)
            ready.put(None)

        th = threading.Thread(target=producer, daemon=True)
        th.start()
        batches = [0]
        errors = []

        def worker(k):
            while True:
                batch = ready.get()
                if batch is None:
                    ready.put(None)  # for the other workers
                    return
                try:
                    rows = self.search(batch, worker=k, record=record)
                except Exception as e:  # noqa: BLE001 - reported after the run
                    errors.append(e)
                    rows = [{} for _ in batch]
                self._deliver(batch, rows, now())
                with self._lock:
                    batches[0] += 1

        ws = [threading.Thread(target=worker, args=(k,), daemon=True) for k in range(self.workers)]
        for w in ws:
            w.start()
        for w in ws:
            w.join()
        th.join()
        if errors:
            raise errors[0]
        wall = now()
        lat = np.array([(r.done - r.arrival) * 1e3 for r in requests])
        mbp = sum(r.n_bases for r in requests) / 1e6
        return {"requests": len(requests), "parts": nparts[0], "batches": batches[0], "workers": self.workers,
                "wall_s": round(wall, 3), "query_mbp": round(mbp, 3), "mbp_per_s": round(mbp / wall, 3),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_batch_parts": round(nparts[0] / max(batches[0], 1), 2),
                "admission": {"mem_budget_bytes": self.mem_budget, "per_batch_share_bytes": self.share,
                              "min_batch_cap_bytes": int(min(caps)) if caps else 0,
                              "bytes_per_position": round(self.bytes_per_position, 1),
                              "largest_batch_positions": admitted[0],
                              "part_positions": self.part_positions,
                              "workspace_bytes": {k: w for k, (_, w) in sorted(self._ws.items())}}}


def make_requests(rspec, target, n_requests: int, rate: float, seed: int, min_reads: int = 1,
                  max_reads: int = 10_000) -> list:
    """Exactly n_requests requests of log-uniform read counts in [min_reads,
    max_reads], arriving as a Poisson process of `rate` per second, with
    reads drawn from the generator spec `rspec` against `target`."""
    from . import gen

    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(min_reads), np.log(max_reads), n_requests))).astype(np.int64)
    sizes = np.maximum(sizes, 1)
    gaps = rng.exponential(1.0 / rate, n_requests)
    r = gen.make_reads(dataclasses.replace(rspec, n_reads=int(sizes.sum()), seed=int(rspec.seed) + seed), target)
    out, i, t = [], 0, 0.0
    for k, g in zip(sizes, gaps):
        t += g
        lo, hi = int(r.offsets[i]), int(r.offsets[i + k])
        out.append(Request(len(out), r.bases[lo:hi], (r.offsets[i:i + k + 1] - lo).astype(np.int64), t))
        i += int(k)
    return out


def poisson_requests(reads_bases: np.ndarray, reads_offsets: np.ndarray, rate: float, seed: int,
                     min_reads: int = 1, max_reads: int = 10_000) -> list:
    """Cuts a read set into requests of log-uniform size arriving as a
    Poisson process of `rate` requests per second."""
    rng = np.random.default_rng(seed)
    n = len(reads_offsets) - 1
    out, i, t = [], 0, 0.0
    while i < n:
        k = int(round(np.exp(rng.uniform(np.log(min_reads), np.log(max_reads)))))
        k = max(1, min(k, n - i))
        t += rng.exponential(1.0 / rate)
        lo, hi = int(reads_offsets[i]), int(reads_offsets[i + k])
        out.append(Request(len(out), reads_bases[lo:hi], (reads_offsets[i:i + k + 1] - lo).astype(np.int64), t))
        i += k
    return out
