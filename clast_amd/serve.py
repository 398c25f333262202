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

Admission is by device memory.  At start the server reads the free device
memory (hipMemGetInfo, with the index already resident) and keeps a budget of
90% of it less a reserve; each of the `workers` device threads may hold one
batch, so a batch is closed before its estimated device bytes pass the
budget's share of one worker, or the free memory left at that moment.  The
estimate is bytes per query k-mer position (the query set's own tables plus
the search workspace), learned from the workers: after each search a worker
reads its workspace size (it grows to the largest batch it has run and is
kept), and the estimate is the largest ratio of workspace to positions seen.
A request larger than a whole share is admitted alone.

Host preparation of the next batch (list building, windows) overlaps the
device search of the current one: a preparation thread feeds a queue and
`workers` device threads drain it, each with its own library streams, so
that many batches are in flight on the GPU at once.  ctypes releases the GIL
during native calls.
"""
from __future__ import annotations

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

    @property
    def n_bases(self) -> int:
        return int(self.offsets[-1] - self.offsets[0])


@dataclass
class Prepared:
    requests: list
    wins: list               # query windows (read pair ranges) per request
    positions: int


class Server:
    """Continuous-batching search over a resident index, admitted by a device
    memory budget."""

    # starting estimate of device bytes per query k-mer position, before any
    # search has run: 16 B cell record, 4 B strand, and room for seeds and hits
    START_BYTES_PER_POSITION = 160.0
    # device bytes per query slot of the query set itself (class bytes, 4-bit
    # copy) on top of the positions' share
    BYTES_PER_SLOT = 2.0

    def __init__(self, params: native.Params, index: native.Index, q_vram: float = 2e6,
                 workers: int = 2, mem_fraction: float = 0.9, reserve_bytes: int = 2 << 30,
                 max_positions: int | None = None):
        self.p = params
        self.index = index
        self.q_vram = q_vram
        self.workers = max(1, workers)
        self.max_positions = max_positions
        free, total = native.mem_info()
        self.reserve = reserve_bytes
        self.mem_budget = max(0, int((free - reserve_bytes) * mem_fraction))
        self.share = self.mem_budget // self.workers
        self.bytes_per_position = self.START_BYTES_PER_POSITION
        self._ws = {}        # worker -> (largest positions run, workspace bytes)
        self._lock = threading.Lock()
        self.records = []    # per batch: (requests, positions, kernel times, counts)

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

    def prepare(self, reqs: list) -> Prepared:
        wins, npos = [], 0
        for r in reqs:
            g = native.read_gateway(np.diff(r.offsets), self.p.lmer)
            wins.append(native.windows(g, self.q_vram, True))
            npos += self.positions(r)
        return Prepared(reqs, wins, npos)

    def slots(self, r: Request) -> int:
        ln = np.diff(r.offsets)
        return int(2 * (ln + max(self.p.lmer - 1, 0)).sum())

    def search(self, prep: Prepared, worker: int = 0, record: bool = False) -> None:
        """One launch chain for the whole batch; rows split per request."""
        owner = [k for k, wl in enumerate(prep.wins) for _ in wl]
        if not owner:
            for r in prep.requests:
                r.rows = {}
            return
        # one upload of the batch's reads; each request's windows shifted to
        # its reads' place in the batch, list IDs numbered per request
        bases = np.concatenate([r.bases[r.offsets[0]:r.offsets[-1]] for r in prep.requests])
        offs = [np.zeros(1, np.int64)]
        bwins, ids, first = [], [], 0
        for r, wl in zip(prep.requests, prep.wins):
            offs.append((r.offsets[1:] - r.offsets[0]).astype(np.int64) + offs[-1][-1])
            bwins += [(first + a, first + b) for a, b in wl]
            ids += [2 * a for a, _ in wl]
            first += len(r.offsets) - 1
        q = native.Query.from_reads(self.p, bases, np.concatenate(offs), bwins, ids)
        try:
            hits = native.search_batch(self.p, self.index, q)
            rows = hits.arrays()
            if record:
                with self._lock:
                    self.records.append((len(prep.requests), prep.positions, hits.kernel_times(), hits.counts()))
            hits.close()
        finally:
            q.close()
        self._learn(worker, prep.positions)
        req_of_row = np.asarray(owner, dtype=np.int64)[rows["qwin"]] if len(rows["qwin"]) else np.zeros(0, np.int64)
        for k, r in enumerate(prep.requests):
            sel = np.nonzero(req_of_row == k)[0]
            r.rows = {key: v[sel] for key, v in rows.items() if key not in ("qwin",)}

    def search_alone(self, r: Request) -> dict:
        self.search(self.prepare([r]))
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
        pending = list(requests)
        lock = threading.Lock()

        def producer():
            i = 0
            while i < len(pending):
                # wait for the next arrival
                if realtime:
                    wait = pending[i].arrival - now()
                    if wait > 0:
                        time.sleep(wait)
                # admit what has arrived while its estimated device bytes fit
                cap = self._cap()
                batch, npos, nslot = [], 0, 0
                while i < len(pending) and (not realtime or pending[i].arrival <= now()):
                    need, sl = self.positions(pending[i]), self.slots(pending[i])
                    over = self.estimate(npos + need, nslot + sl) > cap or (
                        self.max_positions is not None and npos + need > self.max_positions)
                    if batch and over:
                        break
                    batch.append(pending[i])
                    npos += need
                    nslot += sl
                    i += 1
                with self._lock:
                    admitted[0] = max(admitted[0], npos)
                    caps.append(cap)
                ready.put(self.prepare(batch))
            ready.put(None)

        admitted, caps = [0], []
        th = threading.Thread(target=producer, daemon=True)
        th.start()
        batches = [0]
        errors = []

        def worker(k):
            while True:
                prep = ready.get()
                if prep is None:
                    ready.put(None)  # for the other workers
                    return
                try:
                    self.search(prep, worker=k, record=record)
                except Exception as e:  # noqa: BLE001 - reported after the run
                    errors.append(e)
                t = now()
                with lock:
                    for r in prep.requests:
                        r.done = t
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
        return {"requests": len(requests), "batches": batches[0], "workers": self.workers, "wall_s": round(wall, 3),
                "query_mbp": round(mbp, 3), "mbp_per_s": round(mbp / wall, 3),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_batch": round(len(requests) / max(batches[0], 1), 2),
                "admission": {"mem_budget_bytes": self.mem_budget, "per_batch_share_bytes": self.share,
                              "min_batch_cap_bytes": int(min(caps)) if caps else 0,
                              "bytes_per_position": round(self.bytes_per_position, 1),
                              "largest_batch_positions": admitted[0],
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
