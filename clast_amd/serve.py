"""Serving loop: a resident target index in HBM and continuous batching of
read requests.

Each request is a set of reads (a query RAM chunk of its own).  Its query
windows are the ones the driver would build for that request alone (qVRAM
windows over '+'/'-' pairs, computed from the read lengths); the strands
themselves are built on the device from one upload of the batch's reads
(clast_query_create_reads).  Requests that have
arrived while the device was busy are admitted in arrival order up to a
per-launch budget of query k-mer positions and coalesced into one query set
(one window list, request after request), searched in one launch chain, and
the rows are split back per request.  Because the only coupling between
queries is per (target window, query window) unit and every request keeps its
own windows, each request's rows are identical alone and in the mix.

Host preparation of the next batch (list building, windows) overlaps the
device search of the current one: a preparation thread feeds a queue and
`workers` device threads drain it, each with its own library streams, so
that many batches are in flight on the GPU at once.  ctypes releases the GIL
during native calls.
"""
from __future__ import annotations

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
    """Continuous-batching search over a resident index."""

    def __init__(self, params: native.Params, index: native.Index, q_vram: float = 2e6,
                 max_positions: int = 64_000_000, workers: int = 2):
        self.p = params
        self.index = index
        self.q_vram = q_vram
        self.max_positions = max_positions
        self.workers = max(1, workers)

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

    def search(self, prep: Prepared) -> None:
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
            hits.close()
        finally:
            q.close()
        req_of_row = np.asarray(owner, dtype=np.int64)[rows["qwin"]] if len(rows["qwin"]) else np.zeros(0, np.int64)
        for k, r in enumerate(prep.requests):
            sel = np.nonzero(req_of_row == k)[0]
            r.rows = {key: v[sel] for key, v in rows.items() if key not in ("qwin",)}

    def search_alone(self, r: Request) -> dict:
        self.search(self.prepare([r]))
        return r.rows

    def run(self, requests: list, realtime: bool = True) -> dict:
        """Serves `requests` (sorted by arrival).  realtime: requests become
        visible at their arrival time on the wall clock; otherwise all are
        queued at once.  Returns latency and throughput figures."""
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
                # admit everything that has arrived, up to the position budget
                batch, npos = [], 0
                while i < len(pending) and (not realtime or pending[i].arrival <= now()):
                    need = self.positions(pending[i])
                    if batch and npos + need > self.max_positions:
                        break
                    batch.append(pending[i])
                    npos += need
                    i += 1
                ready.put(self.prepare(batch))
            ready.put(None)

        th = threading.Thread(target=producer, daemon=True)
        th.start()
        batches = [0]
        errors = []

        def worker():
            while True:
                prep = ready.get()
                if prep is None:
                    ready.put(None)  # for the other workers
                    return
                try:
                    self.search(prep)
                except Exception as e:  # noqa: BLE001 - reported after the run
                    errors.append(e)
                t = now()
                with lock:
                    for r in prep.requests:
                        r.done = t
                    batches[0] += 1

        ws = [threading.Thread(target=worker, daemon=True) for _ in range(self.workers)]
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
                "mean_batch": round(len(requests) / max(batches[0], 1), 2)}


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
