#!/usr/bin/env python3
"""Benchmark of the CLAST search path on one MI355X.

Workload (default ``large``): 1 Gbp target (1,000 x 1 Mbp, 10% repeat
content) and 1,000,000 reads of 150 bp, reference defaults (lMer 15, stride
5, repeat 20, width 100, gap 8, global, cutOff 10; tRAM = qRAM = tVRAM = 64
MBase, qVRAM = 2 MBase).  The target's 16 RAM chunks (one window each) are
indexed once and kept resident in HBM; each of the 5 query RAM chunks (~35
query windows each) is uploaded before the timed region.  One step = every
query chunk searched against every target window (clast_search_batch per
query chunk), hits copied back to host memory.

value = query Mbp (one strand, as read) aligned per second over the K timed
steps.  latency = per query-chunk search time (p50/p99 over all timed
searches).  roofline = the seed lookup kernel (k_seed_count) against HBM
peak; its traffic (HBM bytes per launch) comes from the counter runs recorded
in profiles/large_hbm_traffic.txt.  cpu_baseline = the CPU oracle on a bounded sample (one target window
x 16 query windows, OpenMP, searches only), scaled to the whole target; the
same sample is searched on the GPU and must match it bit-exactly.
"""
from __future__ import annotations

import argparse
from concurrent.futures import ThreadPoolExecutor
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_INT32_PEAK_TOPS = 39.3  # 256 CU x 64 lanes x 2.4 GHz, one int32 op per lane per clock


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_target(name: str, scale: float, p, t_ram: float, t_vram: float):
    """Target chunks as the FASTA loader would cut them, their host lists and
    windows, flattened in (chunk, window) order."""
    from clast_amd import gen, native

    tspec, _ = gen.workload(name, scale)
    t = gen.make_target(tspec)
    chunks, keep = native.chunks(t.lengths, int(t_ram))
    out = []
    for a, b in chunks:
        idx = [i for i in range(a, b) if keep[i]]
        if not idx:
            continue
        sub = t.subset(idx[0], idx[-1] + 1)
        lst = native.build_list(sub.bases, sub.offsets, p.lmer, False)
        wins = native.windows(lst.gateway, t_vram, False)
        out.append((lst, wins))
    return t, out


def build_queries(name: str, scale: float, p, t, q_ram: float, q_vram: float):
    from clast_amd import gen, native

    _, rspec = gen.workload(name, scale)
    r = gen.make_reads(rspec, t)
    chunks, keep = native.chunks(r.lengths, int(q_ram / 2))
    out = []
    for a, b in chunks:
        idx = [i for i in range(a, b) if keep[i]]
        if not idx:
            continue
        sub = r.subset(idx[0], idx[-1] + 1)
        lst = native.build_list(sub.bases, sub.offsets, p.lmer, True)
        wins = native.windows(lst.gateway, q_vram, True)
        out.append((lst, wins, int(sub.lengths.sum())))
    return r, out


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload (FETCH_SIZE +
    WRITE_SIZE, calibrated against a plain device copy), as recorded by the
    last counter runs in profiles/large_hbm_traffic.txt; None if absent."""
    path = os.path.join(ROOT, "profiles", "large_hbm_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                part = line.split()
                if len(part) == 3 and part[0] == "traffic_bytes_per_launch" and part[1] == kernel:
                    return float(part[2])
    except OSError:
        pass
    return None


def full_size_checks(p, args, tchunks, qchunks, indexes, queries) -> dict:
    """Size-independent checks on every row of one full step (untimed): the
    rows of each search with the bench's threads equal those one search at a
    time; rows are in the reference's output order (query pair ascending,
    score descending: CHostResultHolder.cu:100-109), each E-value is
    K D qLen exp(-lambda score)
    (krnlCalculateEvalue.cu:18-21) within 1e-9, survivors pass the bad-hit
    filters (E <= cutOff, qHL >= lMer + stride; ...deleteBadHits.cu), and
    hits lie inside their sequences with match <= both lengths; rows that
    share a (target, tIdx, query, qIdx) start are counted."""
    from clast_amd import native

    keys = [(qi, ii) for qi in range(len(queries)) for ii in range(len(indexes))]

    def collect(streams):
        res = {}

        def run(part):
            for qi, ii in part:
                h = native.search_batch(p, indexes[ii], queries[qi])
                res[(qi, ii)] = h.arrays()
                h.close()

        parts = [keys[k::streams] for k in range(streams)]
        with ThreadPoolExecutor(max_workers=streams) as ex:
            list(ex.map(run, parts))
        return res

    many, one = collect(max(args.streams, 1)), collect(1)
    fields = native.HIT_FIELDS + ("evalue", "twin", "qwin")
    threads_equal = all(np.array_equal(many[k][f], one[k][f]) for k in keys for f in fields)
    # window -> (target chunk, window in chunk) for a merged index
    wchunk = [c for c, (_, w) in enumerate(tchunks) for _ in w]
    rows = 0
    order = filters = bounds = True
    repeated = 0
    emax = 0.0
    for qi, ii in keys:
        a = one[(qi, ii)]
        n = len(a["tID"])
        rows += n
        if n == 0:
            continue
        pair, sc = a["qID"] // 2, a["score"].astype(np.int64)
        order &= bool(np.all((pair[1:] > pair[:-1]) | ((pair[1:] == pair[:-1]) & (sc[1:] <= sc[:-1]))))
        # rows sharing a start: seeds on diagonals too far apart for the
        # duplicate filter can extend to one start; the reference keeps
        # both (no row dedup after the DP), so this is a count, not a check
        tup = np.stack([a["twin"], a["tID"], a["tIdx"], a["qID"], a["qIdx"]], axis=1)
        repeated += n - len(np.unique(tup, axis=0))
        qlen = qchunks[qi][0].length[a["qID"]].astype(np.float64)
        exp_e = p.K * p.total_db_size * qlen * np.exp(-p.lam * sc.astype(np.float64))
        emax = max(emax, float(np.max(np.abs(a["evalue"] - exp_e) / np.maximum(exp_e, 1e-300))))
        filters &= bool(np.all(a["qHL"] >= p.lmer + p.stride))
        if p.cut_off != -1:
            filters &= bool(np.all(a["evalue"] <= p.cut_off))
        chunk = np.array([wchunk[w] for w in a["twin"]]) if len(indexes) == 1 else np.full(n, ii)
        tlen = np.array([tchunks[c][0].length[t] for c, t in zip(chunk, a["tID"])], dtype=np.int64)
        bounds &= bool(np.all((a["qIdx"] >= 0) & (a["qIdx"] + a["qHL"] <= qlen) & (a["tIdx"] >= 0) &
                              (a["tIdx"].astype(np.int64) + a["tHL"] <= tlen) &
                              (a["match"] <= np.minimum(a["qHL"], a["tHL"]))))
    return {"rows": rows, "threads_equal_one_at_a_time": bool(threads_equal), "output_order": bool(order),
            "rows_sharing_a_start": int(repeated), "evalue_max_rel": emax, "bad_hit_filters": bool(filters),
            "within_sequences": bool(bounds)}


def cpu_baseline(p, tchunks, qchunks, n_twin_total: int, threads: int):
    """The oracle (OpenMP, one query window per thread) on one target window
    x `threads` query windows.  Like the GPU figure it counts the searches
    only: the oracle builds that window's table first, timed apart.  Query
    Mbp/s is scaled by the number of target windows every read meets.  The
    GPU searches the same units and must match bit-exactly."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from clast_amd import native

    tl, twins = tchunks[0]
    ql, qwins, _ = qchunks[0]
    tw, qws = twins[0], qwins[:threads]
    q_bases = sum(int(ql.length[2 * a:2 * b:2].sum()) for a, b in qws)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = oracle.search(p, tl, [tw], ql, qws)
    exp = o["final"]
    ix = native.Index(p, tl, [tw])
    q = native.Query(p, ql, qws)
    got = native.search_batch(p, ix, q).arrays()
    ix.close()
    q.close()
    exact = len(got["tID"]) == len(exp["tID"]) and all(
        np.array_equal(got[k], exp[k]) for k in ("tID", "tIdx", "qID", "qIdx", "tHL", "qHL", "match", "score"))
    rel = float(np.max(np.abs(got["evalue"] - exp["evalue"]) / np.maximum(exp["evalue"], 1e-300))) \
        if exact and len(exp["tID"]) else 0.0
    mbps = q_bases / 1e6 / (o["search_s"] * n_twin_total)
    sample = (f"oracle on {threads} threads: 1 of {n_twin_total} target windows ({tw[1] - tw[0]} seqs, table "
              f"built in {o['index_s']:.1f} s, not counted) x {len(qws)} query windows "
              f"({qws[-1][1] - qws[0][0]} reads, {q_bases / 1e6:.2f} Mbp) searched in {o['search_s']:.2f} s; "
              f"Mbp/s scaled by {n_twin_total} target windows")
    return {"value": round(mbps, 4), "unit": "query Mbp/s aligned", "cores": threads, "kind": "port",
            "sample": sample}, {"rows": int(len(exp["tID"])), "bit_exact": bool(exact), "evalue_max_rel": rel}


def serve_bench(args) -> int:
    """The serving loop on the 3 Gbp workload (40% repeats, 1% low
    complexity): the index of every target window resident in HBM, requests of
    1..10,000 reads (log-uniform count; lengths 50/50 150 bp and 1-15 kbp)
    arriving as a Poisson process, admitted by a query k-mer budget and
    coalesced into shared launch chains; latency = arrival to rows in host
    memory."""
    import dataclasses

    from clast_amd import gen, native, serve

    native.init(0)
    p = native.Params()
    t0 = time.perf_counter()
    t, tchunks = build_target("serving", args.scale, p, 64e6, 64e6)
    p.total_db_size = float(t.bases.shape[0])
    _, rspec = gen.workload("serving", args.scale)
    rspec = dataclasses.replace(rspec, n_reads=args.serve_reads)
    r = gen.make_reads(rspec, t)
    log(f"generated {t.bases.shape[0] / 1e9:.2f} Gbp target, {len(r)} reads ({r.bases.shape[0] / 1e6:.0f} Mbp) "
        f"in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    ix = native.Index.merged(p, tchunks)
    index_s = time.perf_counter() - t1
    n_twin = sum(len(w) for _, w in tchunks)
    log(f"index over {n_twin} target windows built in {index_s:.2f} s")
    srv = serve.Server(p, ix, q_vram=2e6, max_positions=args.serve_budget, workers=args.serve_workers)
    reqs = serve.poisson_requests(r.bases, r.offsets, rate=args.rate, seed=7, min_reads=1, max_reads=10_000)
    srv.search_alone(serve.Request(-1, reqs[0].bases, reqs[0].offsets))  # warm up
    log(f"{len(reqs)} requests at {args.rate} req/s")
    stats = srv.run(reqs, realtime=True)
    out = {
        "metric": "served query Mbp/s + p50/p99 request latency (arrival to rows in host memory), 1 MI355X",
        "value": stats["mbp_per_s"], "unit": "query Mbp/s", "higher_is_better": True,
        "latency_ms": {"p50": stats["p50_ms"], "p99": stats["p99_ms"]},
        "config": {"workload": f"serving: {t.bases.shape[0] / 1e9:.2f} Gbp target ({len(t)} seqs, 40% repeats), "
                               f"{len(r)} reads, Poisson {args.rate} req/s",
                   "target_windows": n_twin, "admission_budget_kmer_positions": args.serve_budget},
        "serving": stats, "index_build_s": round(index_s, 3), "dtype": "int32", "data": "synthetic",
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large")
    ap.add_argument("--streams", type=int, default=3,
                    help="host threads issuing a step's searches (each with its own library streams)")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full-checks", action="store_true",
                    help="skip the untimed property checks over every row of one full step")
    ap.add_argument("--plain", action="store_true", help="also time the per-window (one launch chain per unit) path")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--index", choices=("merged", "per-chunk"), default="merged",
                    help="one resident index over every target window, or one per target chunk")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only copy MiB of device memory once (the known-size kernel that scales the HBM "
                         "traffic counters) and exit")
    ap.add_argument("--serve", action="store_true", help="run the serving loop instead (Poisson request stream)")
    ap.add_argument("--serve-reads", type=int, default=100_000)
    ap.add_argument("--serve-workers", type=int, default=2, help="device threads draining the batch queue (--serve)")
    ap.add_argument("--rate", type=float, default=50.0, help="requests per second (--serve)")
    ap.add_argument("--serve-budget", type=int, default=64_000_000,
                    help="query k-mer positions admitted per launch chain (--serve)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.ones(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} MiB")
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: there is no multi-GPU path (the search would shard by query "
              "chunk with no exchange during compute).", file=sys.stderr)
        return 2

    if args.serve:
        return serve_bench(args)

    from clast_amd import native

    native.init(0)
    p = native.Params()
    t_ram = t_vram = 64e6
    q_ram, q_vram = 64e6, 2e6
    t0 = time.perf_counter()
    t, tchunks = build_target(args.workload, args.scale, p, t_ram, t_vram)
    p.total_db_size = float(t.bases.shape[0])  # all sequence-line bytes of the target FASTA
    r, qchunks = build_queries(args.workload, args.scale, p, t, q_ram, q_vram)
    log(f"generated {len(t)} targets ({t.bases.shape[0] / 1e6:.0f} Mbp), {len(r)} reads in "
        f"{time.perf_counter() - t0:.1f} s; {len(tchunks)} target chunks, {len(qchunks)} query chunks")

    # resident index: every window of every target chunk, (chunk, window) order
    t1 = time.perf_counter()
    indexes = []
    n_twin = sum(len(w) for _, w in tchunks)
    if args.index == "merged":
        indexes.append(native.Index.merged(p, tchunks))
    else:
        for lst, wins in tchunks:
            indexes.append(native.Index(p, lst, wins))
    index_s = time.perf_counter() - t1
    log(f"index built in {index_s:.2f} s over {n_twin} target windows")
    queries = [native.Query(p, lst, wins) for lst, wins, _ in qchunks]
    total_bases = sum(b for _, _, b in qchunks)
    log(f"{len(queries)} query chunks resident, {total_bases / 1e6:.1f} Mbp")

    jobs = [(q, ix) for q in queries for ix in indexes]

    def run_jobs(mine, record):
        # each search's rows reach host memory while the next search runs;
        # a search's handle is released (after its copy) once the next one
        # has been issued
        rows = 0
        prev = None
        for q, ix in mine:
            h = native.search_batch(p, ix, q)
            rows += len(h)
            if record is not None:
                record.append((h.times(), h.counts()))
            if prev is not None:
                prev.close()
            prev = h
        if prev is not None:
            prev.close()
        return rows

    def step(record, streams):
        # the searches of a step dealt round-robin to `streams` host threads,
        # each with its own library streams, so searches overlap on the GPU;
        # the step ends with every copy done
        if streams <= 1:
            return run_jobs(jobs, record)
        parts = [jobs[k::streams] for k in range(streams)]
        with ThreadPoolExecutor(max_workers=streams) as ex:
            return sum(ex.map(lambda part: run_jobs(part, None), parts))

    for w in range(args.warmup):
        step(None, args.streams)
        log(f"warmup {w + 1}/{args.warmup} done")
    step_ms = []
    rows = 0
    for k in range(args.steps):
        ts = time.perf_counter()
        rows = step(None, args.streams)
        step_ms.append((time.perf_counter() - ts) * 1e3)
        log(f"step {k + 1}/{args.steps}: {step_ms[-1]:.1f} ms, {rows} hits ({args.streams} streams)")
    ms = float(np.mean(step_ms))
    value = total_bases / 1e6 / (ms / 1e3)
    # per-stage kernel times and work counters: one more step with one
    # search at a time, so each kernel's duration is its own
    rec = []
    step(rec, 1)
    # latency of one query chunk alone: call to rows in host memory
    lat = []
    for q in queries:
        for ix in indexes:
            ts = time.perf_counter()
            h = native.search_batch(p, ix, q)
            h.wait()
            lat.append((time.perf_counter() - ts) * 1e3)
            h.close()
    lat = np.array(lat)
    times = np.array([x[0] for x in rec])  # per search: lookup, segments, sort, filt, dp, res, total, raw seeds
    # raw seeds, DP hits, DP cells, rows, 16-bit DP tasks, their cells, long
    # DP tasks (one lane per band), their cells
    counts = np.array([x[1] for x in rec], dtype=np.float64)
    # k_seed_count algorithmic bytes per launch: 16 B per query k-mer position
    # (8 B index key lookup + 8 B cell (offset, count)); positions per launch
    # are taken from the query chunks
    npos = []
    for lst, wins, _ in qchunks:
        ln = lst.length.astype(np.int64)
        npos.append(int(np.clip(ln - p.lmer + 1, 0, None).sum()))
    n_idx = len(indexes)
    pos_per_launch = np.repeat(np.array(npos, dtype=np.float64), n_idx)
    count_ms = times[:, 0].astype(np.float64)
    bytes_count = 16.0 * pos_per_launch
    achieved = float(bytes_count.sum() / (count_ms.sum() / 1e3) / 1e9)
    traffic = measured_traffic("k_seed_count") if args.workload == "large" and args.scale == 1.0 else None
    # per-stage work against MI355X peaks (algorithmic bytes / ops): lookup
    # 16 B per query k-mer; segments 24 B per raw seed
    # (expansion) + 32 B (sort) + 16 B (isolate/duplicate pass) + 16 B per hit
    # out; DP 14 int32 ops per cell; results 40 B per hit in + 48 B per row out
    # (a stage that did not run, e.g. no seeds, reports a rate of 0)
    sec = lambda col: max(float(times[:, col].sum()) / 1e3, 1e-12)  # noqa: E731
    seeds_raw, dp_hits, cells, out_rows = (float(counts[:, c].sum()) for c in range(4))
    seg_bytes = 72.0 * seeds_raw + 16.0 * dp_hits
    res_bytes = 40.0 * dp_hits + 48.0 * out_rows
    stages = {
        "seed_lookup": {"GB/s": round(bytes_count.sum() / sec(0) / 1e9, 1), "peak": HBM_PEAK_GBS},
        "seed_segments": {"GB/s": round(seg_bytes / sec(1) / 1e9, 1), "peak": HBM_PEAK_GBS,
                          "raw_seeds_per_search": int(seeds_raw / len(counts))},
        "dp": {"GCUPS": round(cells / sec(4) / 1e9, 1), "Tops/s": round(14.0 * cells / sec(4) / 1e12, 2),
               "peak_Tops/s": VALU_INT32_PEAK_TOPS, "hits_per_search": int(dp_hits / len(counts)),
               "cells_per_search": int(cells / len(counts)),
               "cells_16bit_frac": round(float(counts[:, 5].sum()) / max(cells, 1.0), 4),
               "cells_band_frac": round(float(counts[:, 7].sum()) / max(cells, 1.0), 4)},
        "results": {"GB/s": round(res_bytes / sec(5) / 1e9, 1), "peak": HBM_PEAK_GBS},
    }
    stage_names = ["seed_lookup", "seed_segments", "global_seed_sort", "global_seed_filters", "dp", "results",
                   "search_total"]
    stage_ms = {n: round(float(times[:, i].sum()), 3) for i, n in enumerate(stage_names)}
    log(f"stage ms per step (one search at a time): {stage_ms}")

    out = {
        "metric": "query Mbp/s aligned (bit-exact hits) + p50/p99 request latency, 1 MI355X",
        "value": round(value, 3),
        "unit": "query Mbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {t.bases.shape[0] / 1e9:.3g} Gbp target ({len(t)} seqs) x {len(r)} reads "
                        f"({total_bases / 1e6:.0f} Mbp), reference defaults, batched resident index",
            "target_windows": n_twin, "query_chunks": len(queries),
            "query_windows": sum(len(w) for _, w, _ in qchunks), "hits": rows,
            "streams": args.streams,
        },
        "latency_ms": {"unit": "one query chunk (~200k reads) alone, call to rows in host memory, searched "
                               "against the resident index of every target window",
                       "p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "stage_ms_per_step": stage_ms,
        "stages": stages,
        "index_build_s": round(index_s, 3),
        "roofline": {"kernel": "k_seed_count", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
    }
    if args.plain:
        tp = time.perf_counter()
        for (lst, wins, _) in qchunks[:1]:
            for (tl, twins) in tchunks:
                for a, b in twins:
                    ix = native.Index(p, tl, [(a, b)])
                    for qa, qb in wins:
                        q = native.Query(p, lst, [(qa, qb)])
                        native.search(p, ix, q, a, 2 * qa).close()
                        q.close()
                    ix.close()
        plain_s = time.perf_counter() - tp
        out["plain_value"] = round(qchunks[0][2] / 1e6 / plain_s, 3)
    if not args.no_full_checks:
        t2 = time.perf_counter()
        out["full_size_checks"] = full_size_checks(p, args, tchunks, qchunks, indexes, queries)
        log(f"full-size checks in {time.perf_counter() - t2:.1f} s: {out['full_size_checks']}")
    if not args.no_cpu_baseline:
        cb, parity = cpu_baseline(p, tchunks, qchunks, n_twin, args.cpu_threads)
        out["cpu_baseline"] = cb
        out["parity_sample"] = parity
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
