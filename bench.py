#!/usr/bin/env python3
"""Benchmark of the CLAST search path on one MI355X.

Headline (the serving configuration): a 3 Gbp target (3,000 x 1 Mbp, 40%
repeats from 1,000 families at 5-25% divergence, 1% (AT)n / (CA)n tracts)
indexed once and kept resident in HBM; requests of 1-10,000 reads
(log-uniform count; lengths half 150 bp, half log-uniform 1-15 kbp; 4%
substitutions, 8% indels) arrive as a Poisson process at --rate requests per
second and are served by continuous batching, admitted by a device-memory
budget (clast_amd/serve.py).  Reference defaults: lMer 15, stride 5, repeat
20, width 100, gap 8, global, cutOff 10, query windows of 2 MBase.

A step is one second of arrivals (--rate requests); W warmup steps are
served first, then K timed steps as one continuous stream.  value = query
Mbp (one strand, as read) served per second of wall time from the first
arrival to the last request's rows in host memory; latency = arrival to rows
in host memory per request (p50 / p99 over all K x rate requests).  Every
request's time includes uploading its reads and building its strands on the
device (clast_query_create_reads) and copying its rows back.

Also reported (untimed checks and extra legs):
* full_size_checks: every served row through the size-independent checks
  (clast_amd/checks.py); mix_equals_alone: a sample of requests searched
  again alone, rows equal;
* capacity: the same requests queued at once (served Mbp/s with no arrival
  gaps);
* roofline: the kernel with the most device time in the timed run, its
  algorithmic work per launch over its hipEvent-timed launch durations;
* cpu_baseline: the oracle (built -march=native on this host, OpenMP) on a
  bounded sample of the same requests against one target window, scaled by
  the number of target windows; the GPU searches the same unit and must match
  it bit-exactly (parity_sample);
* pressure_sweep: the seed-pressure stream (50% random, 25% repeat-family,
  25% reverse-complement reads of 100 bp-20 kbp) at 10-200 requests/s, each
  request's rows equal at every rate;
* large: the 1 Gbp target with 1,000,000 reads of 150 bp in 5 query RAM
  chunks, each step uploading and encoding every chunk's reads on the device
  and copying the rows back, from 3 host threads;
* uneven_batched_vs_plain: long reads (100 bp-20 kbp) searched the
  reference's way, one launch chain per (target window, query window) unit,
  and batched, over the same timed span, rows equal.
"""
from __future__ import annotations

import argparse
from concurrent.futures import ThreadPoolExecutor
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# hardware queues per process (HIP reads this at its first call): each device
# thread of the server issues on three streams (compute, side, copy), so three
# threads want nine queues; with HIP's default of four, streams share queues
# and their kernels wait behind each other (serving p50 about 20% higher).
# Raised to 16 when the environment asks for fewer.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or "4") < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0
CLOCK_GHZ = 2.4
N_CU = 256
# VALU peaks at the 2.4 GHz engine clock: 64 int32 lane-ops per CU per clock,
# 128 16-bit lane-ops when packed (two per 32-bit lane)
VALU_INT32_PEAK_TOPS = N_CU * 64 * CLOCK_GHZ / 1e3
VALU_INT16_PEAK_TOPS = N_CU * 128 * CLOCK_GHZ / 1e3
# int32 operations per DP cell: vertical and horizontal gap terms (2 adds +
# max each), diagonal (add of the substitution score), max(v, h), max with
# the diagonal, and for the match count 2 compares, 2 selects and an add;
# the score pass of a two-pass DP stage computes no match counts: 9
OPS_PER_CELL = 14
OPS_PER_CELL_SCORE = 9
METRIC = "query Mbp/s aligned (bit-exact hits) + p50/p99 request latency, 1 MI355X"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(name: str, kernel: str, cells_per_launch: float = 0.0):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE scaled by a
    plain device copy's counters) from profiles/<name>_hbm_traffic.txt: a
    per-launch figure, or a per-DP-cell one times this run's cells per
    launch (serving launches differ in size from run to run)."""
    path = os.path.join(ROOT, "profiles", f"{name}_hbm_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                part = line.split()
                if len(part) == 3 and part[1] == kernel:
                    if part[0] == "traffic_bytes_per_launch":
                        return float(part[2])
                    if part[0] == "traffic_bytes_per_cell" and cells_per_launch > 0:
                        return round(float(part[2]) * cells_per_launch)
    except OSError:
        pass
    return None


# ---------------------------------------------------------------- serving

def serving_target():
    from clast_amd import configs, gen, native

    tspec, _ = gen.workload("serving")
    t = gen.make_target(tspec)
    p = native.Params(total_db_size=float(t.bases.shape[0]))
    tch = configs.target_chunks(t, p)
    return t, p, tch


def request_checks(p, reqs, tch) -> dict:
    from clast_amd import checks

    wchunk = np.array([c for c, (_, w) in enumerate(tch) for _ in w], dtype=np.int64)
    tlens = [lst.length for lst, _ in tch]
    res = []
    for r in reqs:
        rows = r.rows
        n = len(rows["tID"])
        qlens = np.repeat(np.diff(r.offsets), 2)
        if n == 0:
            res.append(checks.check_rows(p, rows, np.zeros(0), np.zeros(0)))
            continue
        ch = wchunk[rows["twin"]]
        tl = np.empty(n, np.int64)
        for c in np.unique(ch):
            sel = ch == c
            tl[sel] = tlens[c][rows["tID"][sel]]
        res.append(checks.check_rows(p, rows, qlens[rows["qID"]], tl))
    m = checks.merge(res)
    m["passed"] = checks.passed(m)
    return m, res


def outside_row_vs_oracle(p, tch, reqs, res) -> dict:
    """The smallest request holding a row that runs past an end of its
    target sequence (a row the reference's rules allow, counted by the
    checks): that request's rows in that target window against the oracle."""
    from clast_amd import configs, native

    cand = [(len(r.offsets), k) for k, (r, c) in enumerate(zip(reqs, res)) if c["rows_outside_target"]]
    if not cand:
        return {"checked": False, "reason": "no such row"}
    k = min(cand)[1]
    req, i = reqs[k], res[k]["first_outside_target"]
    twin = int(req.rows["twin"][i])
    off = 0
    for lst, wins in tch:
        if twin < off + len(wins):
            tl, tw = lst, wins[twin - off]
            break
        off += len(wins)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    ql = native.build_list(req.bases, req.offsets, p.lmer, True)
    qwins = native.windows(ql.gateway, configs.Q_VRAM, True)
    exp = oracle.search(p, tl, [tw], ql, qwins)["final"]
    mine = {key: v[req.rows["twin"] == twin] for key, v in req.rows.items()}
    row = {key: int(req.rows[key][i]) for key in ("tID", "tIdx", "tHL", "qID", "qIdx", "qHL", "score")}
    return {"checked": True, "request_reads": len(req.offsets) - 1, "target_window": twin, "row": row,
            "target_length": int(tl.length[row["tID"]]), "rows": int(len(exp["tID"])),
            "bit_exact": rows_equal(mine, exp, windows=False)}


def rows_equal(a: dict, b: dict, windows: bool = True) -> bool:
    """Rows equal (integers exactly, E-values within 1e-9 relative), with
    the target window column unless windows is False."""
    from clast_amd import native

    keys = native.HIT_FIELDS + (("twin",) if windows else ())
    if len(a["tID"]) != len(b["tID"]):
        return False
    if not all(np.array_equal(a[k], b[k]) for k in keys):
        return False
    if len(a["tID"]) == 0:
        return True
    rel = np.abs(a["evalue"] - b["evalue"]) / np.maximum(np.abs(b["evalue"]), 1e-300)
    return bool(np.all(rel <= 1e-9))


def kernel_roofline(records, name: str) -> tuple[dict, dict]:
    """Per-kernel device time over the timed batches and the roofline of the
    kernel with the most time."""
    kt = np.array([r[2] for r in records], dtype=np.float64).reshape(-1, 8)
    dpk = np.array([r[4] for r in records], dtype=np.float64).reshape(-1, 12)
    pos = np.array([r[1] for r in records], dtype=np.float64)
    # each DP kernel's own cells over its own event-timed span (the
    # past-the-end k_dp16 launch, k_dp and k_dp_band run beside k_dp16 on a
    # second stream)
    ops16 = np.where(dpk[:, 11] > 0, OPS_PER_CELL_SCORE, OPS_PER_CELL)
    cells = {"k_dp16<8>": dpk[:, 0], "k_dp16<8,edge>": dpk[:, 1], "dp_second_pass": dpk[:, 9],
             "k_dp<8,false>": dpk[:, 2], "k_dp_band<8,false>": dpk[:, 3]}
    kernels = {
        # name: (ms, work, unit, peak, bound)
        "k_seed_count": (kt[:, 0], 16.0 * pos, "GB/s", HBM_PEAK_GBS, "hbm"),
        "k_dp16<8>": (dpk[:, 4], ops16 * dpk[:, 0], "Tops/s", VALU_INT16_PEAK_TOPS, "valu_int16"),
        "k_dp16<8,edge>": (dpk[:, 5], ops16 * dpk[:, 1], "Tops/s", VALU_INT16_PEAK_TOPS, "valu_int16"),
        "dp_second_pass": (dpk[:, 10], OPS_PER_CELL * dpk[:, 9], "Tops/s", VALU_INT16_PEAK_TOPS, "valu_int16"),
        "k_dp<8,false>": (dpk[:, 6], OPS_PER_CELL * dpk[:, 2], "Tops/s", VALU_INT32_PEAK_TOPS, "valu_int32"),
        "k_dp_band<8,false>": (dpk[:, 7], OPS_PER_CELL * dpk[:, 3], "Tops/s", VALU_INT32_PEAK_TOPS, "valu_int32"),
    }
    per = {}
    for k, (ms, work, unit, peak, bound) in kernels.items():
        launched = ms > 0
        tot_ms = float(ms.sum())
        per[k] = {"ms": round(tot_ms, 3), "launches": int(launched.sum()),
                  "avg_us": round(1e3 * tot_ms / max(int(launched.sum()), 1), 2)}
        if tot_ms > 0:
            scale = 1e9 if unit == "GB/s" else 1e12
            per[k][unit] = round(float(work[launched].sum()) / (tot_ms / 1e3) / scale, 3)
    per["dp_stage"] = {"ms": round(float(kt[:, 4].sum()), 3)}
    per["seed_generation"] = {"ms": round(float(kt[:, 5].sum()), 3)}
    per["results"] = {"ms": round(float(kt[:, 6].sum()), 3)}
    per["search_total"] = {"ms": round(float(kt[:, 7].sum()), 3)}
    dom = max(kernels, key=lambda k: per[k]["ms"])
    ms, work, unit, peak, bound = kernels[dom]
    launched = ms > 0
    scale = 1e9 if unit == "GB/s" else 1e12
    achieved = float(work[launched].sum()) / (float(ms.sum()) / 1e3) / scale if ms.sum() > 0 else 0.0
    roof = {"kernel": dom, "bound": bound, "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": unit,
            "frac": round(achieved / peak, 4), "launches": int(launched.sum()),
            "avg_launch_us": per[dom]["avg_us"],
            "traffic": measured_traffic(name, dom.split("<")[0],
                                        float(cells[dom][launched].sum()) / max(int(launched.sum()), 1)
                                        if dom in cells else 0.0),
            "work_per_unit": ("16 B per query k-mer position" if unit == "GB/s" else
                              f"{OPS_PER_CELL} int32-equivalent ops per DP cell (band width x rows), "
                              f"{OPS_PER_CELL_SCORE} in the score pass of a two-pass DP stage"),
            "two_pass_searches": int((dpk[:, 11] > 0).sum()), "searches": int(dpk.shape[0])}
    if bound == "valu_int16":
        roof["peak_note"] = (f"packed 16-bit VALU: {N_CU} CU x 128 lane-ops per clock at {CLOCK_GHZ} GHz; "
                             f"the int32 peak is {VALU_INT32_PEAK_TOPS:.1f}")
    return roof, per


def cpu_baseline_serving(p, tch, reqs, sample_mbp: float, threads: int):
    """The oracle on the first requests up to sample_mbp of reads against
    target window 0 (its table built first, timed apart), scaled by the
    number of target windows; the GPU searches the same unit."""
    from clast_amd import configs, native

    build = "build_native"
    try:
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "native"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, timeout=300)
    except (subprocess.SubprocessError, OSError) as e:
        log(f"native oracle build failed ({e}); timing the portable build")
        build = "build"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.use_build(build)
    sel, mbp = [], 0.0
    for r in reqs:
        sel.append(r)
        mbp += r.n_bases / 1e6
        if mbp >= sample_mbp:
            break
    bases = np.concatenate([r.bases[r.offsets[0]:r.offsets[-1]] for r in sel])
    offs = [np.zeros(1, np.int64)]
    for r in sel:
        offs.append((r.offsets[1:] - r.offsets[0]).astype(np.int64) + offs[-1][-1])
    offs = np.concatenate(offs)
    ql = native.build_list(bases, offs, p.lmer, True)
    qwins = native.windows(ql.gateway, configs.Q_VRAM, True)
    tl, twins = tch[0]
    tw = twins[0]
    n_twin = sum(len(w) for _, w in tch)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = oracle.search(p, tl, [tw], ql, qwins)
    exp = o["final"]
    ix = native.Index(p, tl, [tw])
    q = native.Query.from_reads(p, bases, offs, qwins)
    got = native.search_batch(p, ix, q).arrays()
    ix.close()
    q.close()
    exact = rows_equal(got, exp)
    value = mbp / (o["search_s"] * n_twin)
    sample = (f"oracle ({build}, {threads} OpenMP threads) on {len(sel)} requests ({len(offs) - 1} reads, "
              f"{mbp:.1f} Mbp) x target window 0 of {n_twin} ({tw[1] - tw[0]} seqs; table built in "
              f"{o['index_s']:.1f} s, not counted), searched in {o['search_s']:.1f} s; Mbp/s scaled by "
              f"{n_twin} target windows")
    rel = 0.0
    if exact and len(exp["tID"]):
        rel = float(np.max(np.abs(got["evalue"] - exp["evalue"]) / np.maximum(exp["evalue"], 1e-300)))
    return ({"value": round(value, 5), "unit": "query Mbp/s aligned", "cores": threads,
             "nproc": os.cpu_count(), "kind": "port", "sample": sample},
            {"rows": int(len(exp["tID"])), "bit_exact": bool(exact), "evalue_max_rel": rel})


def pressure_sweep(p, ix, t, rates, n_requests: int, workers: int, part_positions: int) -> dict:
    """The seed-pressure request stream at each arrival rate: the same
    requests, arrival gaps scaled; each request's rows must be equal at
    every rate (bit-exact under any admission pattern)."""
    from clast_amd import configs, gen, serve

    _, rspec = gen.workload("pressure")
    base = serve.make_requests(rspec, t, n_requests, 1.0, seed=5)
    arrivals = np.array([r.arrival for r in base])
    mbp = sum(r.n_bases for r in base) / 1e6
    ref = None
    points = []
    exact = True
    for rate in rates:
        reqs = [serve.Request(r.rid, r.bases, r.offsets, float(a / rate)) for r, a in zip(base, arrivals)]
        srv = serve.Server(p, ix, q_vram=configs.Q_VRAM, workers=workers, part_positions=part_positions)
        st = srv.run(reqs, realtime=True)
        if ref is None:
            ref = [r.rows for r in reqs]
        else:
            exact &= all(rows_equal(r.rows, e) for r, e in zip(reqs, ref))
        points.append({"rate_req_s": rate, "offered_mbp_s": round(mbp * rate / max(arrivals[-1], 1e-9), 2),
                       "served_mbp_s": st["mbp_per_s"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                       "batches": st["batches"], "mean_batch_parts": st["mean_batch_parts"],
                       "wall_s": st["wall_s"]})
        log(f"pressure {rate} req/s: {points[-1]}")
    return {"requests": n_requests, "query_mbp": round(mbp, 2),
            "rows": int(sum(len(r["tID"]) for r in ref)), "rows_equal_at_every_rate": bool(exact),
            "points": points}


# ---------------------------------------------------------------- large (1 Gbp x 1 M reads)

def stage_rates(kt, cnt, dpk, npos) -> dict:
    """Each stage of one search at a time against its bound, summed over the
    searches, with algorithmic work by the conventions of the hot-path
    table: seed lookup 16 B per query k-mer position; seed generation
    (expansion, sort and the isolate / duplicate compaction, fused per
    strand) 24 B per raw seed + 32 B per seed sorted + 16 B in and 16 B out
    per seed filtered (raw seeds stand for the seeds past the corner filter,
    which are fewer, so the figure is an upper bound) + 16 B per hit out;
    results (E-value, bad-hit filters, dedup, order) 40 B in and 40 B out
    per hit; DP as cells x 14 int32-equivalent ops (9 in the score pass of a
    two-pass stage, whose second pass's cells count 14) against the packed
    16-bit VALU peak."""
    raw, hits = float(cnt[:, 0].sum()), float(cnt[:, 1].sum())
    ops = float((cnt[:, 2] * np.where(dpk[:, 11] > 0, OPS_PER_CELL_SCORE, OPS_PER_CELL)).sum()
                + OPS_PER_CELL * dpk[:, 9].sum())
    work = {"seed_lookup": (kt[:, 0], 16.0 * float(npos.sum()), "hbm"),
            "seed_generation": (kt[:, 5], (24.0 + 32.0 + 32.0) * raw + 16.0 * hits, "hbm"),
            "dp": (kt[:, 4], ops, "valu_int16"),
            "results": (kt[:, 6], 80.0 * hits, "hbm")}
    out = {}
    for k, (ms, w, bound) in work.items():
        sec = float(ms.sum()) / 1e3
        if bound == "hbm":
            rate = w / sec / 1e9 if sec > 0 else 0.0
            out[k] = {"ms": round(1e3 * sec, 3), "GB": round(w / 1e9, 3), "GB/s": round(rate, 1),
                      "frac": round(rate / HBM_PEAK_GBS, 4), "bound": bound}
        else:
            rate = w / sec / 1e12 if sec > 0 else 0.0
            out[k] = {"ms": round(1e3 * sec, 3), "Tops": round(w / 1e12, 3), "Tops/s": round(rate, 2),
                      "frac": round(rate / VALU_INT16_PEAK_TOPS, 4), "bound": bound}
    return out


def large_leg(steps: int, warmup: int, streams: int) -> dict:
    """Each step: every query RAM chunk's reads uploaded and encoded on the
    device (clast_query_create_reads), searched against the resident index
    of all 16 target windows, rows copied to host memory; chunks dealt to
    `streams` host threads.  Also timed with the query sets resident before
    the step (resident_value), and per kernel with one search at a time."""
    from clast_amd import configs, native

    t0 = time.perf_counter()
    t, r, p = configs.workload("large")
    tch = configs.target_chunks(t, p)
    rch = configs.read_chunks(r, p)
    log(f"large: generated in {time.perf_counter() - t0:.1f} s: {len(tch)} target chunks, {len(rch)} query chunks")
    t1 = time.perf_counter()
    ix = native.Index.merged(p, tch)
    index_s = time.perf_counter() - t1
    total = sum(c.n_bases for c in rch)

    def job(ch, record=None):
        q = native.Query.from_reads(p, ch.bases, ch.offsets, ch.wins)
        h = native.search_batch(p, ix, q)
        h.wait()
        n = len(h)
        if record is not None:
            record.append((h.kernel_times(), h.counts(), ch, h.dp_kernels()))
        h.close()
        q.close()
        return n

    def step(pool):
        return sum(pool.map(job, rch))

    with ThreadPoolExecutor(max_workers=streams) as pool:
        for _ in range(warmup):
            step(pool)
        ms = []
        for _ in range(steps):
            ts = time.perf_counter()
            rows = step(pool)
            ms.append((time.perf_counter() - ts) * 1e3)
        # the query sets resident before the step (the inputs-in-HBM figure)
        qs = [native.Query.from_reads(p, c.bases, c.offsets, c.wins) for c in rch]

        def job_res(q):
            h = native.search_batch(p, ix, q)
            h.wait()
            h.close()

        ms_res = []
        for _ in range(steps):
            ts = time.perf_counter()
            list(pool.map(job_res, qs))
            ms_res.append((time.perf_counter() - ts) * 1e3)
        for q in qs:
            q.close()
    rec = []
    for ch in rch:
        job(ch, rec)
    ix.close()
    kt = np.array([x[0] for x in rec], dtype=np.float64)
    cnt = np.array([x[1] for x in rec], dtype=np.float64)
    dpk = np.array([x[3] for x in rec], dtype=np.float64)
    npos = np.array([float(2 * np.clip(np.diff(x[2].offsets) - p.lmer + 1, 0, None).sum()) for x in rec])
    seed_ms = float(kt[:, 0].sum())
    achieved = 16.0 * npos.sum() / (seed_ms / 1e3) / 1e9
    dp16_ms = float(dpk[:, 4].sum())
    step_ms = float(np.mean(ms))
    return {
        "workload": f"large: {t.bases.shape[0] / 1e9:.3g} Gbp target ({len(t)} seqs, 10% repeats) x {len(r)} reads "
                    f"of 150 bp ({total / 1e6:.0f} Mbp), {len(rch)} query RAM chunks, "
                    f"{sum(len(c.wins) for c in rch)} query windows, {sum(len(w) for _, w in tch)} target windows",
        "value": round(total / 1e6 / (step_ms / 1e3), 2), "unit": "query Mbp/s", "ms_per_step": round(step_ms, 3),
        "steps": steps, "streams": streams, "rows": int(rows),
        "timed_span": "reads in host memory -> upload + device encoding -> search -> rows in host memory",
        "resident_value": round(total / 1e6 / (float(np.mean(ms_res)) / 1e3), 2),
        "resident_note": "query sets already in HBM when the step starts",
        "index_build_s": round(index_s, 3),
        "kernels_one_search_at_a_time_ms": {
            "k_seed_count": round(seed_ms, 3), "k_dp16": round(dp16_ms, 3),
            "k_dp_int32": round(float(kt[:, 2].sum()), 3), "dp_stage": round(float(kt[:, 4].sum()), 3),
            "seed_generation": round(float(kt[:, 5].sum()), 3), "results": round(float(kt[:, 6].sum()), 3),
            "search_total": round(float(kt[:, 7].sum()), 3)},
        "seed_lookup_roofline": {"kernel": "k_seed_count", "bound": "hbm", "achieved": round(achieved, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                                 "traffic": measured_traffic("large", "k_seed_count")},
        "dp16": {"GCUPS": round(float(dpk[:, 0].sum()) / (dp16_ms / 1e3) / 1e9, 1) if dp16_ms else 0.0,
                 "cells_16bit_frac": round(float(cnt[:, 5].sum()) / max(float(cnt[:, 2].sum()), 1.0), 4)},
        "stages": stage_rates(kt, cnt, dpk, npos),
    }


# ---------------------------------------------------------------- uneven: batched vs plain

def uneven_leg(n_chunks: int, streams: int) -> dict:
    """The long-read configuration (1 Gbp target, 200,000 reads of 100 bp -
    20 kbp log-uniform, 20% random), the first n_chunks query RAM chunks,
    timed from reads in host memory to rows in host memory two ways over
    the same resident target:
    * plain: the reference's schedule (CHostSchedular.cu:55-127), one launch
      chain per (target window, query window) unit: each unit's query
      window uploaded (clast_query_create) and searched alone against that
      window's index (clast_search), rows copied back, one unit at a time;
    * batched: per query chunk, the reads uploaded and encoded on the device
      once (clast_query_create_reads) and every unit searched in one launch
      chain against the merged index (clast_search_batch: persistent /
      autotuned DP, captured task ordering), chunks from `streams` threads.
    The plain rows, appended in (target window, query window) order and
    stable-sorted by (query pair, score descending) as fixResult does, must
    equal the batched rows."""
    from clast_amd import configs, native

    t, r, p = configs.workload("uneven")
    tch = configs.target_chunks(t, p)
    rch = configs.read_chunks(r, p)[:n_chunks]
    merged = native.Index.merged(p, tch)
    units = [(tl, (a, b), native.Index(p, tl, [(a, b)])) for tl, wins in tch for a, b in wins]
    mbp = sum(c.n_bases for c in rch) / 1e6
    # warm both paths (kernel loading, autotuner trials, graph capture)
    native.search_batch(p, merged, native.Query.from_reads(p, rch[0].bases, rch[0].offsets, rch[0].wins)).close()

    def plain_chunk(ch):
        parts = []
        for tl, (a, b), ixu in units:
            for qa, qb in ch.wins:
                q = native.Query(p, ch.lst, [(qa, qb)])
                h = native.search(p, ixu, q, a, 2 * qa)
                parts.append(h.arrays())
                h.close()
                q.close()
        cat = {k: np.concatenate([x[k] for x in parts]) for k in native.HIT_FIELDS + ("evalue",)}
        order = np.lexsort((-cat["score"].astype(np.int64), cat["qID"] // 2))
        return {k: v[order] for k, v in cat.items()}

    for ch in rch:
        _ = ch.lst  # host lists built before timing (the plain path uploads from them)
    ts = time.perf_counter()
    plain = [plain_chunk(ch) for ch in rch]
    plain_s = time.perf_counter() - ts

    def batched_chunk(ch):
        q = native.Query.from_reads(p, ch.bases, ch.offsets, ch.wins)
        h = native.search_batch(p, merged, q)
        rows = h.arrays()
        h.close()
        q.close()
        return rows

    ts = time.perf_counter()
    with ThreadPoolExecutor(max_workers=streams) as pool:
        batched = list(pool.map(batched_chunk, rch))
    batched_s = time.perf_counter() - ts
    equal = all(len(a["tID"]) == len(b["tID"]) and all(np.array_equal(a[k], b[k]) for k in native.HIT_FIELDS)
                for a, b in zip(plain, batched))
    merged.close()
    for _, _, ixu in units:
        ixu.close()
    nunits = len(units) * sum(len(c.wins) for c in rch)
    return {"workload": f"uneven: 1 Gbp target x the first {len(rch)} query RAM chunks of 200,000 reads "
                        f"(100 bp-20 kbp log-uniform, 20% random): {sum(len(c.offsets) - 1 for c in rch)} reads, "
                        f"{mbp:.1f} Mbp, {nunits} (target window, query window) units",
            "plain_mbp_s": round(mbp / plain_s, 3), "batched_mbp_s": round(mbp / batched_s, 3),
            "batched_over_plain": round(plain_s / batched_s, 2), "rows": int(sum(len(b["tID"]) for b in batched)),
            "rows_equal": bool(equal),
            "timed_span": "reads in host memory -> rows in host memory, target index resident in both"}


# ---------------------------------------------------------------- random-gather calibration

def calibrate_gather(n: int, reps: int = 10) -> dict:
    """The seed lookup's random loads measured apart from it: n uniform
    random 8-byte reads (torch's index_select over a 1-D int64 array, one
    dispatch each) from arrays of 128 MiB (the presence bitmap's size,
    resident in the 256 MB Infinity Cache), 2 GiB and 16 GiB (the direct
    cell table's size: its pages outrun the address-translation caches).
    Each dispatch also streams its index array (8 B per access) and writes
    its output (8 B per access).  Reported per array: the fastest and the
    mean of `reps` hipEvent-timed dispatches and random reads per second."""
    import torch

    dev = torch.device("cuda", 0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    out = {"accesses_per_dispatch": n}
    for name, log2_words in (("gather_128MiB", 24), ("gather_2GiB", 28), ("gather_16GiB", 31)):
        src = torch.zeros(1 << log2_words, dtype=torch.int64, device=dev)
        idx = torch.randint(0, 1 << log2_words, (n,), device=dev, generator=gen)
        for _ in range(2):
            torch.index_select(src, 0, idx)
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            torch.index_select(src, 0, idx)
            e1.record()
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        best = min(ms)
        out[name] = {"best_ms": round(best, 4), "mean_ms": round(float(np.mean(ms)), 4),
                     "G_reads_per_s": round(n / (best / 1e3) / 1e9, 2)}
        log(f"{name}: {out[name]}")
        del src, idx
        torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------- main

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed seconds of arrivals (rate requests each)")
    ap.add_argument("--warmup", type=int, default=3, help="warmup seconds of arrivals")
    ap.add_argument("--rate", type=float, default=50.0, help="requests per second")
    ap.add_argument("--workers", type=int, default=3, help="device threads draining the batch queue")
    ap.add_argument("--part-positions", type=int, default=16_000_000,
                    help="requests are cut into parts of at most this many query k-mer positions")
    ap.add_argument("--capacity-requests", type=int, default=300)
    ap.add_argument("--alone-sample", type=int, default=12)
    ap.add_argument("--cpu-sample-mbp", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=min(os.cpu_count() or 1, 16))
    ap.add_argument("--sweep-rates", default="10,25,50,100,200")
    ap.add_argument("--sweep-requests", type=int, default=80)
    ap.add_argument("--large-steps", type=int, default=3)
    ap.add_argument("--uneven-chunks", type=int, default=2)
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-checks", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-capacity", action="store_true")
    ap.add_argument("--dp-log", default="", metavar="PATH",
                    help="write every serving search's DP kernel figures (clast_hits_dp_kernels), in search "
                         "order, one JSON list a line (for matching a profile's dispatches; use --workers 1)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving headline (only the other legs, e.g. to profile them alone)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only copy MiB of device memory once (the known-size kernel that scales the HBM "
                         "traffic counters) and exit")
    ap.add_argument("--calibrate-gather", type=int, default=0, metavar="N",
                    help="only time N random 8-byte reads from arrays of 128 MiB, 2 GiB and 16 GiB "
                         "(the seed lookup's random loads) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.ones(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} MiB")
        return 0
    if args.calibrate_gather:
        print(json.dumps(calibrate_gather(args.calibrate_gather)), flush=True)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: there is no multi-GPU path (the search would shard by query "
              "request or by target window, with no exchange during compute).", file=sys.stderr)
        return 2

    from clast_amd import configs, gen, native, serve

    native.init(0)
    if args.no_serve:
        out = {"metric": METRIC, "value": None, "note": "serving headline skipped (--no-serve)"}
        if not args.no_large:
            out["large"] = large_leg(args.large_steps, 1, 3)
        if not args.no_uneven:
            out["uneven_batched_vs_plain"] = uneven_leg(args.uneven_chunks, 3)
        print(json.dumps(out), flush=True)
        return 0
    t0 = time.perf_counter()
    t, p, tch = serving_target()
    n_twin = sum(len(w) for _, w in tch)
    log(f"serving target: {t.bases.shape[0] / 1e9:.2f} Gbp, {len(t)} seqs, {n_twin} target windows "
        f"({time.perf_counter() - t0:.1f} s)")
    t1 = time.perf_counter()
    ix = native.Index.merged(p, tch)
    index_s = time.perf_counter() - t1
    log(f"index over {n_twin} target windows built in {index_s:.2f} s")
    _, rspec = gen.workload("serving")
    n_warm = int(round(args.warmup * args.rate))
    n_req = int(round(args.steps * args.rate))
    t2 = time.perf_counter()
    warm = serve.make_requests(rspec, t, max(n_warm, 1), args.rate, seed=100)
    reqs = serve.make_requests(rspec, t, n_req, args.rate, seed=1)
    mbp = sum(r.n_bases for r in reqs) / 1e6
    log(f"{n_warm} warmup + {n_req} timed requests ({mbp:.0f} Mbp timed) generated in "
        f"{time.perf_counter() - t2:.1f} s")
    dp_log = []
    if n_warm:
        srv = serve.Server(p, ix, q_vram=configs.Q_VRAM, workers=args.workers, part_positions=args.part_positions)
        ws = srv.run(warm, realtime=True, record=bool(args.dp_log))
        dp_log += [list(map(float, r[4])) for r in srv.records]
        log(f"warmup: {ws['requests']} requests, {ws['mbp_per_s']} Mbp/s, p99 {ws['p99_ms']} ms")
    srv = serve.Server(p, ix, q_vram=configs.Q_VRAM, workers=args.workers, part_positions=args.part_positions)
    stats = srv.run(reqs, realtime=True, record=True)
    if args.dp_log:
        dp_log += [list(map(float, r[4])) for r in srv.records]
        with open(args.dp_log, "w") as f:
            for row in dp_log:
                f.write(json.dumps(row) + "\n")
    log(f"timed: {stats}")
    wall = stats["wall_s"]
    roof, per_kernel = kernel_roofline(srv.records, "serving")
    cnt = np.array([r[3] for r in srv.records], dtype=np.float64).reshape(-1, 8)
    cells16_frac = float(cnt[:, 5].sum()) / max(float(cnt[:, 2].sum()), 1.0)
    out = {
        "metric": METRIC,
        "value": stats["mbp_per_s"],
        "unit": "query Mbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16" if cells16_frac >= 0.5 else "int32",
        "dp_cells_int16_frac": round(cells16_frac, 4),
        "dtype_note": "integer scores: DP cells whose range proof allows it (dp_cells_int16_frac of all cells) run "
                      "in packed 16-bit lanes with results identical to the int32 kernel's, the rest in int32",
        "data": "synthetic",
        "config": {
            "workload": f"serving: {t.bases.shape[0] / 1e9:.2f} Gbp target ({len(t)} seqs, 40% repeats, 1% low "
                        f"complexity) resident in HBM; {n_req} requests of 1-10,000 reads (150 bp / 1-15 kbp), "
                        f"Poisson {args.rate:g} req/s; reference defaults",
            "target_windows": n_twin, "requests": n_req, "query_mbp": round(mbp, 2), "workers": args.workers,
            "step": f"one second of arrivals ({args.rate:g} requests)",
        },
        "latency_ms": {"p50": stats["p50_ms"], "p99": stats["p99_ms"],
                       "unit": f"per request, arrival to rows in host memory ({n_req} requests)"},
        "serving": stats,
        "index_build_s": round(index_s, 3),
        "roofline": roof,
        "kernels": per_kernel,
        "dp_tuner": native.tuner_report().splitlines(),
    }
    if not args.no_checks:
        t3 = time.perf_counter()
        out["full_size_checks"], per_req = request_checks(p, reqs, tch)
        out["outside_target_row_vs_oracle"] = outside_row_vs_oracle(p, tch, reqs, per_req)
        order = sorted(range(len(reqs)), key=lambda k: len(reqs[k].offsets))
        pick = [order[int(i)] for i in np.linspace(0, len(order) - 1, min(args.alone_sample, len(order)))]
        same = 0
        for k in pick:
            alone = srv.search_alone(serve.Request(reqs[k].rid, reqs[k].bases, reqs[k].offsets))
            same += rows_equal(reqs[k].rows, alone)
        out["mix_equals_alone"] = {"requests": len(pick), "equal": same}
        log(f"checks in {time.perf_counter() - t3:.1f} s: {out['full_size_checks']}, alone {same}/{len(pick)}")
    if not args.no_capacity:
        cap = [serve.Request(r.rid, r.bases, r.offsets, 0.0) for r in reqs[:args.capacity_requests]]
        cs = serve.Server(p, ix, q_vram=configs.Q_VRAM, workers=args.workers,
                          part_positions=args.part_positions).run(cap, realtime=False)
        out["capacity"] = {"requests": cs["requests"], "query_mbp": cs["query_mbp"], "mbp_per_s": cs["mbp_per_s"],
                           "wall_s": cs["wall_s"], "batches": cs["batches"],
                           "note": "the first requests of the timed stream, all queued at once"}
        log(f"capacity: {out['capacity']}")
    if not args.no_cpu_baseline:
        cb, parity = cpu_baseline_serving(p, tch, reqs, args.cpu_sample_mbp, args.cpu_threads)
        out["cpu_baseline"] = cb
        out["parity_sample"] = parity
        log(f"cpu baseline: {cb}; parity {parity}")
    if not args.no_sweep:
        rates = [float(x) for x in args.sweep_rates.split(",") if x]
        out["pressure_sweep"] = pressure_sweep(p, ix, t, rates, args.sweep_requests, args.workers,
                                              args.part_positions)
    ix.close()
    if not args.no_large:
        out["large"] = large_leg(args.large_steps, 1, 3)
        log(f"large: {out['large']['value']} Mbp/s ({out['large']['resident_value']} resident)")
    if not args.no_uneven:
        out["uneven_batched_vs_plain"] = uneven_leg(args.uneven_chunks, 3)
        log(f"uneven: {out['uneven_batched_vs_plain']}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
