"""Size-independent checks of result rows, for workloads too large to run
through the oracle row by row.

Each search's rows must be in the reference's output order (query pair
ascending, score descending: CHostResultHolder.cu:100-109 with
utilResultSorting.cuh:6-26), carry E = K * D * qLen * exp(-lambda * score)
(krnlCalculateEvalue.cu:18-21) within 1e-9 relative, pass the bad-hit
filters (E <= cutOff unless cutOff is -1, qHL >= lMer + stride:
CDeviceHitList_deleteBadHits.cuh:32-48), lie inside the query with a match
count no larger than either hit length, overlap their target sequence, and
hold no two neighbours equal in (window, tID, tIdx, qID, qIdx) (the
post-sort dedup, CDeviceHitList.cu:51-105, leaves none).

Two things the reference allows are counted, not checked: rows that share a
start with a row that is not their neighbour (nothing dedups them), and rows
that run past an end of their target sequence (start before 0, or end past
the length).  Those come from the global extensions, which must reach the
query's ends: a deletion step onto a position outside the sequence costs
only the gap (no substitution is scored for a deleted base,
krnlAlignment.cu:56-76), and on a read whose indels moved its path more than
`gap` off the seed diagonal (the corner filter only bounds the seed
diagonal) that can beat the -1000 'X' cells.
"""
from __future__ import annotations

import numpy as np


def check_rows(p, a: dict, qlen: np.ndarray, tlen: np.ndarray) -> dict:
    """a: one search's rows in output order; qlen / tlen: each row's query
    and target sequence lengths.  Returns the checks' outcomes."""
    n = len(a["tID"])
    out = {"rows": n, "output_order": True, "evalue_max_rel": 0.0, "bad_hit_filters": True,
           "within_sequences": True, "no_adjacent_duplicates": True, "rows_sharing_a_start": 0,
           "rows_outside_target": 0, "first_outside_target": -1}
    if n == 0:
        return out
    pair, sc = a["qID"] // 2, a["score"].astype(np.int64)
    out["output_order"] = bool(np.all((pair[1:] > pair[:-1]) | ((pair[1:] == pair[:-1]) & (sc[1:] <= sc[:-1]))))
    ql = np.asarray(qlen, dtype=np.float64)
    exp_e = p.K * p.total_db_size * ql * np.exp(-p.lam * sc.astype(np.float64))
    out["evalue_max_rel"] = float(np.max(np.abs(a["evalue"] - exp_e) / np.maximum(exp_e, 1e-300)))
    filt = bool(np.all(a["qHL"] >= p.lmer + p.stride))
    if p.cut_off != -1:
        filt &= bool(np.all(a["evalue"] <= p.cut_off))
    out["bad_hit_filters"] = filt
    tl = np.asarray(tlen, dtype=np.int64)
    ti = a["tIdx"].astype(np.int64)
    out["within_sequences"] = bool(np.all(
        (a["qIdx"] >= 0) & (a["qIdx"].astype(np.int64) + a["qHL"] <= qlen) & (ti < tl) & (ti + a["tHL"] > 0) &
        (a["match"] <= np.minimum(a["qHL"], a["tHL"]))))
    past = np.nonzero((ti < 0) | (ti + a["tHL"] > tl))[0]
    out["rows_outside_target"] = int(len(past))
    out["first_outside_target"] = int(past[0]) if len(past) else -1
    tw = a.get("twin", np.zeros(n, np.int32))
    tup = np.stack([tw, a["tID"], a["tIdx"], a["qID"], a["qIdx"]], axis=1).astype(np.int64)
    out["no_adjacent_duplicates"] = bool(not np.any(np.all(tup[1:] == tup[:-1], axis=1))) if n > 1 else True
    out["rows_sharing_a_start"] = int(n - len(np.unique(tup, axis=0)))
    return out


def merge(results: list[dict]) -> dict:
    """The checks of several searches as one record."""
    m = {"rows": 0, "output_order": True, "evalue_max_rel": 0.0, "bad_hit_filters": True,
         "within_sequences": True, "no_adjacent_duplicates": True, "rows_sharing_a_start": 0,
         "rows_outside_target": 0}
    for r in results:
        m["rows"] += r["rows"]
        m["rows_sharing_a_start"] += r["rows_sharing_a_start"]
        m["rows_outside_target"] += r["rows_outside_target"]
        m["evalue_max_rel"] = max(m["evalue_max_rel"], r["evalue_max_rel"])
        for k in ("output_order", "bad_hit_filters", "within_sequences", "no_adjacent_duplicates"):
            m[k] = m[k] and r[k]
    return m


def passed(m: dict) -> bool:
    return (m["output_order"] and m["bad_hit_filters"] and m["within_sequences"] and m["no_adjacent_duplicates"]
            and m["evalue_max_rel"] <= 1e-9)
