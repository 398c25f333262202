"""Shared helpers of the tests: small seeded workloads, list/window
construction, and comparisons of GPU rows against the oracle's."""
from __future__ import annotations

import numpy as np

from clast_amd import gen, native

FIELDS = ("tID", "tIdx", "qID", "qIdx", "tHL", "qHL", "match", "score")


def workload(n_t=6, t_len=30_000, n_r=300, seed=7, n_rate=0.001, repeat_frac=0.0, read_len=150, sub=0.04,
             indel=0.01, random_frac=0.0, len_mode=gen.LEN_FIXED, len_b=150, lowc=0.0):
    t = gen.make_target(gen.TargetSpec(seed, n_t, t_len, n_rate=n_rate, repeat_frac=repeat_frac,
                                       n_families=8 if repeat_frac else 0, fam_min=200, fam_max=800,
                                       div_min=0.02, div_max=0.10, lowc_frac=lowc))
    r = gen.make_reads(gen.ReadSpec(seed + 100, n_r, len_mode=len_mode, len_a=read_len, len_b=len_b, sub_rate=sub,
                                    indel_rate=indel, random_frac=random_frac, repeat_read_frac=0.3), t)
    return t, r


def lists(t, r, lmer=15):
    return native.build_list(t.bases, t.offsets, lmer, False), native.build_list(r.bases, r.offsets, lmer, True)


def db_size(t) -> float:
    # every sequence line byte of the target FASTA (no '>' lines)
    return float(t.bases.shape[0])


def assert_rows_equal(got: dict, exp: dict, fields=FIELDS, evalue=True, what=""):
    n = len(exp["tID"])
    assert len(got["tID"]) == n, f"{what}: {len(got['tID'])} rows vs {n} expected"
    for k in fields:
        if k in exp:
            bad = np.nonzero(got[k] != exp[k])[0]
            assert bad.size == 0, f"{what}: column {k} differs at rows {bad[:10]}: {got[k][bad[:5]]} vs {exp[k][bad[:5]]}"
    if evalue and "evalue" in exp and "evalue" in got and n:
        e, g = exp["evalue"], got["evalue"]
        rel = np.abs(g - e) / np.maximum(np.abs(e), 1e-300)
        assert np.all((rel <= 1e-9) | ((e == 0) & (g == 0))), f"{what}: E-value rel err {rel.max()}"


def sorted_rows(d: dict, fields=FIELDS) -> np.ndarray:
    cols = [d[k] for k in fields if k in d]
    a = np.stack(cols, axis=1) if cols and len(cols[0]) else np.zeros((0, len(cols)), np.int32)
    order = np.lexsort(a.T[::-1]) if len(a) else np.arange(0)
    return a[order]
