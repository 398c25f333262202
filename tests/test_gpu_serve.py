"""Serving: coalesced requests return the same rows as each request alone,
and those rows are the oracle's for that request; searches issued from
several host threads at once return the rows of the same searches run one
after another."""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from clast_amd import gen, native, serve
from common import FIELDS, assert_rows_equal, db_size, workload

pytestmark = pytest.mark.gpu


def test_mix_equals_alone_equals_oracle():
    native.init(0)
    t, r = workload(n_t=8, t_len=25_000, n_r=900, repeat_frac=0.2, random_frac=0.2, len_mode=gen.LEN_LOGUNIFORM,
                    read_len=100, len_b=2000, indel=0.03)
    p = native.Params(total_db_size=db_size(t))
    tl = native.build_list(t.bases, t.offsets, p.lmer, False)
    twins = native.windows(tl.gateway, 60_000, query=False)
    ix = native.Index(p, tl, twins)
    srv = serve.Server(p, ix, q_vram=40_000, max_positions=400_000)
    reqs = serve.poisson_requests(r.bases, r.offsets, rate=1e6, seed=3, min_reads=1, max_reads=300)
    assert len(reqs) >= 6
    print(f"serve {len(reqs)} requests over {len(twins)} target windows", flush=True)
    stats = srv.run(reqs, realtime=False)
    assert stats["batches"] < len(reqs)  # requests were coalesced
    mixed = [r.rows for r in reqs]
    for req, rows in zip(reqs, mixed):
        alone = srv.search_alone(serve.Request(req.rid, req.bases, req.offsets))
        assert_rows_equal(rows, alone, what=f"request {req.rid} mix vs alone")
        np.testing.assert_array_equal(rows["twin"], alone["twin"])
        ql = native.build_list(req.bases, req.offsets, p.lmer, True)
        qwins = native.windows(ql.gateway, 40_000, query=True)
        exp = oracle.search(p, tl, twins, ql, qwins)["final"]
        assert_rows_equal(rows, exp, what=f"request {req.rid} vs oracle")


def test_concurrent_threads_equal_single_thread():
    # Each host thread searches on its own library streams and scratch; four
    # threads issuing searches at once over one resident index (the same
    # query set twice, and two others) return the rows of the searches run
    # one after another, and those are the oracle's.
    native.init(0)
    t, r = workload(n_t=6, t_len=30_000, n_r=800, repeat_frac=0.1, len_mode=gen.LEN_LOGUNIFORM, read_len=100,
                    len_b=1500, indel=0.02, seed=21)
    p = native.Params(total_db_size=db_size(t))
    tl = native.build_list(t.bases, t.offsets, p.lmer, False)
    twins = native.windows(tl.gateway, 60_000, query=False)
    ix = native.Index(p, tl, twins)
    parts = [(0, 300), (300, 550), (550, 800)]
    qsets = []
    for a, b in parts:
        lo, hi = int(r.offsets[a]), int(r.offsets[b])
        ql = native.build_list(r.bases[lo:hi], (r.offsets[a:b + 1] - lo).astype(np.int64), p.lmer, True)
        qsets.append((ql, native.windows(ql.gateway, 40_000, query=True)))
    queries = [native.Query(p, ql, qw) for ql, qw in qsets]
    alone = [native.search_batch(p, ix, q).arrays() for q in queries]
    jobs = [0, 1, 2, 0] * 3
    print(f"concurrent: {len(jobs)} searches from 4 threads", flush=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        got = list(ex.map(lambda k: native.search_batch(p, ix, queries[k]).arrays(), jobs))
    for k, rows in zip(jobs, got):
        assert_rows_equal(rows, alone[k], what=f"query set {k} concurrent vs alone")
    for k, (ql, qw) in enumerate(qsets):
        exp = oracle.search(p, tl, twins, ql, qw)["final"]
        assert len(exp["tID"]) > 0
        assert_rows_equal(alone[k], exp, what=f"query set {k} vs oracle")
