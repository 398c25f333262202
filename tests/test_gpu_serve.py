"""Serving: coalesced requests return the same rows as each request alone,
and those rows are the oracle's for that request."""
from __future__ import annotations

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
