"""clast_search_batch_noalloc through the C ABI: the rows of the batched
search (clast_search_batch) on two calls in a row, the second with a smaller
query set (the scratch then still holds the first call's contents), and the
documented refusals (too few result rows, a query set the scratch cannot
hold)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from clast_amd import gen, native
from common import assert_rows_equal, db_size, lists, workload

pytestmark = pytest.mark.gpu

CLAST_ERR_LIMIT = 4


def _equal(got, exp, what):
    assert_rows_equal(got, exp, what=what)
    for k in ("twin", "qwin"):
        np.testing.assert_array_equal(got[k], exp[k], err_msg=f"{what}: {k}")


def test_noalloc_rows_equal_batched_twice_second_smaller():
    native.init(0)
    t, r = workload(n_t=6, t_len=30_000, n_r=700, repeat_frac=0.1, len_mode=gen.LEN_LOGUNIFORM, read_len=100,
                    len_b=3000, indel=0.02, seed=31)
    p = native.Params(total_db_size=db_size(t))
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 60_000, query=False)
    qwins = native.windows(ql.gateway, 40_000, query=True)
    assert len(qwins) >= 4
    ix = native.Index(p, tl, twins)
    big = native.Query(p, ql, qwins)
    small = native.Query(p, ql, qwins[:2])
    ref_big = native.search_batch(p, ix, big).arrays()
    ref_small = native.search_batch(p, ix, small).arrays()
    assert len(ref_small["tID"]) < len(ref_big["tID"])
    print(f"noalloc: {len(ref_big['tID'])} then {len(ref_small['tID'])} rows", flush=True)
    rc, got = native.search_batch_noalloc(p, ix, big, len(ref_big["tID"]) + 5)
    assert rc == 0, native.last_error()
    _equal(got, ref_big, "first call (all query windows)")
    rc, got = native.search_batch_noalloc(p, ix, small, len(ref_small["tID"]))
    assert rc == 0, native.last_error()
    _equal(got, ref_small, "second call (two query windows)")
    exp = oracle.search(p, tl, twins, ql, qwins[:2])["final"]
    assert_rows_equal(got, exp, what="second call vs oracle")
    # too small a row capacity: refused, with the rows needed reported
    rc, got = native.search_batch_noalloc(p, ix, big, 3)
    assert rc == CLAST_ERR_LIMIT and got["needed"] == len(ref_big["tID"])
    # and the entry point still gives the rows afterwards
    rc, got = native.search_batch_noalloc(p, ix, small, len(ref_small["tID"]))
    assert rc == 0, native.last_error()
    _equal(got, ref_small, "after a refused call")
    for h in (big, small, ix):
        h.close()


def test_noalloc_refuses_a_query_set_over_the_scratch():
    # 1,700 reads of 10 kbp: 2 x 1,700 x 9,986 k-mer positions, 16 B each,
    # over the 256 MiB scratch; refused before any device work
    native.init(0)
    rng = np.random.default_rng(3)
    t = gen.Seqs(rng.choice(list(b"ACGT"), 50_000).astype(np.uint8), np.array([0, 50_000], np.int64))
    p = native.Params(total_db_size=50_000.0)
    tl = native.build_list(t.bases, t.offsets, p.lmer, False)
    ix = native.Index(p, tl, [(0, 1)])
    n, ln = 1_700, 10_000
    bases = rng.choice(list(b"ACGT"), n * ln).astype(np.uint8)
    offs = np.arange(n + 1, dtype=np.int64) * ln
    q = native.Query.from_reads(p, bases, offs, [(0, n)])
    assert 16 * 2 * n * (ln - p.lmer + 1) > native.NOALLOC_SCRATCH_BYTES
    rc, got = native.search_batch_noalloc(p, ix, q, 1000)
    assert rc == CLAST_ERR_LIMIT, (rc, native.last_error())
    assert "split it" in native.last_error()
    for h in (q, ix):
        h.close()
