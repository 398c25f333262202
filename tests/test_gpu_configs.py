"""Every benchmark configuration at its own size, cut into the chunks and
windows the reference's driver would use (configs.py).

* small (10 Mbp x 10,000 reads): every row against the oracle, through the
  batched search, the per-unit search and the query strands built on the
  device.
* large (1 Gbp x 1,000,000 reads) and uneven (1 Gbp x 200,000 reads of
  100 bp - 20 kbp): every row of one full pass through the size-independent
  checks (checks.py), and two (target window x query window) units each
  against the oracle, row for row, both as the unit alone and as that unit's
  rows inside the batched pass.
* serving (3 Gbp resident, requests of 1 - 10,000 reads) and seed pressure
  (the same target, 50% random / 25% repeat / 25% reverse-complement reads):
  every request served in a mix, a sample of requests again alone (rows
  equal), every row through the checks, and one request's unit against the
  oracle.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import oracle
from clast_amd import checks, configs, gen, native, serve
from common import assert_rows_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    native.init(0)


def _say(msg):
    print(msg, flush=True)


def _tlen_table(tch):
    """(global window) -> target chunk, and each chunk's sequence lengths."""
    wchunk = np.array([c for c, (_, w) in enumerate(tch) for _ in w], dtype=np.int64)
    return wchunk, [lst.length for lst, _ in tch]


def _row_checks(p, rows, qlens, wchunk, tlens):
    n = len(rows["tID"])
    if n == 0:
        return checks.check_rows(p, rows, np.zeros(0), np.zeros(0))
    ch = wchunk[rows["twin"]]
    tl = np.empty(n, np.int64)
    for c in np.unique(ch):
        sel = ch == c
        tl[sel] = tlens[c][rows["tID"][sel]]
    return checks.check_rows(p, rows, qlens[rows["qID"]], tl)


def _unit_rows(rows, twin, qwin):
    sel = (rows["twin"] == twin) & (rows["qwin"] == qwin)
    return {k: v[sel] for k, v in rows.items()}


def _oracle_unit(p, tl, tw, ql, qw):
    return oracle.search(p, tl, [tw], ql, [qw])["final"]


def _unit_vs_oracle(p, tch, twin_global, chunk, qwin, batched_rows):
    """One unit: the GPU's per-unit search and the unit's rows inside the
    batched pass, both against the oracle."""
    off = 0
    for lst, wins in tch:
        if twin_global < off + len(wins):
            tl, tw = lst, wins[twin_global - off]
            break
        off += len(wins)
    ql, qw = chunk.lst, chunk.wins[qwin]
    exp = _oracle_unit(p, tl, tw, ql, qw)
    ix = native.Index(p, tl, [tw])
    q = native.Query(p, ql, [qw])
    got = native.search(p, ix, q, tw[0], 2 * qw[0]).arrays()
    ix.close()
    q.close()
    assert_rows_equal(got, exp, what=f"unit ({twin_global}, {qwin}) alone vs oracle")
    inside = _unit_rows(batched_rows, twin_global, qwin)
    assert_rows_equal(inside, exp, what=f"unit ({twin_global}, {qwin}) inside the batch vs oracle")
    return len(exp["tID"])


def test_config_small_every_row_vs_oracle():
    t, r, p = configs.workload("small")
    tch = configs.target_chunks(t, p)
    rch = configs.read_chunks(r, p)
    assert len(tch) == 1 and len(rch) == 1
    (tl, twins), ch = tch[0], rch[0]
    _say(f"small: {len(twins)} target window(s) x {len(ch.wins)} query windows")
    exp = oracle.search(p, tl, twins, ch.lst, ch.wins)["final"]
    assert len(exp["tID"]) > 9_000
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ch.lst, ch.wins)
    got = native.search_batch(p, ix, q).arrays()
    assert_rows_equal(got, exp, what="small batched vs oracle")
    np.testing.assert_array_equal(got["twin"], exp["twin"])
    qr = native.Query.from_reads(p, ch.bases, ch.offsets, ch.wins)
    got_r = native.search_batch(p, ix, qr).arrays()
    assert_rows_equal(got_r, exp, what="small, strands built on the device, vs oracle")
    for h in (q, qr):
        h.close()
    # the reference's own schedule: one launch chain per unit
    for a, (tb, te) in enumerate(twins):
        ixu = native.Index(p, tl, [(tb, te)])
        for k, (qa, qb) in enumerate(ch.wins):
            qu = native.Query(p, ch.lst, [(qa, qb)])
            rows = native.search(p, ixu, qu, tb, 2 * qa).arrays()
            assert_rows_equal(rows, _oracle_unit(p, tl, (tb, te), ch.lst, (qa, qb)), what=f"small unit ({a}, {k})")
            qu.close()
        ixu.close()
    ix.close()


def _full_pass(name, units):
    t, r, p = configs.workload(name)
    tch = configs.target_chunks(t, p)
    rch = configs.read_chunks(r, p)
    wchunk, tlens = _tlen_table(tch)
    _say(f"{name}: {len(t)} targets in {len(tch)} chunks ({len(wchunk)} windows), {len(r)} reads in "
         f"{len(rch)} chunks ({sum(len(c.wins) for c in rch)} query windows)")
    ix = native.Index.merged(p, tch)
    res = []
    kept = {}
    for k, ch in enumerate(rch):
        q = native.Query.from_reads(p, ch.bases, ch.offsets, ch.wins)
        rows = native.search_batch(p, ix, q).arrays()
        q.close()
        qlens = np.repeat(np.diff(ch.offsets), 2)
        res.append(_row_checks(p, rows, qlens, wchunk, tlens))
        if any(u[1] == k for u in units):
            kept[k] = rows
        _say(f"  chunk {k}: {len(rows['tID'])} rows")
    ix.close()
    m = checks.merge(res)
    _say(f"{name}: {m}")
    assert checks.passed(m), m
    assert m["rows"] > 0
    n_unit = 0
    for twin, chunk, qwin in units:
        n_unit += _unit_vs_oracle(p, tch, twin, rch[chunk], qwin, kept[chunk])
    assert n_unit > 0
    return m


def test_config_large_full_pass():
    # units: the first target window with the first query window, and a
    # middle target window with a middle query window of another chunk
    m = _full_pass("large", [(0, 0, 0), (7, 2, 17)])
    assert m["rows"] > 1_000_000


def test_config_uneven_full_pass():
    _full_pass("uneven", [(0, 0, 0), (9, 11, 5)])


@pytest.fixture(scope="module")
def big_target():
    """The 3 Gbp target of the serving workloads, indexed once."""
    tspec, _ = gen.workload("serving")
    t = gen.make_target(tspec)
    p = native.Params(total_db_size=float(t.bases.shape[0]))
    tch = configs.target_chunks(t, p)
    ix = native.Index.merged(p, tch)
    _say(f"3 Gbp target: {len(t)} sequences, {sum(len(w) for _, w in tch)} windows")
    yield t, p, tch, ix
    ix.close()


def _serve_case(big_target, name, n_reads, alone_every):
    t, p, tch, ix = big_target
    _, rspec = gen.workload(name)
    r = gen.make_reads(dataclasses.replace(rspec, n_reads=n_reads), t)
    # parts of at most 4 M positions: the large requests are split over
    # several batches, the alone runs below are not
    srv = serve.Server(p, ix, q_vram=configs.Q_VRAM, workers=2, part_positions=4_000_000)
    reqs = serve.poisson_requests(r.bases, r.offsets, rate=50.0, seed=11, min_reads=1, max_reads=10_000)
    _say(f"{name}: {len(reqs)} requests, {len(r)} reads ({r.bases.shape[0] / 1e6:.1f} Mbp)")
    stats = srv.run(reqs, realtime=False)
    _say(f"{name}: {stats}")
    wchunk, tlens = _tlen_table(tch)
    res = []
    for req in reqs:
        qlens = np.repeat(np.diff(req.offsets), 2)
        res.append(_row_checks(p, req.rows, qlens, wchunk, tlens))
    m = checks.merge(res)
    _say(f"{name}: {m}")
    assert checks.passed(m), m
    assert m["rows"] > 0
    for req in reqs[::alone_every]:
        alone = srv.search_alone(serve.Request(req.rid, req.bases, req.offsets))
        assert_rows_equal(req.rows, alone, what=f"{name} request {req.rid} mix vs alone")
        np.testing.assert_array_equal(req.rows["twin"], alone["twin"])
    # one small request's unit (the target window with most of its rows),
    # and the window of the first row that runs past an end of its target sequence (a
    # row the reference allows; counted above), each against the oracle
    small = min(reqs, key=lambda q: (len(q.offsets) > 200, -len(q.rows["tID"])))
    twin = int(np.bincount(small.rows["twin"]).argmax()) if len(small.rows["twin"]) else 0
    _request_window_vs_oracle(p, tch, small, twin, name)
    past = [(req, c["first_outside_target"]) for req, c in zip(reqs, res) if c["rows_outside_target"]]
    if past:
        req, i = min(past, key=lambda x: len(x[0].offsets))
        _say(f"{name}: request {req.rid} row {i} runs past an end of its target sequence: "
             f"{ {k: int(req.rows[k][i]) for k in ('twin', 'tID', 'tIdx', 'tHL', 'qID', 'qIdx', 'qHL', 'score')} }")
        _request_window_vs_oracle(p, tch, req, int(req.rows["twin"][i]), name)
    return m


def _request_window_vs_oracle(p, tch, req, twin, name):
    """A request's rows in one target window: the unit searched alone on the
    GPU and the request's rows in the mix restricted to that window, both
    against the oracle (request windows as the server cuts them)."""
    ql = native.build_list(req.bases, req.offsets, p.lmer, True)
    qwins = native.windows(ql.gateway, configs.Q_VRAM, True)
    off = 0
    for lst, wins in tch:
        if twin < off + len(wins):
            tl, tw = lst, wins[twin - off]
            break
        off += len(wins)
    exp = oracle.search(p, tl, [tw], ql, qwins)["final"]
    ixu = native.Index(p, tl, [tw])
    q = native.Query.from_reads(p, req.bases, req.offsets, qwins)
    got = native.search_batch(p, ixu, q).arrays()
    for h in (q, ixu):
        h.close()
    assert_rows_equal(got, exp, what=f"{name} request {req.rid} x target window {twin} vs oracle")
    mine = {k: v[req.rows["twin"] == twin] for k, v in req.rows.items()}
    assert_rows_equal(mine, exp, what=f"{name} request {req.rid} in the mix, window {twin}, vs oracle")


def test_config_serving(big_target):
    _serve_case(big_target, "serving", 12_000, 3)


def test_config_pressure(big_target):
    _serve_case(big_target, "pressure", 12_000, 3)
