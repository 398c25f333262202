"""Pins of the CPU oracle (no GPU): known answers, brute force on tiny inputs
and closed forms, each chosen so that a plausible slip in the oracle (a
swapped base code, an off-by-one window, a lost gap term, the wrong filter
predecessor) fails one of them."""
from __future__ import annotations

import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from clast_amd import gen, native
from common import lists, workload

P0 = native.Params()


# ---------------------------------------------------------------- k-mer code

def test_kmer_code_known_answers():
    assert oracle.kmer_code(b"A" * 15, 15) == 0
    assert oracle.kmer_code(b"T" * 15, 15) == 4 ** 15 - 1
    assert oracle.kmer_code(b"ACGT", 4) == 0b00011011  # MSB first, A0 C1 G2 T3
    assert oracle.kmer_code(b"CA", 2) == 4
    assert oracle.kmer_code(b"ACGN", 4) == -1
    assert oracle.kmer_code(b"acgt", 4) == -1  # lower case is not a base here
    assert oracle.kmer_code(b"ACGTACGTACGTACGTACGTACGTACGTACG", 31) == int("0123" * 7 + "012", 4)
    assert oracle.kmer_code(b"ACGT", 4, avail=3) == -1  # past the window end


# ---------------------------------------------------------------- index

def brute_index(p, tl, a, b):
    """Strided positions of window [a, b) grouped by k-mer, by brute force."""
    lo, hi = int(tl.gateway[a]), int(tl.gateway[b])
    win = tl.base[lo:hi].tobytes()
    n = len(win)
    cells = {}
    for i in range(n // p.stride):
        pos = i * p.stride
        s = win[pos:pos + p.lmer]
        ok = len(s) == p.lmer and all(c in b"ACGT" for c in s)
        key = int(s.translate(bytes.maketrans(b"ACGT", b"0123")), 4) if ok else -1
        cells.setdefault(key, []).append(pos)
    return cells


@pytest.mark.parametrize("lmer,stride,cut", [(15, 5, 20), (6, 2, 3), (8, 1, -1)])
def test_index_brute_force(lmer, stride, cut):
    t, r = workload(n_t=3, t_len=1500, n_r=1, n_rate=0.02, repeat_frac=0.3)
    p = native.Params(lmer=lmer, stride=stride, cut_repeat=cut)
    tl, _ = lists(t, r, lmer)
    ix = oracle.index(p, tl, 1, 3)
    cells = brute_index(p, tl, 1, 3)
    keys = sorted(cells)
    np.testing.assert_array_equal(ix["key"], keys)
    pre = [len(cells[k]) for k in keys]
    # offsets are taken before the repeat cut
    np.testing.assert_array_equal(ix["gateway_index"], np.concatenate([[0], np.cumsum(pre)[:-1]]))
    exp_size = [(c if (cut == -1 or c < cut) else 0) for c in pre]
    if keys[0] == -1:
        exp_size[0] = 0
    np.testing.assert_array_equal(ix["cell_size"], exp_size)
    np.testing.assert_array_equal(ix["positions"], np.concatenate([cells[k] for k in keys]))
    assert sum(pre) == (int(tl.gateway[3]) - int(tl.gateway[1])) // stride


# ---------------------------------------------------------------- seeds and filters

def test_seed_stages_brute_force():
    t, r = workload(n_t=3, t_len=3000, n_r=40, repeat_frac=0.2, sub=0.05)
    p = native.Params(total_db_size=9000.0, width=60)
    tl, ql = lists(t, r, p.lmer)
    o = oracle.search(p, tl, [(0, 3)], ql, [(0, 40)], dump=True)
    cells = brute_index(p, tl, 0, 3)
    tg = tl.gateway

    def tpos(pos):
        s = int(np.searchsorted(tg, pos, side="right")) - 1
        return s, int(pos - tg[s])

    qwin = ql.base[: int(ql.gateway[80])].tobytes()
    raw = []
    for i in range(len(qwin) - p.lmer + 1):
        s = qwin[i:i + p.lmer]
        if not all(c in b"ACGT" for c in s):
            continue
        key = int(s.translate(bytes.maketrans(b"ACGT", b"0123")), 4)
        c = cells.get(key, [])
        if len(c) >= p.cut_repeat:
            continue
        qs = int(np.searchsorted(ql.gateway, i, side="right")) - 1
        for pos in c:
            ts, ti = tpos(pos)
            raw.append((ts, ti, qs, i - int(ql.gateway[qs])))
    got_raw = list(zip(*[o["raw"][k].tolist() for k in ("tID", "tIdx", "qID", "qIdx")]))
    assert got_raw == raw
    # boundary: query end, target end, corner (global)
    L, g = p.lmer, p.gap
    bnd = []
    for ts, ti, qs, qi in raw:
        qlen, tlen = int(ql.length[qs]), int(tl.length[ts])
        if qi + L > qlen or ti + L > tlen:
            continue
        d = ti - qi
        if d + g < 0 or d + qlen - g > tlen:
            continue
        bnd.append((ts, ti, qs, qi))
    assert list(zip(*[o["boundary"][k].tolist() for k in ("tID", "tIdx", "qID", "qIdx")])) == bnd
    srt = sorted(bnd, key=lambda s: (s[2], s[0], s[1] - s[3], s[3]))
    assert list(zip(*[o["sorted"][k].tolist() for k in ("tID", "tIdx", "qID", "qIdx")])) == srt

    # isolate / duplicate as chains: a seed is "linked" to its predecessor when
    # both share (qID, tID) and lie within `width` on the target
    def linked(a, b):
        return a[2] == b[2] and a[0] == b[0] and abs(a[1] - b[1]) <= p.width and (b[1] - b[3]) - (a[1] - a[3]) <= g

    iso = [srt[0]] + [srt[i] for i in range(1, len(srt)) if linked(srt[i], srt[i - 1])] if len(srt) > 1 else srt
    dup = [iso[0]] + [iso[i] for i in range(1, len(iso)) if not linked(iso[i], iso[i - 1])] if len(iso) > 1 else iso
    assert list(zip(*[o["isolate"][k].tolist() for k in ("tID", "tIdx", "qID", "qIdx")])) == iso
    assert list(zip(*[o["duplicate"][k].tolist() for k in ("tID", "tIdx", "qID", "qIdx")])) == dup
    assert len(dup) < len(iso) < len(srt)


def test_isolate_duplicate_known_answer():
    """A hand-built chain layout on one query strand: the read carries three
    exact 15-mer matches to t0 close together (one chain), one far away
    (isolated), and two more close together (second chain)."""
    rng = np.random.default_rng(5)
    tgt = bytes(rng.choice(list(b"ACGT"), 20000).astype(np.uint8))
    # query strand '+' = concatenation of target pieces
    pieces = [tgt[1000:1015], tgt[1020:1035], tgt[1040:1055], tgt[9000:9015], tgt[15000:15015], tgt[15030:15045]]
    q = b"".join(pieces)
    t = gen.Seqs(np.frombuffer(tgt, np.uint8).copy(), np.array([0, len(tgt)], np.int64))
    r = gen.Seqs(np.frombuffer(q, np.uint8).copy(), np.array([0, len(q)], np.int64))
    p = native.Params(stride=1, local=1, cut_off=-1, width=100)
    tl, ql = lists(t, r, p.lmer)
    o = oracle.search(p, tl, [(0, 1)], ql, [(0, 1)], dump=True)
    srt = list(zip(o["sorted"]["tIdx"].tolist(), o["sorted"]["qIdx"].tolist()))
    exact = {(1000, 0), (1020, 15), (1040, 30), (9000, 45), (15000, 60), (15030, 75)}
    assert exact <= set(srt)
    iso = set(zip(o["isolate"]["tIdx"].tolist(), o["isolate"]["qIdx"].tolist()))
    dup = list(zip(o["duplicate"]["tIdx"].tolist(), o["duplicate"]["qIdx"].tolist()))
    # element 0 of the window survives the isolate filter; chain members with
    # a near predecessor survive; the lone seed at 9000 does not
    assert srt[0] == (1000, 0) and srt[0] in iso
    assert (1020, 15) in iso and (1040, 30) in iso and (15030, 75) in iso
    assert (9000, 45) not in iso
    # duplicate keeps one representative per chain: the chain's first
    # isolate survivor (element 0 for the first chain)
    chain1 = [s for s in srt if s[0] < 2000 and s in iso]
    chain2 = [s for s in srt if 14000 < s[0] < 16000 and s in iso]
    assert [s for s in dup if s[0] < 2000] == [(1000, 0)] == chain1[:1]
    assert [s for s in dup if 14000 < s[0] < 16000] == chain2[:1]
    assert (15000, 60) not in iso  # first seed of its chain: isolated


# ---------------------------------------------------------------- banded DP

def align(p, tgt: bytes, q: bytes, tIdx: int, qIdx: int):
    t = gen.Seqs(np.frombuffer(tgt, np.uint8).copy(), np.array([0, len(tgt)], np.int64))
    r = gen.Seqs(np.frombuffer(q, np.uint8).copy(), np.array([0, len(q)], np.int64))
    tl, ql = lists(t, r, p.lmer)
    return oracle.align_one(p, tl, (0, 1), ql, (0, 1), 0, tIdx, 0, qIdx)


RNG = np.random.default_rng(11)
TGT = bytes(RNG.choice(list(b"ACGT"), 3000).astype(np.uint8))


def test_dp_exact_match():
    q = TGT[1000:1150]
    for qidx in (0, 40, 135):
        assert align(P0, TGT, q, 1000 + qidx, qidx) == (1000, 0, 150, 150, 150, 150)


def test_dp_substitution_and_n():
    q = bytearray(TGT[1000:1150])
    q[100] = next(c for c in b"ACGT" if c != q[100])
    # one mismatch: 149 matches + (-3)
    assert align(P0, TGT, bytes(q), 1000, 0) == (1000, 0, 150, 150, 149, 146)
    q[100] = ord("N")  # N scores as a mismatch and counts no match
    assert align(P0, TGT, bytes(q), 1000, 0) == (1000, 0, 150, 150, 149, 146)
    q[100] = ord("E")  # the -1 class
    assert align(P0, TGT, bytes(q), 1000, 0) == (1000, 0, 150, 150, 148, 148)


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_dp_deletion_affine(k):
    # query misses k target bases after its first 100: one gap of length k
    q = TGT[1000:1100] + TGT[1100 + k:1150 + k]
    assert len(q) == 150
    got = align(P0, TGT, q, 1000, 0)
    assert got == (1000, 0, 150 + k, 150, 150, 150 + (-5) + k * (-2))


@pytest.mark.parametrize("k", [1, 3, 8])
def test_dp_insertion_affine(k):
    ins = bytes(c for c in (b"ACGT" * 4)[:k])
    q = TGT[1000:1100] + ins + TGT[1100:1150]
    got = align(P0, TGT, q, 1000, 0)
    t_idx, q_idx, thl, qhl, m, s = got
    assert (t_idx, q_idx, thl, qhl) == (1000, 0, 150, 150 + k)
    # best of: a k-gap, or absorbing inserted bases that happen to match
    assert s >= 150 - 5 - 2 * k and m >= 150


def test_dp_left_extension_and_band_floor():
    # the read's first base sits before the target start: the global left
    # extension must pay the bad-area 'X' (-1000) for that cell
    q = b"A" + TGT[0:149]
    # cheaper than an 'X' cell: one leading query base as a 1-gap (149 - 7)
    assert align(P0, TGT, q, 50, 51) == (0, 0, 149, 150, 149, 142)
    # a 10-base overhang (> band 8) cannot be one gap: below the 10-gap score
    got = align(P0, TGT, b"ACGTACGTAC" + TGT[0:140], 60, 70)
    assert got[3] == 150 and got[5] < 140 - 5 - 2 * 10
    # gap of 9 > band 8: cannot be one gap; the score is far below the 8-gap one
    q9 = TGT[1000:1100] + TGT[1109:1159]
    assert align(P0, TGT, q9, 1000, 0)[5] < 150 - 5 - 2 * 8


# ---------------------------------------------------------------- E-value and order

def test_evalue_closed_form_and_order():
    t, r = workload(n_t=4, t_len=8000, n_r=120)
    p = native.Params(total_db_size=32000.0)
    tl, ql = lists(t, r, p.lmer)
    o = oracle.search(p, tl, [(0, 2), (2, 4)], ql, [(0, 60), (60, 120)])["final"]
    assert len(o["tID"]) > 50
    qlen = ql.length[o["qID"]]
    e = np.array([p.K * p.total_db_size * int(n) * math.exp(-p.lam * int(s)) for n, s in zip(qlen, o["score"])])
    np.testing.assert_allclose(o["evalue"], e, rtol=1e-15)
    assert np.all(o["evalue"] <= p.cut_off)
    assert np.all(o["qHL"] >= p.lmer + p.stride)
    # fixResult order: query pair ascending, then score descending
    key = o["qID"] // 2
    assert np.all(np.diff(key) >= 0)
    same = np.diff(key) == 0
    assert np.all(np.diff(o["score"])[same] <= 0)


def test_oracle_tsv_row_format(tmp_path):
    q = TGT[500:650]
    (tmp_path / "t.fa").write_bytes(b">chrT first\n" + TGT + b"\n")
    (tmp_path / "q.fa").write_bytes(b">read1\n" + q[:70] + b"\n" + q[70:] + b"\n")
    out = tmp_path / "o.tsv"
    subprocess.run([oracle.EXE, "-t", str(tmp_path / "t.fa"), "-q", str(tmp_path / "q.fa"), "-o", str(out)],
                   check=True, stdout=subprocess.DEVNULL)
    rows = out.read_text().splitlines()
    e = 0.71 * 3000 * 150 * math.exp(-1.37 * 150)
    assert rows[0] == f"read1\t0\t150\t+\tchrT first\t500\t150\t150(100%)\t150\t{e:g}"


def test_total_db_size(tmp_path):
    f = tmp_path / "t.fa"
    f.write_bytes(b">a\nACGT\r\nAC\n@12\n>b>c\nGG\n")
    # every byte of every line without '>', '\r' and '@' lines included
    assert oracle.total_db_size([str(f)]) == 5 + 2 + 3 + 2


# ---------------------------------------------------------------- local-mode DP
#
# localAlign{Backward,Forward} (krnlAlignment.cu:312-533): the end node is the
# best cell over every row and band, scanned row by row and band by band with
# a strict '<' against a running best that starts at score 0
# (calcIdx = -1, endPoint = gap); the extension is calcIdx + 1 rows and
# endPoint - gap target columns.  Each expected tuple below is worked by hand
# from that rule and the scores M = +1, X = -3, gap = -5 - 2k (common.hpp).

PL = native.Params(local=1)


def _other(c: int) -> bytes:
    return bytes([next(b for b in b"ACGT" if b != c)])


def test_local_dp_stops_at_last_match():
    # 85 exact rows after the seed, then 'N' rows (each -3 on the diagonal,
    # gaps cost at least 7): the best cell is row 84 on the centre band
    # (score 85), so the extension is 84 + 1 rows and no target shift.
    # A lost '+1' gives qHL 99; a best cell taken at the last row gives 150.
    q = TGT[1000:1100] + b"N" * 50
    assert align(PL, TGT, q, 1000, 0) == (1000, 0, 100, 100, 100, 100)


def test_local_dp_all_negative_rows_extend_nothing():
    # every cell after the seed is below zero: no cell beats the initial
    # best of 0, calcIdx stays -1, so the extension is 0 rows and
    # endPoint - gap = 0 columns.  Starting the best from -infinity instead
    # would extend by one row (qHL 16, score 12).
    q = TGT[1000:1015] + b"N" * 30
    assert align(PL, TGT, q, 1000, 0) == (1000, 0, 15, 15, 15, 15)


def test_local_dp_equal_score_tie_keeps_first_cell():
    # rows 0..34 exact (centre score 35 at row 34), row 35 a mismatch (32),
    # rows 36..38 exact (35 again at row 38), then 'N'.  Strict '<' keeps the
    # first 35 (row 34): 35 rows, score 15 + 35.  With '<=' the later tie
    # wins: qHL 54, match 53.
    q = TGT[1000:1050] + _other(TGT[1050]) + TGT[1051:1054] + b"N" * 20
    assert align(PL, TGT, q, 1000, 0) == (1000, 0, 50, 50, 50, 50)


def test_local_dp_forward_across_deletion():
    # query = 20 N | target 1000..1017 | target 1020..1099 | 20 N (the query
    # lacks target bases 1018-1019), seed at query 60 <-> target 1042.
    # Backward: query 75..117 exact (43 rows) then N: +43, no shift.
    # Forward: 22 exact rows (query 59..38), one 2-base gap (-5 - 2*2 = -9),
    # 18 exact rows (query 37..20) -> best 22 - 9 + 18 = 31 at row 39 on band
    # offset +2: 40 rows, 42 target columns.  Totals: tIdx 1042 - 42 = 1000,
    # qIdx 60 - 40 = 20, tHL 15 + 43 + 42, qHL 15 + 43 + 40, match 15 + 43
    # + 40, score 15 + 43 + 31 (= 98 matches - 9 for the gap).
    q = b"N" * 20 + TGT[1000:1018] + TGT[1020:1100] + b"N" * 20
    assert align(PL, TGT, q, 1042, 60) == (1000, 20, 100, 98, 98, 89)


# ---------------------------------------------------------------- result dedup
#
# deleteHits_duplicateResult (CDeviceHitList.cu:32-128, is_duplicate
# CDeviceHitList.cuh:32-49): after resultSorting (stable, query pair
# ascending, score descending) a row is dropped when its (tID, tIdx, qID,
# qIdx) equals the row just before it in that order.

def _two_chain_read(rng):
    """A 3,000-base random target region R = 400 bases whose read carries
    exact ends (0..139, 260..399) and a middle with a substitution every 10
    bases (no exact 15-mer there): two seed chains on one diagonal more than
    `width` apart, each leaving one hit after the isolate and duplicate
    filters.  Global extension of either hit covers the whole read on the
    diagonal: tHL = qHL = 400, 388 matches, score 388 - 3 * 12 = 352."""
    region = bytes(rng.choice(list(b"ACGT"), 400).astype(np.uint8))
    read = bytearray(region)
    for i in range(145, 260, 10):
        read[i] = _other(read[i])[0]
    return region, bytes(read)


def _search_final(p, seqs, read):
    t = gen.Seqs(np.frombuffer(b"".join(seqs), np.uint8).copy(),
                 np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.int64))
    r = gen.Seqs(np.frombuffer(read, np.uint8).copy(), np.array([0, len(read)], np.int64))
    tl, ql = lists(t, r, p.lmer)
    return oracle.search(p, tl, [(0, len(seqs))], ql, [(0, 1)])["final"]


def test_result_dedup_adjacent_duplicate_collapses():
    rng = np.random.default_rng(21)
    region, read = _two_chain_read(rng)
    bg = bytes(rng.choice(list(b"ACGT"), 3000).astype(np.uint8))
    seq0 = bg[:1000] + region + bg[1400:]
    p = native.Params(total_db_size=3000.0)
    f = _search_final(p, [seq0], read)
    # the two chains' hits extend to one start with one score and sit next to
    # each other after the result sort: one row survives
    plus = f["qID"] == 0
    rows = list(zip(*[f[k][plus].tolist() for k in ("tID", "tIdx", "qIdx", "tHL", "qHL", "match", "score")]))
    assert rows == [(0, 1000, 0, 400, 400, 388, 352)]


def test_result_dedup_keeps_duplicates_that_are_not_adjacent():
    # The same region copied into a second target sequence at the same stride
    # phase (slot 3000 + 14 joint + 501 = 3515, a multiple of 5 like 1000), so
    # both copies give the same seeds.  Chain representatives after the
    # filters: in tID 0 chain A keeps its first seed (qIdx 0, the window's
    # element 0) and chain B its second (qIdx 265); in tID 1 both chains
    # keep their second seed (qIdx 5 and 265).  All four extend to one start
    # per copy with score 352, so the result order is the DP order: seed qIdx
    # descending, ties in seed order (tID): B0 B1 A1 A0.  B1 and A1 are
    # neighbours with equal (tID, tIdx, qID, qIdx): A1 goes.  B0 and A0 are
    # equal but not neighbours: both stay.  A dedup over all rows would leave
    # 2 rows, a missed neighbour 4, an order by tID first (B0 A0 B1 A1) 2.
    rng = np.random.default_rng(21)
    region, read = _two_chain_read(rng)
    bg = bytes(rng.choice(list(b"ACGT"), 3000).astype(np.uint8))
    bg1 = bytes(rng.choice(list(b"ACGT"), 3000).astype(np.uint8))
    seq0 = bg[:1000] + region + bg[1400:]
    seq1 = bg1[:501] + region + bg1[901:]
    assert (3000 + 14 + 501) % 5 == 0 and 1000 % 5 == 0
    p = native.Params(total_db_size=6000.0)
    f = _search_final(p, [seq0, seq1], read)
    plus = f["qID"] == 0
    rows = list(zip(*[f[k][plus].tolist() for k in ("tID", "tIdx", "qIdx", "tHL", "qHL", "match", "score")]))
    a0, a1 = (0, 1000, 0, 400, 400, 388, 352), (1, 501, 0, 400, 400, 388, 352)
    assert rows == [a0, a1, a0]


# ---------------------------------------------------------------- printResult
#
# CHostResultHolder::printResult (CHostResultHolder.cu:131-243): rows in
# fixResult order; a row equal in all ten printed fields to the last printed
# row is skipped and not counted; per query label the first row always
# prints and at most numOut rows print (numOut -1: all).

def _run_oracle_cli(tmp_path, targets: bytes, queries: bytes, *flags):
    (tmp_path / "t.fa").write_bytes(targets)
    (tmp_path / "q.fa").write_bytes(queries)
    out = tmp_path / "o.tsv"
    subprocess.run([oracle.EXE, "-t", str(tmp_path / "t.fa"), "-q", str(tmp_path / "q.fa"), "-o", str(out),
                    *[str(f) for f in flags]], check=True, stdout=subprocess.DEVNULL)
    return out.read_text().splitlines()


def _tsv_case():
    """Targets: chrA split into two overlapping records (@0 and @600, the
    preprocessDB layout) so that a read inside the overlap maps to both with
    identical printed rows; chrB holds the read with one substitution, chrC
    with two.  The read: 150 bases of chrA 700..849."""
    rng = np.random.default_rng(33)
    a = bytes(rng.choice(list(b"ACGT"), 1500).astype(np.uint8))
    read = a[700:850]
    b = bytes(rng.choice(list(b"ACGT"), 1000).astype(np.uint8))
    c = bytes(rng.choice(list(b"ACGT"), 1000).astype(np.uint8))
    rb = bytearray(read)
    rb[75] = _other(rb[75])[0]
    rc = bytearray(read)
    rc[50] = _other(rc[50])[0]
    rc[100] = _other(rc[100])[0]
    b = b[:300] + bytes(rb) + b[450:]
    c = c[:200] + bytes(rc) + c[350:]
    fa = (b">chrA\n@0\n" + a[:1000] + b"\n>chrA\n@600\n" + a[600:] + b"\n>chrB\n" + b + b"\n>chrC\n" + c + b"\n")
    return fa, b">r1\n" + read + b"\n", read


def _expected_rows(db):
    def row(tl, ti, m, s):
        e = 0.71 * db * 150 * math.exp(-1.37 * s)
        return f"r1\t0\t150\t+\t{tl}\t{ti}\t150\t{m}({m * 100 / 150:g}%)\t{s}\t{e:g}"
    # chrA at 700 (printed once: the overlap's second record gives the same
    # ten fields), chrB at 300 (149 matches, 149 - 3), chrC at 200 (148, 148 - 6)
    return [row("chrA", 700, 150, 150), row("chrB", 300, 149, 146), row("chrC", 200, 148, 142)]


@pytest.mark.parametrize("num_out,n", [(-1, 3), (1, 1), (2, 2), (0, 1), (5, 3)])
def test_print_num_out_and_equal_row_skip(tmp_path, num_out, n):
    fa, qa, _ = _tsv_case()
    (tmp_path / "db.fa").write_bytes(fa)
    db = oracle.total_db_size([str(tmp_path / "db.fa")])
    # one target window per record (tVRAM 0.001 MBase), as split targets run
    rows = _run_oracle_cli(tmp_path, fa, qa, "-numOut", num_out, "-tVRAM", 0.001)
    plus = [r for r in rows if r.split("\t")[3] == "+"]
    # the skipped duplicate does not count toward numOut: numOut 2 prints
    # chrA and chrB, not chrA alone
    assert plus == _expected_rows(db)[:n]
