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
