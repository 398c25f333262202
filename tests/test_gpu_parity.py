"""GPU parity: every stage of the HIP pipeline against the CPU oracle on the
same seeded inputs (integers bit-exact, E-values within 1e-9 relative)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from clast_amd import gen, native
from common import FIELDS, assert_rows_equal, db_size, lists, sorted_rows, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    print("init device 0", flush=True)
    native.init(0)


def expected_entries(p, tl, twins):
    """Index entries implied by the oracle's CDeviceHashTable of each window:
    cells with a non-zero (post-cut) size, positions whose k-mer does not run
    past the sequence end, in (key, window, position) order."""
    rows = []
    for w, (a, b) in enumerate(twins):
        ix = oracle.index(p, tl, a, b)
        lo = int(tl.gateway[a])
        gw = tl.gateway[a:b + 1] - lo
        for k in range(len(ix["key"])):
            c = int(ix["cell_size"][k])
            if c <= 0 or ix["key"][k] < 0:
                continue
            o = int(ix["gateway_index"][k])
            for pos in ix["positions"][o:o + c]:
                s = int(np.searchsorted(gw, pos, side="right")) - 1
                tidx = int(pos - gw[s])
                if tidx + p.lmer <= int(tl.length[a + s]):
                    rows.append((int(ix["key"][k]), w, a + s, tidx))
    rows.sort(key=lambda r: (r[0], r[1]))  # stable: windows in order, positions ascending
    return np.array(rows, dtype=np.int64).reshape(-1, 4)


@pytest.mark.parametrize("lmer,stride,cut", [(15, 5, 20), (11, 3, 4), (20, 4, -1), (31, 7, 20), (13, 2, 6)])
def test_index_entries(lmer, stride, cut):
    t, r = workload(n_t=5, t_len=12_000, n_r=10, repeat_frac=0.2, lowc=0.02)
    p = native.Params(lmer=lmer, stride=stride, cut_repeat=cut, total_db_size=db_size(t))
    tl, _ = lists(t, r, lmer)
    twins = [(0, 2), (2, 5)]
    print(f"index lmer={lmer} stride={stride} cut={cut}", flush=True)
    ix = native.Index(p, tl, twins)
    got = ix.entries()
    exp = expected_entries(p, tl, twins)
    assert len(got["key"]) == len(exp)
    np.testing.assert_array_equal(got["key"], exp[:, 0])
    np.testing.assert_array_equal(got["window"], exp[:, 1])
    np.testing.assert_array_equal(got["tID"], exp[:, 2])
    np.testing.assert_array_equal(got["tIdx"], exp[:, 3])


@pytest.mark.parametrize("lmer,form", [(15, "auto"), (12, "bitmap"), (17, "auto")])
def test_index_save_load(tmp_path, lmer, form, monkeypatch):
    # an index written to a file and read back (its cell tables rebuilt, in
    # the form the environment asks for) holds the same entries and gives
    # the same rows as the index it was saved from, and the oracle's
    t, r = workload(n_t=5, t_len=12_000, n_r=200, repeat_frac=0.2, seed=19)
    p = native.Params(lmer=lmer, total_db_size=db_size(t))
    tl, ql = lists(t, r, lmer)
    twins = [(0, 2), (2, 5)]
    qwins = native.windows(ql.gateway, 30_000, query=True)
    print(f"index save/load lmer={lmer} form={form}", flush=True)
    ix = native.Index(p, tl, twins)
    path = str(tmp_path / "index.bin")
    ix.save(path)
    monkeypatch.setenv("CLAST_INDEX_FORM", form)
    ix2 = native.Index.load(p, path)
    e1, e2 = ix.entries(), ix2.entries()
    for k in e1:
        np.testing.assert_array_equal(e1[k], e2[k])
    q = native.Query(p, ql, qwins)
    got1 = native.search_batch(p, ix, q).arrays()
    got2 = native.search_batch(p, ix2, q).arrays()
    for h in (q, ix, ix2):
        h.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    assert_rows_equal(got2, got1, what="loaded vs built")
    assert_rows_equal(got2, o["final"], what="loaded vs oracle")


def run_unit(p, t, r, tw=None, qw=None, keep=True):
    """keep=True retains the stages (every strand on the global path);
    keep=False runs the per-strand LDS path."""
    tl, ql = lists(t, r, p.lmer)
    tw = tw or (0, tl.n_seq)
    qw = qw or (0, ql.n_seq // 2)
    ix = native.Index(p, tl, [tw])
    q = native.Query(p, ql, [qw])
    hits = native.search_batch(p, ix, q, keep_stages=keep)
    o = oracle.search(p, tl, [tw], ql, [qw], dump=True)
    return hits, o


SETTINGS = [
    dict(),
    dict(local=1),
    dict(gap=2),
    dict(gap=16, lmer=12),
    dict(lmer=21, stride=3),
    dict(lmer=31, stride=7),
    dict(lmer=13, stride=2, cut_repeat=6),
    dict(cut_repeat=-1, width=30),
    dict(cut_off=-1),
]


@pytest.mark.parametrize("kw", SETTINGS, ids=[str(k) or "default" for k in SETTINGS])
def test_unit_stages(kw):
    t, r = workload(n_t=6, t_len=25_000, n_r=250, repeat_frac=0.1, indel=0.02)
    p = native.Params(total_db_size=db_size(t), **kw)
    print(f"unit stages {kw}", flush=True)
    hits, o = run_unit(p, t, r)
    # seeds after the three boundary filters, in the reference's order
    assert_rows_equal(hits.stage(native.STAGE_SEEDS), o["boundary"], FIELDS[:4], what="seeds")
    assert_rows_equal(hits.stage(native.STAGE_SORTED), o["sorted"], FIELDS[:4], what="sorted seeds")
    assert_rows_equal(hits.stage(native.STAGE_HITS), o["duplicate"], FIELDS[:4], what="hits")
    # DP outputs: same multiset of extended hits
    np.testing.assert_array_equal(sorted_rows(hits.stage(native.STAGE_ALIGNED)), sorted_rows(o["aligned"]))
    got = hits.arrays()
    assert_rows_equal(got, o["final"], what="final")
    assert len(o["final"]["tID"]) > 0
    fast, _ = run_unit(p, t, r, keep=False)
    assert_rows_equal(fast.arrays(), o["final"], what="final, per-strand path")


def test_plain_search_matches_batch_unit():
    t, r = workload(n_t=4, t_len=20_000, n_r=200)
    p = native.Params(total_db_size=db_size(t))
    tl, ql = lists(t, r, p.lmer)
    ix = native.Index(p, tl, [(0, 4)])
    q = native.Query(p, ql, [(50, 200)])
    got = native.search(p, ix, q, 0, 100).arrays()
    o = oracle.search(p, tl, [(0, 4)], ql, [(50, 200)])
    assert_rows_equal(got, o["final"], what="clast_search")


def test_batch_many_windows():
    t, r = workload(n_t=9, t_len=15_000, n_r=400, repeat_frac=0.15, random_frac=0.1)
    p = native.Params(total_db_size=db_size(t))
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 40_000, query=False)
    qwins = native.windows(ql.gateway, 30_000, query=True)
    assert len(twins) > 2 and len(qwins) > 2
    print(f"batch {len(twins)}x{len(qwins)} windows", flush=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    got = native.search_batch(p, ix, q).arrays()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert_rows_equal(got, o["final"], what="batch")
    np.testing.assert_array_equal(got["twin"], o["final"]["twin"])
    # the same rows, unit by unit, through the plain entry point
    for w, (a, b) in enumerate(twins[:2]):
        ixw = native.Index(p, tl, [(a, b)])
        for qa, qb in qwins[:2]:
            qq = native.Query(p, ql, [(qa, qb)])
            one = native.search(p, ixw, qq, a, 2 * qa).arrays()
            ou = oracle.search(p, tl, [(a, b)], ql, [(qa, qb)])
            assert_rows_equal(one, ou["final"], what=f"unit {w},{qa}")


@pytest.mark.parametrize("n_t", [40, 70])
def test_element0_many_units(n_t):
    # Element 0 of every (target window, query window) unit: target windows
    # of one sequence each (40: unit candidates as per-strand window masks;
    # 70: over 64 windows, candidates through unit_first directly) and query
    # windows of a few reads, so one wavefront of strands spans several query
    # windows.
    t, r = workload(n_t=n_t, t_len=3_000, n_r=300, repeat_frac=0.1, random_frac=0.1, seed=31)
    p = native.Params(total_db_size=db_size(t))
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 3_100, query=False)
    qwins = native.windows(ql.gateway, 3_000, query=True)
    assert len(twins) == n_t and len(qwins) > 20
    print(f"element 0: {len(twins)}x{len(qwins)} units", flush=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    got = native.search_batch(p, ix, q).arrays()
    q.close()
    ix.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    assert_rows_equal(got, o["final"], what=f"{n_t} target windows")
    np.testing.assert_array_equal(got["twin"], o["final"]["twin"])


def test_uneven_lengths_and_local():
    t, r = workload(n_t=4, t_len=40_000, n_r=60, len_mode=gen.LEN_LOGUNIFORM, read_len=100, len_b=3000,
                    random_frac=0.2, indel=0.05)
    for local in (0, 1):
        p = native.Params(total_db_size=db_size(t), local=local)
        print(f"uneven local={local}", flush=True)
        hits, o = run_unit(p, t, r, keep=False)
        assert_rows_equal(hits.arrays(), o["final"], what=f"uneven local={local}")


def test_edge_inputs():
    # reads shorter than lMer, reads of N, reads with X and other letters
    t, _ = workload(n_t=3, t_len=10_000, n_r=1)
    seqs = [b"ACGT", b"N" * 40, t.seq(0)[100:260], t.seq(1)[500:650].replace(b"A", b"X", 3),
            t.seq(2)[1000:1150].replace(b"C", b"R", 5), b"", t.seq(0)[9900:10000]]
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.int64)
    r = gen.Reads(bases, off)
    p = native.Params(total_db_size=db_size(t))
    for keep in (True, False):
        hits, o = run_unit(p, t, r, keep=keep)
        assert_rows_equal(hits.arrays(), o["final"], what=f"edge keep={keep}")
    # a query set with no k-mer position at all
    short = gen.Reads(np.frombuffer(b"ACGTACGT", np.uint8).copy(), np.array([0, 8], np.int64))
    hits, o = run_unit(p, t, short)
    assert len(hits) == 0 and len(o["final"]["tID"]) == 0


def test_result_sort():
    rng = np.random.default_rng(3)
    n = 5000
    rows = {k: rng.integers(0, 1000, n).astype(np.int32) for k in FIELDS}
    rows["qID"] = rng.integers(0, 200, n).astype(np.int32)
    rows["score"] = rng.integers(-50, 50, n).astype(np.int32)
    rows["evalue"] = rng.random(n)
    order = np.lexsort((np.arange(n), -rows["score"].astype(np.int64), rows["qID"] // 2))
    exp = {k: v[order] for k, v in rows.items()}
    native.result_sort(rows)
    for k in exp:
        np.testing.assert_array_equal(rows[k], exp[k])


def test_long_reads_mix_fast_and_global_paths():
    # at stride 2 (about one seed per two bases on the true strand): 6-10 kbp
    # reads carry > 2048 seeds and take the global path, 2.5-3.8 kbp reads
    # (1024-2048 seeds) and 0.8-1.5 kbp reads (256-1024) the big LDS pass,
    # 150 bp reads the first LDS pass, all in one query set
    t, _ = workload(n_t=3, t_len=60_000, n_r=1)
    long_r = gen.make_reads(gen.ReadSpec(21, 12, len_mode=gen.LEN_LOGUNIFORM, len_a=6000, len_b=10_000,
                                         sub_rate=0.01, indel_rate=0.002), t)
    big_r = gen.make_reads(gen.ReadSpec(24, 10, len_mode=gen.LEN_LOGUNIFORM, len_a=2500, len_b=3800,
                                        sub_rate=0.01, indel_rate=0.002), t)
    mid_r = gen.make_reads(gen.ReadSpec(23, 20, len_mode=gen.LEN_LOGUNIFORM, len_a=800, len_b=1500,
                                        sub_rate=0.02, indel_rate=0.005), t)
    short_r = gen.make_reads(gen.ReadSpec(22, 80, sub_rate=0.03, indel_rate=0.01), t)
    parts = [long_r, big_r, mid_r, short_r]
    bases = np.concatenate([x.bases for x in parts])
    off = [np.zeros(1, np.int64)]
    for x in parts:
        off.append(x.offsets[1:] + off[-1][-1])
    r = gen.Reads(bases, np.concatenate(off).astype(np.int64))
    for local in (0, 1):
        p = native.Params(total_db_size=db_size(t), local=local, stride=2)
        tl, ql = lists(t, r, p.lmer)
        ix = native.Index(p, tl, [(0, 3)])
        q = native.Query(p, ql, [(0, 61), (61, 122)])
        print(f"long reads local={local}", flush=True)
        got = native.search_batch(p, ix, q)
        assert got.times()[7] > 3000  # seeds
        o = oracle.search(p, tl, [(0, 3)], ql, [(0, 61), (61, 122)])
        assert_rows_equal(got.arrays(), o["final"], what=f"long local={local}")


@pytest.mark.parametrize("n_t", [40, 70])
def test_global_path_many_windows(n_t):
    # Strands over 2048 seeds take the global path, whose seed writer spreads
    # each round's seeds over its workgroup and records unit candidates as
    # one window mask per strand (40 target windows) or one note per seed
    # (70: over 64 windows).  Target windows of one sequence each with 30%
    # repeats (cells of many entries), long reads at stride 2 next to short
    # ones, and query windows of a few reads, so that global-path strands are
    # the first strands of some units.
    t = gen.make_target(gen.TargetSpec(41, n_t, 12_000, n_rate=0.001, repeat_frac=0.3, n_families=6,
                                       fam_min=300, fam_max=1500, div_min=0.0, div_max=0.05))
    long_r = gen.make_reads(gen.ReadSpec(42, 14, len_mode=gen.LEN_LOGUNIFORM, len_a=5000, len_b=9000,
                                         sub_rate=0.01, indel_rate=0.002, repeat_read_frac=0.3), t)
    short_r = gen.make_reads(gen.ReadSpec(43, 60, sub_rate=0.03, indel_rate=0.01), t)
    bases = np.concatenate([long_r.bases, short_r.bases])
    off = np.concatenate([long_r.offsets, long_r.offsets[-1] + short_r.offsets[1:]]).astype(np.int64)
    r = gen.Reads(bases, off)
    p = native.Params(total_db_size=db_size(t), stride=2)
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 12_100, query=False)
    qwins = native.windows(ql.gateway, 20_000, query=True)
    assert len(twins) == n_t and len(qwins) > 4
    print(f"global path: {len(twins)}x{len(qwins)} units", flush=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    got = native.search_batch(p, ix, q)
    assert got.times()[2] > 0  # the global path's seed sort ran
    rows = got.arrays()
    q.close()
    ix.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    assert_rows_equal(rows, o["final"], what=f"global path, {n_t} target windows")
    np.testing.assert_array_equal(rows["twin"], o["final"]["twin"])


@pytest.mark.parametrize("gap", [2, 8, 16])
def test_dp_16bit_lanes(gap, monkeypatch):
    # The DP runs a task in packed 16-bit lanes when its query has at most
    # 30000 bases, neither its query strand nor its target sequence holds a
    # star (E F I J L O P Q Z) or X letter, and its band stays inside the
    # target sequence and window; every other task takes the int32 kernel.
    # Both kinds in one set: a target sequence and some reads with such
    # letters, reads from sequence ends, read lengths 60-1500 (two tasks of
    # unequal rows share a lane).  Rows equal the oracle's with the 16-bit
    # path on and off.
    monkeypatch.setenv("CLAST_DP_BAND", "off")  # no extension leaves the 16-bit kernels for k_dp_band
    t, r = workload(n_t=5, t_len=20_000, n_r=240, len_mode=gen.LEN_LOGUNIFORM, read_len=60, len_b=1500,
                    indel=0.02, seed=11)
    tb = t.bases.copy()
    o1 = int(t.offsets[1])
    tb[o1 + 3000:o1 + 20_000:997] = ord("X")
    tb[o1 + 3500:o1 + 20_000:1499] = ord("E")
    t = gen.Target(tb, t.offsets)
    seqs = [r.seq(i) for i in range(len(r))]
    for i in range(0, len(seqs), 7):  # star letters in every 7th read
        b = bytearray(seqs[i])
        b[len(b) // 2] = ord("Q")
        seqs[i] = bytes(b)
    for k in (0, 2, 3, 4):  # reads from both ends of clean sequences
        seqs += [t.seq(k)[:180], t.seq(k)[-180:], t.seq(k)[5:90], t.seq(k)[-95:-4]]
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    r = gen.Reads(bases, off)
    p = native.Params(total_db_size=db_size(t), gap=gap)
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 45_000, query=False)
    qwins = native.windows(ql.gateway, 40_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    print(f"dp16 gap={gap}: {len(twins)}x{len(qwins)} windows", flush=True)
    h = native.search_batch(p, ix, q)
    got, cnt = h.arrays(), h.counts()
    monkeypatch.setenv("CLAST_DP16", "0")
    h0 = native.search_batch(p, ix, q)
    got0, cnt0 = h0.arrays(), h0.counts()
    ix.close()
    q.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    ntask = 2 * int(cnt[1])
    assert 0 < cnt[4] < ntask and 0 < cnt[5] < cnt[2], f"16-bit tasks {cnt[4]} of {ntask}"
    assert cnt0[4] == 0 and cnt0[2] == cnt[2]
    assert_rows_equal(got, o["final"], what=f"dp16 gap {gap}")
    assert_rows_equal(got0, o["final"], what=f"int32 dp gap {gap}")


@pytest.mark.parametrize("with_n", [False, True], ids=["no_n", "n_in_two"])
@pytest.mark.parametrize("gap", [3, 4, 5, 6, 7, 8, 9, 12, 16])
def test_dp16_query_profiles(gap, with_n, monkeypatch):
    # The 16-bit score pass with 9 or more bands takes its diagonal terms
    # from per-row query profiles when every task of a 64-pair chunk lies on
    # a target sequence without N; other chunks use the class-code xor.
    # Gaps 4-8 (four rows per loop step) look the profiles up in a lane
    # table, wider ones build them per row, gap 3 (7 bands) has none; every
    # gap reads its class streams from the mirrored 4-bit arrays.
    # Every 5th read carries an N (the query's other class, which matches no
    # target letter).  Targets: no N at all (every chunk on profiles), or N
    # in two of six sequences (both kinds of chunk in one launch).  Rows equal
    # the oracle's and the int32 kernel's.
    monkeypatch.setenv("CLAST_DP_BAND", "off")
    t, r = workload(n_t=6, t_len=20_000, n_r=300, n_rate=0.0, len_mode=gen.LEN_LOGUNIFORM, read_len=60,
                    len_b=1200, indel=0.02, seed=21)
    if with_n:
        tb = t.bases.copy()
        for k in (1, 4):
            tb[int(t.offsets[k]) + 500:int(t.offsets[k + 1]) - 500:331] = ord("N")
        t = gen.Target(tb, t.offsets)
    seqs = [r.seq(i) for i in range(len(r))]
    for i in range(0, len(seqs), 5):
        b = bytearray(seqs[i])
        b[len(b) // 3] = ord("N")
        seqs[i] = bytes(b)
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    r = gen.Reads(bases, off)
    p = native.Params(total_db_size=db_size(t), gap=gap)
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 45_000, query=False)
    qwins = native.windows(ql.gateway, 40_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    print(f"profiles gap={gap} with_n={with_n}: {len(twins)}x{len(qwins)} windows", flush=True)
    h = native.search_batch(p, ix, q)
    got, cnt = h.arrays(), h.counts()
    monkeypatch.setenv("CLAST_DP16", "0")
    h0 = native.search_batch(p, ix, q)
    got0 = h0.arrays()
    ix.close()
    q.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    assert cnt[4] > 0, "no task in 16-bit lanes"
    assert_rows_equal(got, o["final"], what=f"profiles gap {gap}")
    assert_rows_equal(got0, o["final"], what=f"int32 dp gap {gap}")


@pytest.mark.parametrize("local,gap", [(0, 2), (0, 8), (0, 16), (1, 8), (1, 16)])
def test_dp_lane_per_band(local, gap, monkeypatch):
    # Extensions of at least 1024 rows run one lane per band (two tasks per
    # lane group, interleaved a step apart); shorter ones take the other DP
    # kernels.  Reads of 1-6 kbp, some with star or X letters, some from
    # sequence ends (bands that leave the target), in one set.  Rows equal
    # the oracle's with every long extension on the lane-per-band kernel
    # (CLAST_DP_BAND=all), with only those the 16-bit kernel cannot take
    # (rest: here the reads with X letters and those running off a sequence
    # end; in local mode all), with none (off), and with the default (the autotuner's pick).
    t, _ = workload(n_t=3, t_len=50_000, n_r=1, seed=13)
    long_r = gen.make_reads(gen.ReadSpec(31, 16, len_mode=gen.LEN_LOGUNIFORM, len_a=1100, len_b=6000,
                                         sub_rate=0.02, indel_rate=0.01), t)
    short_r = gen.make_reads(gen.ReadSpec(32, 40, sub_rate=0.03, indel_rate=0.01), t)
    seqs = [long_r.seq(i) for i in range(len(long_r))] + [short_r.seq(i) for i in range(len(short_r))]
    for i in range(0, len(long_r), 5):
        b = bytearray(seqs[i])
        b[len(b) // 3] = ord("X")
        b[len(b) // 2] = ord("J")
        seqs[i] = bytes(b)
    seqs += [t.seq(0)[:1500], t.seq(2)[-1400:], t.seq(1)[:1200]]
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    r = gen.Reads(bases, off)
    p = native.Params(total_db_size=db_size(t), gap=gap, local=local)
    tl, ql = lists(t, r, p.lmer)
    twins = [(0, 3)]
    qwins = native.windows(ql.gateway, 60_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    print(f"band local={local} gap={gap}: {len(qwins)} query windows", flush=True)
    runs = {}
    for mode in ("all", "rest", "off", None):
        if mode is None:
            monkeypatch.delenv("CLAST_DP_BAND", raising=False)
        else:
            monkeypatch.setenv("CLAST_DP_BAND", mode)
        h = native.search_batch(p, ix, q)
        runs[mode] = (h.arrays(), h.counts())
    monkeypatch.setenv("CLAST_DP_BAND", "1")
    with pytest.raises(native.ClastError, match="CLAST_DP_BAND"):
        native.search_batch(p, ix, q)
    monkeypatch.delenv("CLAST_DP_BAND")
    ix.close()
    q.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    c_all, c_rest, c_off, c_def = (runs[m][1] for m in ("all", "rest", "off", None))
    # (the default is whatever the DP autotuner picked for this kind)
    assert c_all[6] >= c_rest[6] > 0 and c_off[6] == 0, (c_all[6], c_rest[6], c_off[6])
    assert c_all[2] == c_rest[2] == c_off[2] == c_def[2]
    if local:
        assert c_rest[6] == c_all[6]  # no 16-bit kernel in local mode
    for mode, (got, _) in runs.items():
        assert_rows_equal(got, o["final"], what=f"band mode {mode} local={local} gap={gap}")


@pytest.mark.parametrize("lmer", [12, 15])
def test_query_from_reads(lmer):
    # Query strands and joints built on the device from the reads as given
    # equal the host-built list (reverse complement: A<->T, C<->G, any other
    # byte N; joints: the strand's first bases, N past a short strand): the
    # same rows as the host path and the oracle, over several windows, with
    # lowercase, X, star and N letters, reads shorter than lMer and an empty
    # read.
    t, r = workload(n_t=4, t_len=20_000, n_r=220, len_mode=gen.LEN_LOGUNIFORM, read_len=60, len_b=900, seed=17)
    seqs = [r.seq(i) for i in range(len(r))]
    seqs[3] = seqs[3].lower()
    seqs[5] = seqs[5][:40] + b"X" + seqs[5][41:]
    seqs[7] = seqs[7][:30] + b"NNNNNNNN" + seqs[7][38:]
    seqs[9] = seqs[9][:50] + b"EJ" + seqs[9][52:]
    seqs += [b"ACG", b"", t.seq(1)[200:210], t.seq(2)[300:600].lower().upper()]
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    p = native.Params(total_db_size=db_size(t), lmer=lmer)
    tl = native.build_list(t.bases, t.offsets, lmer, False)
    ql = native.build_list(bases, off, lmer, True)
    twins = native.windows(tl.gateway, 40_000, query=False)
    qwins = native.windows(ql.gateway, 30_000, query=True)
    g = native.read_gateway(np.diff(off), lmer)
    np.testing.assert_array_equal(g, ql.gateway)
    assert native.windows(g, 30_000, query=True) == qwins
    assert len(qwins) > 1
    ix = native.Index(p, tl, twins)
    q_host = native.Query(p, ql, qwins)
    q_dev = native.Query.from_reads(p, bases, off, qwins)
    print(f"query from reads lmer={lmer}: {len(qwins)} windows", flush=True)
    got_host = native.search_batch(p, ix, q_host).arrays()
    got_dev = native.search_batch(p, ix, q_dev).arrays()
    for q in (q_host, q_dev):
        q.close()
    ix.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    assert_rows_equal(got_dev, got_host, what=f"device vs host encoding lmer {lmer}")
    assert_rows_equal(got_dev, o["final"], what=f"device encoding vs oracle lmer {lmer}")


@pytest.mark.parametrize("lmer", [12, 15])
def test_query_strands_as_given(lmer):
    # Query strands taken as given rather than a read and its reverse
    # complement: pairs of unequal lengths (the lookup walks them in order)
    # and equal-length pairs of unrelated reads (walked interleaved, but
    # mirrored positions are not each other's reverse complement), against
    # the oracle over the same list.
    t, r = workload(n_t=4, t_len=20_000, n_r=200, len_mode=gen.LEN_LOGUNIFORM, read_len=60, len_b=600, seed=29)
    seqs = [r.seq(i) for i in range(len(r))]
    for i in range(0, 80, 2):
        n = min(len(seqs[i]), len(seqs[i + 1]))
        seqs[i], seqs[i + 1] = seqs[i][:n], seqs[i + 1][:n]
    lens = np.array([len(x) for x in seqs])
    assert np.sum(lens[0::2] == lens[1::2]) >= 40 and np.sum(lens[0::2] != lens[1::2]) >= 40
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    p = native.Params(total_db_size=db_size(t), lmer=lmer)
    tl = native.build_list(t.bases, t.offsets, lmer, False)
    ql = native.build_list(bases, off, lmer, False)
    twins = [(0, tl.n_seq)]
    qwins = native.windows(ql.gateway, 20_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    print(f"strands as given lmer={lmer}: {len(qwins)} windows", flush=True)
    got = native.search_batch(p, ix, q).arrays()
    q.close()
    ix.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    assert_rows_equal(got, o["final"], what=f"strands as given lmer {lmer}")


@pytest.mark.parametrize("form", ["direct", "bitmap"])
@pytest.mark.parametrize("lmer", [12, 15, 17])
def test_index_forms(form, lmer, monkeypatch):
    # both cell-lookup forms (direct prefix table / bitmap + rank), with and
    # without prefix narrowing (lMer 17 keeps 30 of 34 key bits)
    monkeypatch.setenv("CLAST_INDEX_FORM", form)
    t, r = workload(n_t=6, t_len=20_000, n_r=300, repeat_frac=0.15, indel=0.02)
    p = native.Params(total_db_size=db_size(t), lmer=lmer)
    tl, ql = lists(t, r, lmer)
    twins = native.windows(tl.gateway, 50_000, query=False)
    qwins = native.windows(ql.gateway, 30_000, query=True)
    print(f"index form {form} lmer={lmer}: {len(twins)}x{len(qwins)} windows", flush=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    got = native.search_batch(p, ix, q).arrays()
    ix.close()
    q.close()
    o = oracle.search(p, tl, twins, ql, qwins)
    assert len(o["final"]["tID"]) > 0
    assert_rows_equal(got, o["final"], what=f"form {form} lmer {lmer}")


@pytest.mark.parametrize("cut_off", [10.0, 1e-200])
def test_large_pair_all_rows_filtered(cut_off):
    # one read whose region sits in 300 target sequences (no repeat cut):
    # its '+' strand has 300 hits, a pair over 256 hits, which takes the
    # sort path of the result stage.  With a tiny cutOff every row fails the
    # E-value cut and the pair leaves no row (the sort path has nothing to
    # order); with the default one every copy gives a row.  Both = oracle.
    rng = np.random.default_rng(44)
    region = rng.choice(list(b"ACGT"), 150).astype(np.uint8)
    seqs = [np.concatenate([rng.choice(list(b"ACGT"), 60).astype(np.uint8), region,
                            rng.choice(list(b"ACGT"), 90).astype(np.uint8)]) for _ in range(300)]
    t = gen.Seqs(np.concatenate(seqs), np.arange(301, dtype=np.int64) * 300)
    r = gen.Seqs(region.copy(), np.array([0, 150], np.int64))
    p = native.Params(cut_repeat=-1, cut_off=cut_off, total_db_size=float(len(t.bases)))
    tl, ql = lists(t, r, p.lmer)
    print(f"large pair, cutOff {cut_off}", flush=True)
    ix = native.Index(p, tl, [(0, 300)])
    q = native.Query(p, ql, [(0, 1)])
    got = native.search_batch(p, ix, q).arrays()
    exp = oracle.search(p, tl, [(0, 300)], ql, [(0, 1)])["final"]
    for h in (q, ix):
        h.close()
    assert len(exp["tID"]) == (300 if cut_off == 10.0 else 0)
    assert_rows_equal(got, exp, what=f"large pair cutOff {cut_off}")


@pytest.mark.parametrize("gap", [2, 8, 16])
def test_dp_16bit_past_sequence_end(gap, monkeypatch):
    # Reads that run off a target sequence: the last part of a sequence
    # followed by 1 to `gap` random bases (the backward extension's band
    # reaches past the end, whose positions k_dp scores as 'X'), and 1 to
    # `gap` random bases followed by the first part of a sequence (the
    # forward band reaches before 0), 150 bp and 1.5-4 kbp long (a longer
    # overhang fails the corner filter: no hit).  Those tasks run in the
    # 16-bit kernel with 'X' marked past the end: every task with rows runs
    # in 16-bit lanes, and the rows equal the oracle's and the int32 kernel's.
    monkeypatch.setenv("CLAST_DP_BAND", "off")  # no extension leaves the 16-bit kernels for k_dp_band
    rng = np.random.default_rng(100 + gap)
    t, _ = workload(n_t=4, t_len=30_000, n_r=1, n_rate=0.0, seed=17)
    seqs = []
    for k in range(4):
        s = t.seq(k)
        for length in (150, 1500, 4000):
            for over in sorted({1, 2, max(1, gap // 2), gap}):
                junk = bytes(rng.choice(list(b"ACGT"), over).astype(np.uint8))
                seqs.append(s[-(length - over):] + junk)
                seqs.append(junk + s[:length - over])
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    r = gen.Reads(bases, off)
    p = native.Params(total_db_size=db_size(t), gap=gap)
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 70_000, query=False)
    qwins = native.windows(ql.gateway, 60_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    print(f"dp16 past the end gap={gap}: {len(twins)}x{len(qwins)} windows", flush=True)
    h = native.search_batch(p, ix, q)
    got, cnt = h.arrays(), h.counts()
    monkeypatch.setenv("CLAST_DP16", "0")
    got0 = native.search_batch(p, ix, q).arrays()
    ix.close()
    q.close()
    o = oracle.search(p, tl, twins, ql, qwins)["final"]
    assert len(o["tID"]) >= len(seqs)
    assert_rows_equal(got, o, what=f"dp16 past the end, gap {gap}")
    assert_rows_equal(got0, o, what=f"int32 dp past the end, gap {gap}")
    # every task with rows took the 16-bit kernels (each hit has one or
    # two), but for gap > lMer - 1: then a backward band from a seed at the
    # sequence start also reads before it (the side k_dp does not guard,
    # real bases of the window), and such a task stays on the int32 kernel
    if gap < p.lmer:
        assert cnt[4] >= cnt[1], (cnt[4], cnt[1])
        assert cnt[5] == cnt[2]
    else:
        assert 0 < cnt[5] < cnt[2]


def _dp_mix_workload(seed):
    """Short and long reads (1-3 kbp: extensions over 1024 rows), some off a
    sequence end, some with X / star letters, over two target windows."""
    t, _ = workload(n_t=4, t_len=40_000, n_r=1, seed=seed)
    tb = t.bases.copy()
    o1 = int(t.offsets[1])
    tb[o1 + 2000:o1 + 40_000:1777] = ord("X")
    t = gen.Target(tb, t.offsets)
    long_r = gen.make_reads(gen.ReadSpec(seed + 1, 40, len_mode=gen.LEN_LOGUNIFORM, len_a=1100, len_b=3000,
                                         sub_rate=0.03, indel_rate=0.03), t)
    short_r = gen.make_reads(gen.ReadSpec(seed + 2, 120, sub_rate=0.04, indel_rate=0.01), t)
    seqs = [long_r.seq(i) for i in range(len(long_r))] + [short_r.seq(i) for i in range(len(short_r))]
    seqs += [t.seq(0)[-1400:] + b"ACG", b"TTGA" + t.seq(2)[:1300], t.seq(3)[-200:]]
    b = bytearray(seqs[3])
    b[len(b) // 3] = ord("Q")
    seqs[3] = bytes(b)
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    return t, gen.Reads(bases, off)


def test_dp_every_configuration_gives_the_oracle_rows(monkeypatch):
    # Each DP launch configuration the autotuner chooses between: the 16-bit
    # kernel's schedule (static snake / persistent work-stealing), which
    # long extensions run one lane per band (off / rest / all) and the radix
    # digit of the task-key sort (8 / 5 bits).  Rows equal the oracle's under
    # every one, and each search reports the configuration it ran with.
    t, r = _dp_mix_workload(61)
    p = native.Params(total_db_size=db_size(t))
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 90_000, query=False)
    qwins = native.windows(ql.gateway, 80_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    exp = oracle.search(p, tl, twins, ql, qwins)["final"]
    assert len(exp["tID"]) > 100
    monkeypatch.setenv("CLAST_TUNE", "0")
    for sched in ("static", "persist"):
        for band in ("off", "rest", "all"):
            for radix in (8, 5):
                monkeypatch.setenv("CLAST_DP_SCHED", sched)
                monkeypatch.setenv("CLAST_DP_BAND", band)
                monkeypatch.setenv("CLAST_RADIX_BITS", str(radix))
                print(f"dp config {sched} {band} {radix}", flush=True)
                h = native.search_batch(p, ix, q)
                assert h.dp_config() == (sched == "persist", ("off", "rest", "all").index(band), radix)
                assert_rows_equal(h.arrays(), exp, what=f"dp config {sched}/{band}/{radix}")
                h.close()
    ix.close()
    q.close()


def test_dp_tuner_keeps_the_rows(monkeypatch):
    # With trials on (the default), the first search of a kind times the
    # configurations and keeps the fastest; its rows and those of the next
    # search (which reuses the choice) equal the oracle's.
    for k in ("CLAST_TUNE", "CLAST_DP_SCHED", "CLAST_DP_BAND", "CLAST_RADIX_BITS"):
        monkeypatch.delenv(k, raising=False)
    t, r = _dp_mix_workload(71)
    p = native.Params(total_db_size=db_size(t), gap=7)  # a kind no other test tunes
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 90_000, query=False)
    qwins = native.windows(ql.gateway, 80_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    exp = oracle.search(p, tl, twins, ql, qwins)["final"]
    print("dp tuner", flush=True)
    h1 = native.search_batch(p, ix, q)
    rep = native.tuner_report()
    line = [x for x in rep.splitlines() if x.startswith("gap=7 local=0 bucket=2")]
    assert line, rep
    h2 = native.search_batch(p, ix, q)
    assert h1.dp_config() == h2.dp_config()
    want = "sched=%d band=%d radix=%d" % h2.dp_config()
    assert want in line[0], (want, line[0])
    assert_rows_equal(h1.arrays(), exp, what="tuned first search")
    assert_rows_equal(h2.arrays(), exp, what="tuned second search")
    for h in (h1, h2, ix, q):
        h.close()


def test_dp_order_graph_replays_equal_eager(monkeypatch):
    # The DP stage's task-ordering chain is captured into a hipGraph per
    # shape and replayed.  One query set searched four times (the first
    # search captures, the others replay), between them searches of a
    # second query set, with the graphs off (kernel by kernel) and on, with
    # synchronisations in between: every replay's rows equal the
    # kernel-by-kernel rows and the oracle's.
    for k, v in (("CLAST_TUNE", "0"), ("CLAST_DP_SCHED", "persist"), ("CLAST_DP_BAND", "rest")):
        monkeypatch.setenv(k, v)
    t, r = _dp_mix_workload(81)
    p = native.Params(total_db_size=db_size(t))
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 90_000, query=False)
    qwins = native.windows(ql.gateway, 80_000, query=True)
    ix = native.Index(p, tl, twins)
    qa = native.Query(p, ql, qwins)
    qb = native.Query(p, ql, qwins[:1])
    exp_a = oracle.search(p, tl, twins, ql, qwins)["final"]
    exp_b = oracle.search(p, tl, twins, ql, qwins[:1])["final"]
    monkeypatch.setenv("CLAST_GRAPH", "0")
    eager_a = native.search_batch(p, ix, qa).arrays()
    eager_b = native.search_batch(p, ix, qb).arrays()
    assert_rows_equal(eager_a, exp_a, what="eager a")
    assert_rows_equal(eager_b, exp_b, what="eager b")
    monkeypatch.delenv("CLAST_GRAPH")
    print("dp order graph replays", flush=True)
    for k in range(4):
        got_a = native.search_batch(p, ix, qa).arrays()
        assert_rows_equal(got_a, eager_a, what=f"graph replay {k} of a")
        monkeypatch.setenv("CLAST_GRAPH", "0" if k % 2 else "1")
        got_b = native.search_batch(p, ix, qb).arrays()
        assert_rows_equal(got_b, eager_b, what=f"search {k} of b")
        monkeypatch.delenv("CLAST_GRAPH")
    for h in (qa, qb, ix):
        h.close()


def test_dp_persistent_schedule_gap_limit(monkeypatch):
    # The persistent 16-bit schedule takes gaps up to 8: the tuner does not
    # try it for wider bands, and forcing it there is refused before any DP
    # launch; the static schedule gives the oracle's rows.
    t, r = workload(n_t=2, t_len=20_000, n_r=60, seed=17)
    p = native.Params(total_db_size=db_size(t), gap=16)
    tl, ql = lists(t, r, p.lmer)
    twins = native.windows(tl.gateway, 45_000, query=False)
    qwins = native.windows(ql.gateway, 40_000, query=True)
    ix = native.Index(p, tl, twins)
    q = native.Query(p, ql, qwins)
    monkeypatch.setenv("CLAST_DP_SCHED", "persist")
    with pytest.raises(native.ClastError, match="persist takes gaps up to 8"):
        native.search_batch(p, ix, q)
    monkeypatch.delenv("CLAST_DP_SCHED")
    h = native.search_batch(p, ix, q)
    assert h.dp_config()[0] == 0
    got = h.arrays()
    ix.close()
    q.close()
    assert_rows_equal(got, oracle.search(p, tl, twins, ql, qwins)["final"], what="gap 16 static")


@pytest.mark.parametrize("sched", ["static", "persist"])
def test_dp_two_pass_equals_one_pass(sched, monkeypatch):
    # Two-pass DP stage: the 16-bit tasks first as scores only, then the
    # tasks of the hits that pass the bad-hit filters again with match
    # counts (long ones one lane per band, the others in k_dp16, past-the-end
    # ones with 'X').  A target with repeats and reads from clean to noisy
    # (20% substitutions: extensions whose scores fail the E-value cut), some
    # over 1024 rows, some off a sequence end, some with X letters: the
    # two-pass rows equal the one-pass rows and the oracle's, the second pass
    # runs some but not all 16-bit tasks, and without the E-value cut
    # (cutOff -1) the stage runs one pass.
    monkeypatch.setenv("CLAST_TUNE", "0")
    monkeypatch.setenv("CLAST_DP_SCHED", sched)
    monkeypatch.setenv("CLAST_DP_BAND", "off")
    t, _ = workload(n_t=4, t_len=40_000, n_r=1, seed=91, repeat_frac=0.3)
    tb = t.bases.copy()
    o1 = int(t.offsets[1])
    tb[o1 + 2000:o1 + 40_000:1777] = ord("X")
    t = gen.Target(tb, t.offsets)
    parts = [gen.make_reads(gen.ReadSpec(92, 40, len_mode=gen.LEN_LOGUNIFORM, len_a=1100, len_b=3000,
                                         sub_rate=0.03, indel_rate=0.02), t),
             gen.make_reads(gen.ReadSpec(93, 40, len_mode=gen.LEN_LOGUNIFORM, len_a=300, len_b=2500,
                                         sub_rate=0.2, indel_rate=0.02), t),
             gen.make_reads(gen.ReadSpec(94, 150, sub_rate=0.05, indel_rate=0.01, repeat_read_frac=0.5), t)]
    seqs = [x.seq(i) for x in parts for i in range(len(x))]
    seqs += [t.seq(0)[-1400:] + b"ACG", b"TTGA" + t.seq(2)[:1300], t.seq(3)[-200:]]
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    r = gen.Reads(bases, off)
    for cut in (10.0, -1.0):
        p = native.Params(total_db_size=db_size(t), cut_off=cut)
        tl, ql = lists(t, r, p.lmer)
        twins = native.windows(tl.gateway, 90_000, query=False)
        qwins = native.windows(ql.gateway, 80_000, query=True)
        ix = native.Index(p, tl, twins)
        q = native.Query(p, ql, qwins)
        exp = oracle.search(p, tl, twins, ql, qwins)["final"]
        print(f"dp two pass {sched} cutOff {cut}", flush=True)
        got = {}
        for mode in ("1", "0"):
            monkeypatch.setenv("CLAST_DP_TWO_PASS", mode)
            h = native.search_batch(p, ix, q)
            got[mode] = h.arrays()
            d, c = h.dp_kernels(), h.counts()
            assert d[11] == (1.0 if mode == "1" and cut > 0 else 0.0)
            if d[11]:
                assert 0 < d[8] < c[4], (d[8], c[4])
                assert d[9] > 0
            h.close()
        for mode in ("1", "0"):
            assert_rows_equal(got[mode], exp, what=f"two pass={mode} {sched} cutOff {cut}")
        ix.close()
        q.close()
    monkeypatch.setenv("CLAST_DP_TWO_PASS", "2")
    with pytest.raises(native.ClastError, match="CLAST_DP_TWO_PASS"):
        native.search_batch(p, *_tiny_pair(p))


def _tiny_pair(p):
    t, r = workload(n_t=1, t_len=5_000, n_r=5, seed=95)
    tl, ql = lists(t, r, p.lmer)
    return (native.Index(p, tl, native.windows(tl.gateway, 90_000, query=False)),
            native.Query(p, ql, native.windows(ql.gateway, 80_000, query=True)))
