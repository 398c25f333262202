"""Host-side code without a GPU: the C-ABI library loads and exports every
declared entry point, the host list layout (and its gateway computed from
read lengths alone), windows, FASTA chunking, the driver's flag handling,
the preprocessDB and devideQuery tools and the checks of a persisted index
file."""
from __future__ import annotations

import ctypes
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from clast_amd import gen, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLAST = os.path.join(ROOT, "clast_amd", "bin", "clast")


def declared_functions():
    src = open(os.path.join(ROOT, "include", "clast_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(clast_[a-z_0-9]+)\s*\(", src)))


def test_abi_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(os.path.join(ROOT, "clast_amd", "lib", "libclast_hip.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert native.hip().clast_version().startswith(b"libclast_hip")


def py_list(seqs, lmer, query):
    """CHostSeqList layout restated from utilAddSequence.cu / CHostSeqList_query.cu."""
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    out, lens, gw = [], [], [0]
    for s in seqs:
        strands = [s]
        if query:
            rc = s[::-1].translate(comp)
            rc = bytes(c if c in b"ACGT" else ord("N") for c in rc)
            strands.append(rc)
        for x in strands:
            joint = (x + b"N" * lmer)[:lmer - 1]
            out.append(x + joint)
            lens.append(len(x))
            gw.append(gw[-1] + len(x) + lmer - 1)
    return b"".join(out), lens, gw


@pytest.mark.parametrize("query", [False, True])
def test_host_list_layout(query):
    seqs = [b"ACGTTGCA" * 5, b"ACG", b"", b"NNACGTRYKX" * 4, b"T" * 30]
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.int64)
    lst = native.build_list(bases, off, 15, query)
    b, lens, gw = py_list(seqs, 15, query)
    assert lst.base.tobytes() == b
    assert lst.length.tolist() == lens
    assert lst.gateway.tolist() == gw


@pytest.mark.parametrize("lmer", [1, 11, 15, 21])
def test_read_gateway_equals_host_list(lmer):
    # the strand gateway computed from read lengths alone (device-built query
    # sets) equals the host-built list's, short and empty reads included
    rng = np.random.default_rng(lmer)
    seqs = [bytes(rng.choice(list(b"ACGTN"), int(n))) for n in rng.integers(0, 400, 60)] + [b"", b"AC"]
    bases = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.int64)
    lst = native.build_list(bases, off, lmer, True)
    g = native.read_gateway(np.diff(off), lmer)
    assert g.tolist() == lst.gateway.tolist()
    for vram in (300, 5_000, 1e9):
        assert native.windows(g, vram, True) == native.windows(lst.gateway, vram, True)


def py_windows(gateway, vram, query):
    """moveTargetWindow / moveQueryWindow, CHostSchedular.cu:12-35."""
    out = []
    end = (len(gateway) - 1) // 2 if query else len(gateway) - 1
    b, e = 0, 1
    while b < end:
        bi = gateway[2 * b] if query else gateway[b]
        while e < end and (gateway[2 * e + 2] if query else gateway[e + 1]) - bi < vram:
            e += 1
        out.append((b, e))
        b, e = e, e + 1
    return out


@pytest.mark.parametrize("vram", [1, 500, 2000, 10_000, 1e9])
def test_windows(vram):
    rng = np.random.default_rng(1)
    lens = rng.integers(0, 900, 41)
    gw = np.concatenate([[0], np.cumsum(lens + 14)]).astype(np.int32)
    assert native.windows(gw, vram, False) == py_windows(gw.tolist(), vram, False)
    gq = np.concatenate([[0], np.cumsum(np.repeat(lens, 2) + 14)]).astype(np.int32)
    assert native.windows(gq, vram, True) == py_windows(gq.tolist(), vram, True)


def test_chunks_follow_loader_rule():
    # records accumulate while sum + len < budget; longer than the budget: skipped
    lens = np.array([30, 30, 39, 5, 200, 10, 50, 49], np.int64)
    ch, keep = native.chunks(lens, 100)
    assert ch == [(0, 3), (3, 7), (7, 8)]
    assert keep.tolist() == [1, 1, 1, 1, 0, 1, 1, 1]


def test_loader_skips_a_record_of_exactly_the_budget(tmp_path):
    # judgeContinueReading (CFASTALoader.cpp:43-87) adds a record while
    # sum + len < budget and skips one longer than the budget; a record of
    # exactly the budget at the start of a load fits neither way, and the
    # reference's loader would return to it forever.  This loader skips it
    # with a warning, like an over-long one: loads [r0], [r2 r3].
    import ctypes

    f = tmp_path / "q.fa"
    f.write_bytes(b">r0\n" + b"A" * 10 + b"\n>r1\n" + b"C" * 100 + b"\n>r2\n" + b"G" * 20 + b"\n>r3\n" + b"T" * 30
                  + b"\n")
    L = native.host()
    L.clast_host_fasta_chunks.restype = ctypes.c_int32
    L.clast_host_fasta_chunks.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int32,
                                          np.ctypeslib.ndpointer(np.int32), ctypes.c_int32]
    counts = np.zeros(8, np.int32)
    n = L.clast_host_fasta_chunks(str(f).encode(), 100, 0, counts, 8)
    assert counts[:n].tolist() == [1, 2]
    # and the in-memory chunking the package uses agrees
    ch, keep = native.chunks(np.array([10, 100, 20, 30], np.int64), 100)
    assert ch == [(0, 1), (1, 4)] and keep.tolist() == [1, 0, 1, 1]


def run_clast(args, cwd):
    return subprocess.run([CLAST, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=60)


@pytest.mark.parametrize("args,msg", [
    (["-t", "a.fa", "-q", "b.fa"], "no output file"),
    (["-q", "b.fa", "-o", "o"], "no target file"),
    (["-t", "a.fa", "-o", "o"], "no query file"),
    (["-t", "a.fa", "-q", "b.fa", "-o", "o", "-gap", "17"], "too large allowableGap"),
    (["-t", "a.fa", "-q", "b.fa", "-o", "o", "-gap", "1"], "allowableGap must be larger than 1"),
    (["-t", "a.fa", "-q", "b.fa", "-o", "o", "-lMer", "32"], "too large lMerLength"),
    (["-t", "a.fa", "-q", "b.fa", "-o", "o", "-lMer", "x15"], 'Check the after of "-lMer"'),
    (["-t", "a.fa", "-q", "b.fa", "-o", "o", "-cutOff", "1e"], 'Check the after of "-cutOff"'),
    (["-bogus"], "Check the command line parameter"),
    (["-t", "a.fa", "-t"], 'Check the after of "-t"'),
])
def test_driver_rejects_bad_flags(tmp_path, args, msg):
    r = run_clast(args, tmp_path)
    assert r.returncode != 0
    assert msg in r.stdout


def test_driver_prints_settings_before_device(tmp_path):
    # parsing and the settings printout need no GPU; missing files abort in the loader
    r = run_clast(["-t", "a.fa", "b.fa", "-q", "c.fa", "-o", "o.tsv", "-width", "50", "-tRAM", "1.5"], tmp_path)
    assert "allowable seed-seed width: 50" in r.stdout
    assert "RAM usage for target (MB): 1.5" in r.stdout
    assert "No.2 b.fa" in r.stdout


def index_file_head(lmer=15, stride=5, cut=20, nwin=1, nseq=1, nent=0, nslots=0, magic=b"CLASTIX1", version=1):
    """Header of a persisted index file: magic, version, lMer, stride, repeat
    cut, windows, sequences, entries, slots (host byte order, the u64 at
    offset 40)."""
    return struct.pack("<8sI5iI4xQ", magic, version, lmer, stride, cut, nwin, nseq, nent, nslots)


def index_file(head_kw, win_slot, id_begin, seq_slot, seq_len, seq_win, seq_ref, classes, epos, codes):
    return (index_file_head(**head_kw) + np.asarray(win_slot, "<u8").tobytes() + np.asarray(id_begin, "<i4").tobytes()
            + np.asarray(seq_slot, "<u8").tobytes() + np.asarray(seq_len, "<i4").tobytes()
            + np.asarray(seq_win, "<i4").tobytes() + np.asarray(seq_ref, "<i4").tobytes()
            + np.asarray(classes, np.uint8).tobytes() + np.asarray(epos, "<i4").tobytes()
            + np.asarray(codes, "<u8").tobytes())


def test_index_file_rejected_before_device(tmp_path):
    # clast_index_load checks a file on the host before any device work: a
    # missing file, another format, a size that does not match the header,
    # and tables the kernels could not index are CLAST_ERR_IO (5)
    lib = ctypes.CDLL(os.path.join(ROOT, "clast_amd", "lib", "libclast_hip.so"))
    lib.clast_index_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.clast_last_error.restype = ctypes.c_char_p

    def load(data, name="ix.bin"):
        path = tmp_path / name
        if data is not None:
            path.write_bytes(data)
        h = ctypes.c_void_p()
        rc = lib.clast_index_load(os.fsencode(str(path)), ctypes.byref(h))
        return rc, lib.clast_last_error().decode()

    # one window, one 20-base sequence (+ 14 joint slots), two entries
    good = dict(head_kw=dict(nent=2, nslots=34), win_slot=[0, 34], id_begin=[0], seq_slot=[0], seq_len=[20],
                seq_win=[0], seq_ref=[0], classes=[0, 1, 2, 3] * 8 + [0, 1], epos=[[0, 0], [0, 5]], codes=[7, 9])
    assert len(index_file(**good)) == 48 + 16 + 4 + 8 + 12 + 34 + 32

    def bad(**kw):
        d = dict(good)
        d.update(kw)
        return index_file(**d)

    cases = [
        (None, "cannot open"),
        (b"not an index", "truncated"),
        (index_file_head(magic=b"CLASTIX0") + bytes(100), "not an index file"),
        (index_file_head(version=2) + bytes(100), "unsupported version"),
        (index_file_head(lmer=32) + bytes(100), "header out of range"),
        (index_file(**good)[:-8], "size does not match"),
        (index_file(**good) + b"x", "size does not match"),
        (bad(win_slot=[0, 33]), "window table"),
        (bad(seq_len=[40]), "sequence table"),
        (bad(seq_win=[1]), "sequence table"),
        (bad(classes=[0] * 33 + [7]), "class bytes"),
        (bad(codes=[9, 7]), "entry codes"),
        (bad(codes=[7, 1 << 30]), "entry codes"),
        (bad(epos=[[0, 0], [1, 0]]), "entries"),
        (bad(epos=[[0, 0], [0, 6]]), "entries"),
    ]
    for i, (data, msg) in enumerate(cases):
        rc, err = load(data, f"ix{i}.bin")
        assert rc == 5 and msg in err, (i, rc, err)


PRE = os.path.join(ROOT, "clast_amd", "bin", "preprocessDB")


def run_pre(args, cwd):
    return subprocess.run([PRE, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=60)


def fasta_parts(text):
    """(label, start, sequence) of every part of a preprocessDB output."""
    parts = []
    for block in text.split("\n\n"):
        lines = [x for x in block.split("\n") if x]
        if lines:
            assert lines[1].startswith("@")
            parts.append((lines[0], int(lines[1][1:]), "".join(lines[2:])))
    return parts


def test_preprocess_db_known_parts(tmp_path):
    # preprocessDB.cpp's loops worked by hand: parts start at 0 and step
    # part - overlap while a whole part fits, then one part holds the rest
    # from the last start; 60-column lines, a part of exactly 60 bases gets
    # an empty last line; spaces are removed; an empty line ends a record
    # (lines after it until the next '>' are ignored); output is appended.
    (tmp_path / "in.fa").write_text(">x desc\nACGT ACGTAC\n>v\n" + "A" * 60 + "C" * 40 + "\n" + "G" * 30 +
                                    "\n\nTTTT\n>z\nGG\n")
    r = run_pre(["in.fa", "out.fa", "100", "30"], tmp_path)
    assert r.returncode == 0, r.stdout
    exp = (">x desc\n@0\nACGTACGTAC\n\n"
           ">v\n@0\n" + "A" * 60 + "\n" + "C" * 40 + "\n\n"
           ">v\n@70\n" + "C" * 30 + "G" * 30 + "\n\n\n"
           ">z\n@0\nGG\n\n")
    assert (tmp_path / "out.fa").read_text() == exp
    (tmp_path / "in2.fa").write_text(">x\nACGTACGTAC\n")
    assert run_pre(["in2.fa", "out.fa", "4", "1"], tmp_path).returncode == 0
    exp += ">x\n@0\nACGT\n\n>x\n@3\nTACG\n\n>x\n@6\nGTAC\n\n>x\n@9\nC\n\n"
    assert (tmp_path / "out.fa").read_text() == exp


def test_preprocess_db_round_trip(tmp_path):
    # every part is the input at its offset; parts start every part - overlap
    # bases, all whole but the last, which reaches the end (a 700-base
    # sequence gives [0, 700) and the 150 bases from 550)
    rng = np.random.default_rng(5)
    seqs = ["".join(rng.choice(list("ACGTN"), size=n)) for n in (5000, 700, 699, 1)]
    with open(tmp_path / "in.fa", "w") as f:
        for i, s in enumerate(seqs):
            f.write(f">s{i}\n")
            for k in range(0, len(s), 77):
                f.write(s[k:k + 77] + "\n")
    assert run_pre(["in.fa", "out.fa", "700", "150"], tmp_path).returncode == 0
    parts = fasta_parts((tmp_path / "out.fa").read_text())
    for i, s in enumerate(seqs):
        mine = [(st, q) for lab, st, q in parts if lab == f">s{i}"]
        starts = [st for st, _ in mine]
        assert starts == list(range(0, 550 * (len(starts) - 1) + 1, 550))
        for st, q in mine:
            assert q == s[st:st + len(q)]
        # whole parts while one fits, then the rest from the last start
        assert all(len(q) == 700 for _, q in mine[:-1]) and len(mine[-1][1]) <= 700
        assert mine[-1][0] + len(mine[-1][1]) == len(s)


def test_preprocess_db_arguments(tmp_path):
    (tmp_path / "in.fa").write_text(">x\nACGT\n")
    r = run_pre([], tmp_path)
    assert r.returncode == 0 and "usage" in r.stdout
    r = run_pre(["in.fa", "o.fa", "10", "10"], tmp_path)
    assert r.returncode != 0 and "larger than the overlap" in r.stdout
    r = run_pre(["in.fa", "o.fa", "10x", "2"], tmp_path)
    assert r.returncode != 0 and "not an integer" in r.stdout
    assert not (tmp_path / "o.fa").exists()
    # no records: no output file
    (tmp_path / "empty.fa").write_text("ACGT\n")
    assert run_pre(["empty.fa", "o.fa", "10", "2"], tmp_path).returncode == 0
    assert not (tmp_path / "o.fa").exists()


DIV = os.path.join(ROOT, "clast_amd", "bin", "devideQuery")


def test_devide_query_known_parts(tmp_path):
    # devideQuery.cpp's greedy rule worked by hand: each record, in file
    # order, goes to the part with the smallest total so far, the lowest
    # index on ties; written as label, sequence on one line, empty line
    (tmp_path / "q.fa").write_text(">a\nAC GTA\n>b\nCCC\n\n>c\nGG\nGG\n>d\nT\n")
    r = subprocess.run([DIV, "2", "q.fa"], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout
    # totals: a(5) -> 0 [5, 0]; b(3) -> 1 [5, 3]; c(4) -> 1 [5, 7]; d(1) -> 0 [6, 7]
    assert (tmp_path / "q.fa.0").read_text() == ">a\nACGTA\n\n>d\nT\n\n"
    assert (tmp_path / "q.fa.1").read_text() == ">b\nCCC\n\n>c\nGGGG\n\n"
    for args in ([], ["0", "q.fa"], ["2"], ["x", "q.fa"]):
        r = subprocess.run([DIV, *args], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=60)
        assert r.returncode != 0


def test_devide_query_round_trip(tmp_path):
    # every record lands in exactly one part, unchanged; parts differ in
    # total length by at most the longest record
    rng = np.random.default_rng(3)
    recs = [(f">r{i}", "".join(rng.choice(list("ACGT"), size=int(n)))) for i, n in
            enumerate(rng.integers(1, 400, size=300))]
    with open(tmp_path / "q.fa", "w") as f:
        for lab, s in recs:
            f.write(lab + "\n" + "\n".join(s[k:k + 60] for k in range(0, len(s), 60)) + "\n")
    assert subprocess.run([DIV, "7", "q.fa"], cwd=tmp_path, timeout=60).returncode == 0
    got, totals = [], []
    for k in range(7):
        blocks = [b.split("\n") for b in (tmp_path / f"q.fa.{k}").read_text().split("\n\n") if b]
        got += [(b[0], b[1]) for b in blocks]
        totals.append(sum(len(b[1]) for b in blocks))
    assert sorted(got) == sorted(recs)
    assert max(totals) - min(totals) <= max(len(s) for _, s in recs)


def test_poisson_requests_partition_reads():
    from clast_amd import serve

    rng = np.random.default_rng(0)
    lens = rng.integers(50, 400, 5000)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    bases = rng.choice(np.frombuffer(b"ACGT", np.uint8), int(off[-1]))
    reqs = serve.poisson_requests(bases, off, rate=50.0, seed=1, max_reads=400)
    assert sum(len(r.offsets) - 1 for r in reqs) == 5000
    assert b"".join(r.bases.tobytes() for r in reqs) == bases.tobytes()
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0)
    # mean inter-arrival close to 1/rate
    assert 0.01 < float(np.mean(np.diff(arr))) < 0.04


def test_score_threshold_equals_bisection(tmp_path):
    """The per-length score threshold of the E-value cut (closed-form start,
    stepped to the exact boundary) against plain bisection on the same
    expression, over a grid of K, lambda, database size, cut-off and query
    length, with cut-offs exactly on a score's E-value, and every length up
    to 20000 at the reference defaults (tests/threshold_check.cpp)."""
    here = os.path.dirname(os.path.abspath(__file__))
    exe = str(tmp_path / "threshold_check")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(here, "threshold_check.cpp"), "-o", exe], check=True)
    res = subprocess.run([exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout
    m = re.search(r"cases (\d+) mismatches 0", res.stdout)
    assert m and int(m.group(1)) > 20000, res.stdout
