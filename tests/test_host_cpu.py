"""Host-side code without a GPU: the C-ABI library loads and exports every
declared entry point, the host list layout (and its gateway computed from
read lengths alone), windows, FASTA chunking and the driver's flag
handling."""
from __future__ import annotations

import ctypes
import os
import re
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
