"""The clast driver end to end on FASTA files: the per-window path and the
batched path produce the oracle's tab-separated output."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle
from clast_amd import gen
from common import workload

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLAST = os.path.join(ROOT, "clast_amd", "bin", "clast")
PRE = os.path.join(ROOT, "clast_amd", "bin", "preprocessDB")


def write_inputs(d):
    t, r = workload(n_t=7, t_len=30_000, n_r=420, repeat_frac=0.15, random_frac=0.05, indel=0.015)
    gen.write_fasta(os.path.join(d, "t1.fa"), t.subset(0, 4), "chr", width=70)
    # second target file: an '@' offset record, lower-case and odd letters
    with open(os.path.join(d, "t2.fa"), "wb") as f:
        for i in range(4, 7):
            s = t.seq(i)
            f.write(f">chr{i} part\n".encode())
            if i == 5:
                f.write(b"@1000000\n")
            for k in range(0, len(s), 60):
                line = s[k:k + 60]
                if i == 6 and k == 600:
                    line = line.lower()
                if i == 6 and k == 1200:
                    line = line[:10] + b"-*" + line[12:]
                f.write(line + b"\n")
            f.write(b"\n")
    gen.write_fasta(os.path.join(d, "q1.fa"), r.subset(0, 300), "read", width=80)
    gen.write_fasta(os.path.join(d, "q2.fa"), r.subset(300, 420), "more", width=50)
    return ["-t", "t1.fa", "t2.fa", "-q", "q1.fa", "q2.fa"]


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


def same_tsv(a: str, b: str):
    la, lb = a.splitlines(), b.splitlines()
    assert len(la) == len(lb), (len(la), len(lb))
    for i, (x, y) in enumerate(zip(la, lb)):
        if x == y:
            continue
        fx, fy = x.split("\t"), y.split("\t")
        assert fx[:9] == fy[:9], f"row {i}: {x!r} vs {y!r}"
        assert abs(float(fx[9]) - float(fy[9])) <= 1e-5 * abs(float(fy[9])), f"row {i}: {x!r} vs {y!r}"


FLAGS = [
    [],
    ["-tRAM", "0.07", "-tVRAM", "0.05", "-qRAM", "0.03", "-qVRAM", "0.008"],
    ["-numOut", "1", "-qVRAM", "0.01"],
    ["-local", "1", "-gap", "5", "-cutOff", "-1", "-qVRAM", "0.02"],
    ["-lMer", "12", "-stride", "3", "-repeat", "8", "-width", "40", "-tVRAM", "0.1"],
]


@pytest.mark.parametrize("flags", FLAGS, ids=[" ".join(f) or "defaults" for f in FLAGS])
def test_driver_tsv_matches_oracle(tmp_path, flags):
    d = str(tmp_path)
    inputs = write_inputs(d)
    print(f"driver {flags}", flush=True)
    run([oracle.EXE, *inputs, "-o", "exp.tsv", *flags], d)
    run([CLAST, *inputs, "-o", "plain.tsv", *flags], d)
    run([CLAST, *inputs, "-o", "batch.tsv", "-batch", "1", *flags], d)
    exp = open(os.path.join(d, "exp.tsv")).read()
    assert len(exp.splitlines()) > 100
    same_tsv(open(os.path.join(d, "plain.tsv")).read(), exp)
    same_tsv(open(os.path.join(d, "batch.tsv")).read(), exp)


def test_driver_split_target_matches_oracle(tmp_path):
    # targets split by preprocessDB into overlapping '@offset' parts shorter
    # than the sequences (reads fall inside parts and across part
    # boundaries): the driver's rows on the split file equal the oracle
    # CLI's, with target positions in the whole sequence
    d = str(tmp_path)
    t, r = workload(n_t=3, t_len=60_000, n_r=400, repeat_frac=0.1, seed=41)
    gen.write_fasta(os.path.join(d, "t.fa"), t, "chr", width=70)
    gen.write_fasta(os.path.join(d, "q.fa"), r, "read", width=80)
    run([PRE, "t.fa", "parts.fa", "20000", "1000"], d)
    inputs = ["-t", "parts.fa", "-q", "q.fa"]
    print("driver on a split target", flush=True)
    run([oracle.EXE, *inputs, "-o", "exp.tsv"], d)
    run([CLAST, *inputs, "-o", "plain.tsv"], d)
    run([CLAST, *inputs, "-o", "batch.tsv", "-batch", "1"], d)
    exp = open(os.path.join(d, "exp.tsv")).read()
    assert len(exp.splitlines()) > 100
    assert max(int(x.split("\t")[5]) for x in exp.splitlines()) > 40_000
    same_tsv(open(os.path.join(d, "plain.tsv")).read(), exp)
    same_tsv(open(os.path.join(d, "batch.tsv")).read(), exp)
