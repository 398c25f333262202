"""Address and undefined-behaviour sanitizer runs of the CPU code (no GPU):
the host driver's CPU side (flags, FASTA chunk loading, host lists, windows,
TSV writer: clast_amd/host/selftest.cpp) and the CPU oracle's CLI, each built
with -fsanitize=address,undefined and -fno-sanitize-recover=all, so that any
report fails the run.  The sanitized oracle's TSV must equal the regular
build's."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle
from clast_amd import build, gen
from common import workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ORACLE = os.path.join(ROOT, "oracle", "build_san", "clast_oracle")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


@pytest.fixture(scope="module")
def san_builds():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, stdout=subprocess.DEVNULL)
    return build.build_host_selftest(sanitize=True)


def test_host_driver_under_sanitizers(san_builds, tmp_path):
    res = subprocess.run([san_builds, str(tmp_path)], env=ENV, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=120)
    assert res.returncode == 0 and "selftest ok" in res.stdout, res.stdout[-3000:]
    assert "ERROR: AddressSanitizer" not in res.stdout and "runtime error" not in res.stdout


@pytest.mark.parametrize("flags", [[], ["-local", "1", "-gap", "5"], ["-numOut", "1", "-tVRAM", "0.02"]])
def test_oracle_under_sanitizers(san_builds, tmp_path, flags):
    t, r = workload(n_t=3, t_len=12_000, n_r=60, repeat_frac=0.2, seed=5)
    gen.write_fasta(str(tmp_path / "t.fa"), t, "chr")
    gen.write_fasta(str(tmp_path / "q.fa"), r, "read")
    outs = []
    for exe in (SAN_ORACLE, oracle.EXE):
        out = tmp_path / (os.path.basename(os.path.dirname(exe)) + ".tsv")
        res = subprocess.run([exe, "-t", str(tmp_path / "t.fa"), "-q", str(tmp_path / "q.fa"), "-o", str(out),
                              "-threads", "2", *flags], env=ENV, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True, timeout=300)
        assert res.returncode == 0, res.stdout[-3000:]
        assert "ERROR: AddressSanitizer" not in res.stdout and "runtime error" not in res.stdout
        outs.append(out.read_text())
    assert outs[0] == outs[1] and len(outs[0].splitlines()) > 20
