"""Seeded synthetic workloads (libclast_gen.so) and FASTA I/O.

The five workloads of the search path, by name:

* ``small``   — 10 Mbp target as 100 x 100 kbp uniform ACGT with 0.1% N;
  10,000 reads of 150 bp, 50% reverse complement, 5% substitutions, 1% indels
  (geometric length p=0.5, max 3).  Seed 1.
* ``large``   — 1 Gbp target as 1,000 x 1 Mbp with 10% repeat content (200
  families of 300-6,000 bp, copies at 15% divergence); 1,000,000 reads of
  150 bp with 3% substitutions and 0.5% indels.  Seed 2.
* ``uneven``  — the large target; 200,000 reads of log-uniform 100 bp-20 kbp
  length, 8% indels, 4% substitutions, 20% random no-hit reads.  Seed 3.
* ``serving`` — 3 Gbp target (3,000 x 1 Mbp, 40% repeats from 1,000 families
  at 5-25% divergence, 1% (AT)n/(CA)n tracts); requests of log-uniform
  1-10,000 reads, lengths 50/50 150 bp and log-uniform 1-15 kbp.  Seed 4.
* ``pressure`` — the serving target; 50% random reads, 25% from repeat
  copies, 25% reverse complement, lengths log-uniform 100 bp-20 kbp.  Seed 5.

``scale`` shrinks sequence and read counts for tests (sequence lengths stay).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .native import LIB_DIR


class CTargetSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_seq", C.c_int32), ("seq_len", C.c_int64), ("n_rate", C.c_double),
                ("repeat_frac", C.c_double), ("n_families", C.c_int32), ("fam_min", C.c_int32),
                ("fam_max", C.c_int32), ("div_min", C.c_double), ("div_max", C.c_double), ("lowc_frac", C.c_double)]


class CReadSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_reads", C.c_int64), ("len_mode", C.c_int32), ("len_a", C.c_int32),
                ("len_b", C.c_int32), ("len_c", C.c_int32), ("mix_frac", C.c_double), ("sub_rate", C.c_double),
                ("indel_rate", C.c_double), ("indel_p", C.c_double), ("indel_max", C.c_int32),
                ("rc_frac", C.c_double), ("random_frac", C.c_double), ("repeat_read_frac", C.c_double)]


LEN_FIXED, LEN_LOGUNIFORM, LEN_MIXED = 0, 1, 2


@dataclass
class TargetSpec:
    seed: int
    n_seq: int
    seq_len: int
    n_rate: float = 0.0
    repeat_frac: float = 0.0
    n_families: int = 0
    fam_min: int = 300
    fam_max: int = 6000
    div_min: float = 0.0
    div_max: float = 0.0
    lowc_frac: float = 0.0


@dataclass
class ReadSpec:
    seed: int
    n_reads: int
    len_mode: int = LEN_FIXED
    len_a: int = 150
    len_b: int = 150
    len_c: int = 150
    mix_frac: float = 0.5
    sub_rate: float = 0.0
    indel_rate: float = 0.0
    indel_p: float = 0.5
    indel_max: int = 3
    rc_frac: float = 0.5
    random_frac: float = 0.0
    repeat_read_frac: float = 0.0


@dataclass
class Seqs:
    """Concatenated sequences: bases (uint8) and n+1 offsets."""
    bases: np.ndarray
    offsets: np.ndarray

    def __len__(self) -> int:
        return len(self.offsets) - 1

    @property
    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets)

    def subset(self, a: int, b: int) -> "Seqs":
        lo, hi = int(self.offsets[a]), int(self.offsets[b])
        return Seqs(self.bases[lo:hi], self.offsets[a:b + 1] - lo)

    def seq(self, i: int) -> bytes:
        return self.bases[self.offsets[i]:self.offsets[i + 1]].tobytes()


@dataclass
class Target(Seqs):
    copy_pos: np.ndarray | None = None
    copy_len: np.ndarray | None = None


@dataclass
class Reads(Seqs):
    origin: np.ndarray | None = None
    strand: np.ndarray | None = None


_gen = None


def _lib() -> C.CDLL:
    global _gen
    if _gen is None:
        path = os.path.join(LIB_DIR, "libclast_gen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.clast_gen_target_create.restype = vp
        L.clast_gen_target_create.argtypes = [C.POINTER(CTargetSpec)]
        L.clast_gen_target_size.restype = C.c_int64
        L.clast_gen_target_size.argtypes = [vp]
        L.clast_gen_target_ncopies.restype = C.c_int64
        L.clast_gen_target_ncopies.argtypes = [vp]
        L.clast_gen_target_copy.argtypes = [vp, vp, vp, vp]
        L.clast_gen_target_free.argtypes = [vp]
        L.clast_gen_reads_create.restype = vp
        L.clast_gen_reads_create.argtypes = [C.POINTER(CReadSpec), vp, vp, C.c_int32, vp, vp, C.c_int64]
        L.clast_gen_reads_total.restype = C.c_int64
        L.clast_gen_reads_total.argtypes = [vp]
        L.clast_gen_reads_copy.argtypes = [vp, vp, vp, vp, vp]
        L.clast_gen_reads_free.argtypes = [vp]
        _gen = L
    return _gen


def make_target(spec: TargetSpec) -> Target:
    L = _lib()
    cs = CTargetSpec(spec.seed, spec.n_seq, spec.seq_len, spec.n_rate, spec.repeat_frac, spec.n_families,
                     spec.fam_min, spec.fam_max, spec.div_min, spec.div_max, spec.lowc_frac)
    h = L.clast_gen_target_create(C.byref(cs))
    try:
        n = L.clast_gen_target_size(h)
        nc = L.clast_gen_target_ncopies(h)
        bases = np.empty(n, np.uint8)
        cp = np.empty(max(nc, 1), np.int64)
        cl = np.empty(max(nc, 1), np.int32)
        L.clast_gen_target_copy(h, bases.ctypes.data, cp.ctypes.data, cl.ctypes.data)
    finally:
        L.clast_gen_target_free(h)
    offsets = np.arange(spec.n_seq + 1, dtype=np.int64) * spec.seq_len
    return Target(bases, offsets, cp[:nc], cl[:nc])


def make_reads(spec: ReadSpec, target: Target) -> Reads:
    L = _lib()
    cs = CReadSpec(spec.seed, spec.n_reads, spec.len_mode, spec.len_a, spec.len_b, spec.len_c, spec.mix_frac,
                   spec.sub_rate, spec.indel_rate, spec.indel_p, spec.indel_max, spec.rc_frac, spec.random_frac,
                   spec.repeat_read_frac)
    nc = 0 if target.copy_pos is None else len(target.copy_pos)
    cp = target.copy_pos if nc else np.zeros(1, np.int64)
    cl = target.copy_len if nc else np.zeros(1, np.int32)
    off = np.ascontiguousarray(target.offsets, dtype=np.int64)
    h = L.clast_gen_reads_create(C.byref(cs), target.bases.ctypes.data, off.ctypes.data, len(target), cp.ctypes.data,
                                 cl.ctypes.data, nc)
    try:
        n = L.clast_gen_reads_total(h)
        bases = np.empty(max(n, 1), np.uint8)
        offsets = np.empty(spec.n_reads + 1, np.int64)
        origin = np.empty(max(spec.n_reads, 1), np.int64)
        strand = np.empty(max(spec.n_reads, 1), np.int8)
        L.clast_gen_reads_copy(h, bases.ctypes.data, offsets.ctypes.data, origin.ctypes.data, strand.ctypes.data)
    finally:
        L.clast_gen_reads_free(h)
    return Reads(bases[:n], offsets, origin[:spec.n_reads], strand[:spec.n_reads])


def workload(name: str, scale: float = 1.0) -> tuple[TargetSpec, ReadSpec]:
    """Target and read specs of a named workload (module docstring)."""
    def s(n: int) -> int:
        return max(1, int(round(n * scale)))

    if name == "small":
        return (TargetSpec(1, s(100), 100_000, n_rate=0.001),
                ReadSpec(1, s(10_000), sub_rate=0.05, indel_rate=0.01))
    if name in ("large", "uneven"):
        t = TargetSpec(2, s(1000), 1_000_000, repeat_frac=0.10, n_families=200, fam_min=300, fam_max=6000,
                       div_min=0.15, div_max=0.15)
        if name == "large":
            return t, ReadSpec(2, s(1_000_000), sub_rate=0.03, indel_rate=0.005)
        return t, ReadSpec(3, s(200_000), len_mode=LEN_LOGUNIFORM, len_a=100, len_b=20_000, sub_rate=0.04,
                           indel_rate=0.08, random_frac=0.20)
    if name in ("serving", "pressure"):
        t = TargetSpec(4, s(3000), 1_000_000, repeat_frac=0.40, n_families=1000, fam_min=300, fam_max=6000,
                       div_min=0.05, div_max=0.25, lowc_frac=0.01)
        if name == "serving":
            return t, ReadSpec(4, s(10_000), len_mode=LEN_MIXED, len_a=1000, len_b=15_000, len_c=150, mix_frac=0.5,
                               sub_rate=0.04, indel_rate=0.08)
        return t, ReadSpec(5, s(10_000), len_mode=LEN_LOGUNIFORM, len_a=100, len_b=20_000, sub_rate=0.04,
                           indel_rate=0.08, random_frac=0.5, repeat_read_frac=0.5, rc_frac=0.5)
    raise ValueError(f"unknown workload {name!r}")


def write_fasta(path: str, seqs: Seqs, prefix: str, width: int = 80) -> None:
    """FASTA with labels `{prefix}{i}` and `width`-column sequence lines."""
    with open(path, "wb") as f:
        for i in range(len(seqs)):
            f.write(f">{prefix}{i}\n".encode())
            s = seqs.bases[seqs.offsets[i]:seqs.offsets[i + 1]]
            for k in range(0, len(s), width):
                f.write(s[k:k + width].tobytes())
                f.write(b"\n")
