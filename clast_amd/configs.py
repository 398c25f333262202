"""The benchmark workloads at their own sizes, cut into RAM chunks and VRAM
windows the way the reference's driver cuts them.

Chunks follow CFASTALoader (main.cu:40-67: targets by tRAM bases, queries by
qRAM/2 bases, CFASTALoader.cpp:43-87); windows follow CHostSchedular
(moveTargetWindow / moveQueryWindow, CHostSchedular.cu:12-35, windows of at
most tVRAM / qVRAM slots, query windows in '+'/'-' pairs).  The reference
defaults are tRAM = qRAM = tVRAM = 64 MBase and qVRAM = 2 MBase
(common.hpp:12-25).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import gen, native

T_RAM = 64e6
T_VRAM = 64e6
Q_RAM = 64e6
Q_VRAM = 2e6


@dataclass
class ReadChunk:
    """One query RAM chunk: its reads as given (for device encoding), the
    host list (for host encoding and the oracle) and its pair windows."""
    bases: np.ndarray
    offsets: np.ndarray
    first_read: int
    wins: list
    n_bases: int
    _list: native.HostList | None = None
    lmer: int = 15

    @property
    def lst(self) -> native.HostList:
        if self._list is None:
            self._list = native.build_list(self.bases, self.offsets, self.lmer, True)
        return self._list


def target_chunks(t: gen.Seqs, p: native.Params, t_ram: float = T_RAM, t_vram: float = T_VRAM):
    """[(host list, windows)] per target RAM chunk; records longer than the
    RAM budget are skipped as the loader skips them."""
    chunks, keep = native.chunks(t.lengths, int(t_ram))
    out = []
    for a, b in chunks:
        idx = [i for i in range(a, b) if keep[i]]
        if not idx:
            continue
        sub = t.subset(idx[0], idx[-1] + 1)
        lst = native.build_list(sub.bases, sub.offsets, p.lmer, False)
        out.append((lst, native.windows(lst.gateway, t_vram, False)))
    return out


def read_chunks(r: gen.Seqs, p: native.Params, q_ram: float = Q_RAM, q_vram: float = Q_VRAM) -> list[ReadChunk]:
    """Query RAM chunks (qRAM/2 bases of reads each) with the pair windows
    the scheduler would cut from their strands; host lists are built on
    first use."""
    chunks, keep = native.chunks(r.lengths, int(q_ram / 2))
    out = []
    for a, b in chunks:
        idx = [i for i in range(a, b) if keep[i]]
        if not idx:
            continue
        lo, hi = idx[0], idx[-1] + 1
        b0, b1 = int(r.offsets[lo]), int(r.offsets[hi])
        offs = (r.offsets[lo:hi + 1] - b0).astype(np.int64)
        g = native.read_gateway(np.diff(offs), p.lmer)
        out.append(ReadChunk(r.bases[b0:b1], offs, lo, native.windows(g, q_vram, True), b1 - b0, lmer=p.lmer))
    return out


def workload(name: str, scale: float = 1.0):
    """(target, reads, params) of a named workload with the reference
    defaults; the E-value database size is every target base (the FASTA's
    sequence-line bytes, CHostSetting.cu:46-58)."""
    tspec, rspec = gen.workload(name, scale)
    t = gen.make_target(tspec)
    r = gen.make_reads(rspec, t)
    p = native.Params(total_db_size=float(t.bases.shape[0]))
    return t, r, p
