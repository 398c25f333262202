"""In-tree build of the native parts of clast_amd.

* ``libclast_hip.so``  — the gfx950 HIP kernels behind the C ABI of
  ``include/clast_hip.h`` (hipcc, cross-compiled; no GPU needed to build).
* ``libclast_gen.so``  — the seeded synthetic-data generator (g++).
* ``libclast_host.so`` + ``clast`` — the C++ host driver (CLI, FASTA chunking,
  windows, result holder, TSV writer) as a library and an executable, both
  linked against ``libclast_hip.so``.

Everything lands in ``clast_amd/lib`` and ``clast_amd/bin`` so the built files
travel with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
GEN = os.path.join(PKG, "gen")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# outputs are written under a per-process name, then renamed into place
# (builds in parallel test workers do not clobber each other)
_TMP = f".tmp{os.getpid()}"
ARCH = "gfx950"

# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

HIP_UNITS = ["clast_hip.hip", "dp_global.hip", "dp_local.hip", "dp16.hip", "dp16_score.hip", "dp_band.hip"]
HOST_UNITS = ["settings.cpp", "fasta.cpp", "seqlist.cpp", "holder.cpp", "scheduler.cpp", "host_capi.cpp"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def build_hip() -> str:
    os.makedirs(LIB, exist_ok=True)
    objs = []
    cmds = []
    for unit in HIP_UNITS:
        src = os.path.join(CSRC, unit)
        obj = os.path.join(LIB, unit.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIB, "libclast_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out + _TMP])
    os.replace(out + _TMP, out)
    for o in objs:
        os.remove(o)
    return out


def build_gen() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libclast_gen.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", os.path.join(GEN, "clast_gen.cpp"), "-o", out + _TMP])
    os.replace(out + _TMP, out)
    return out


def build_host() -> str:
    """libclast_host.so (driver code as a library), the clast executable and
    the preprocessDB and devideQuery tools."""
    os.makedirs(BIN, exist_ok=True)
    srcs = [os.path.join(HOST, u) for u in HOST_UNITS]
    lib = os.path.join(LIB, "libclast_host.so")
    common = ["g++", "-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}"]
    _run([*common, "-shared", *srcs, f"-L{LIB}", "-lclast_hip", "-Wl,-rpath,$ORIGIN", "-o", lib + _TMP])
    os.replace(lib + _TMP, lib)
    exe = os.path.join(BIN, "clast")
    _run([*common, os.path.join(HOST, "main.cpp"), f"-L{LIB}", "-lclast_host", "-lclast_hip",
          "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + _TMP])
    os.replace(exe + _TMP, exe)
    for src, name in (("preprocess_db.cpp", "preprocessDB"), ("divide_query.cpp", "devideQuery")):
        tool = os.path.join(BIN, name)
        _run(["g++", "-O3", "-std=c++17", os.path.join(HOST, src), "-o", tool + _TMP])
        os.replace(tool + _TMP, tool)
    return exe


SANITIZE = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def build_host_selftest(sanitize: bool = True) -> str:
    """bin/host_selftest_san: the host driver's CPU side (host/selftest.cpp)
    built with the address and undefined-behaviour sanitizers (g++, host
    code only; the HIP library it links is not instrumented)."""
    os.makedirs(BIN, exist_ok=True)
    srcs = [os.path.join(HOST, u) for u in HOST_UNITS] + [os.path.join(HOST, "selftest.cpp")]
    exe = os.path.join(BIN, "host_selftest_san" if sanitize else "host_selftest")
    flags = SANITIZE if sanitize else ["-O2"]
    _run(["g++", *flags, "-std=c++17", f"-I{INCLUDE}", *srcs, f"-L{LIB}", "-lclast_hip",
          "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + _TMP])
    os.replace(exe + _TMP, exe)
    return exe


def build_all() -> None:
    with ThreadPoolExecutor(max_workers=2) as ex:
        f_hip = ex.submit(build_hip)
        f_gen = ex.submit(build_gen)
        f_hip.result()
        f_gen.result()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built", LIB, BIN)
