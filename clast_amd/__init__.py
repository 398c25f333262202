"""clast_amd — CLAST's search pipeline on MI355X (gfx950).

The device stages live in ``libclast_hip.so`` (hand-written HIP kernels behind
the C ABI of ``include/clast_hip.h``); ``libclast_host.so`` and the ``clast``
executable are the C++ host driver around them.  This package holds the
ctypes bindings (:mod:`clast_amd.native`), the seeded workload generator
(:mod:`clast_amd.gen`) and the in-tree build (:mod:`clast_amd.build`).
"""
__version__ = "0.1.0"
