"""MI355X (gfx950) implementation of SIGMo's batched subgraph-isomorphism hot path.

The compute path is ``libsigmo_hip.so`` (hand-written HIP kernels behind the C ABI of
``include/sigmo_hip.h``); :mod:`sigmo_amd.capi` is its ctypes binding, including the seeded
synthetic generator (``HostGraphs.molecules`` / ``HostGraphs.queries``), and
:mod:`sigmo_amd.build` compiles the libraries.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
