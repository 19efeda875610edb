"""MI355X-native RESMCM sparse matrix chain multiplication (row-wise sparsity estimator,
estimate-driven chain-order DP, row-wise Gustavson SpGEMM with uint64 path counts).

The compute path is ``lib/libsmcm.so`` (hand-written HIP kernels for gfx950 behind a C ABI,
declared in ``include/smcm.h``); this package only loads it and mirrors the reference's host
interface (``HeteroGraph``, ``MetaPath``, ``read_metapathesN``, ``build_pair``).
"""
from .hin import HeteroGraph, MetaPath, read_metapathesN  # noqa: F401

__all__ = ["HeteroGraph", "MetaPath", "read_metapathesN"]
