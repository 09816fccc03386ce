"""graphindor_amd -- MI355X-native schedule-driven subgraph enumeration.

The product is native: libgm_hip.so (HIP kernels behind the C ABI of
include/graphindor_amd.h) and libgraphset.so (the host C++ library with the
reference's Graph / Pattern / Schedule_IEP interface), plus the gpu_graph,
gpu_clique and gpu_mc command lines.  This Python package only binds them
(graphindor_amd.native) and builds them (graphindor_amd.build).
"""
from . import native  # noqa: F401

__all__ = ["native"]
