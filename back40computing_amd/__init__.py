"""back40computing_amd — MI355X (gfx950) single-GPU BFS in the shape of b40c.

The product is the shared library ``libb40c_mi355x.so`` (hand-written HIP for
gfx950 behind the C ABI in ``include/b40c_mi355x.h``).  These Python modules
are a thin host mirror of the reference's C++ interface over that ABI.
"""
from . import _lib
from ._lib import (B40CError, CONTRACT_EXPAND, DIRECTION_OPTIMIZING, EDGE_FRONTIERS, EXPAND_CONTRACT,
                   HYBRID, MIXED_FRONTIERS, MODE_DEFAULT, MODE_FUSED, MODE_ITERATIVE, PERSISTENT,
                   TWO_PHASE, VERTEX_FRONTIERS, lib)

__all__ = [
    "B40CError", "CONTRACT_EXPAND", "DIRECTION_OPTIMIZING", "EDGE_FRONTIERS", "EXPAND_CONTRACT", "HYBRID",
    "MIXED_FRONTIERS", "MODE_DEFAULT", "MODE_FUSED", "MODE_ITERATIVE", "PERSISTENT", "TWO_PHASE",
    "VERTEX_FRONTIERS", "lib",
]
