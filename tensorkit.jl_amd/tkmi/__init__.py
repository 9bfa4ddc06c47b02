"""tkmi — MI355X (gfx950) backend for TensorKit's symmetric-tensor hot path.

Host-side mirror of TensorKit's permute/braid and contraction API.  Spaces,
fusion trees and transform tables are built and cached on the host; all
tensor arithmetic runs in libtkmi.so (hand-written HIP kernels) through its
C ABI (include/tkmi.h).
"""
from .sectors import (U1Irrep, SU2Irrep, FermionParity, ProductSector, PlanarTrivial, fZ2U1, Sector,
                      UNIQUE, SIMPLE, BOSONIC, FERMIONIC, NOBRAIDING)
from .spaces import (GradedSpace, ProductSpace, HomSpace, U1Space, SU2Space, PlanarSpace, Vect, otimes,
                     SpaceMismatch, SectorMismatch)
from .fusiontrees import FusionTree, FusionTreeBlock, fusiontrees
from .treetransformers import treebraider, treetransposer, empty_globalcaches, global_cache_info
from . import _lib

__all__ = [
    "U1Irrep", "SU2Irrep", "FermionParity", "ProductSector", "PlanarTrivial", "fZ2U1", "Sector",
    "GradedSpace", "ProductSpace", "HomSpace", "U1Space", "SU2Space", "PlanarSpace", "Vect", "otimes",
    "SpaceMismatch", "SectorMismatch", "FusionTree", "FusionTreeBlock", "fusiontrees",
    "treebraider", "treetransposer", "empty_globalcaches", "global_cache_info",
]

