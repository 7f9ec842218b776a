"""nwchem_amd -- MI355X-native CCSD(T) perturbative triples (the TCE (T) path
of NWChem src/tce/ccsd_t) behind the reference's own kernel surface.

Native code: libtce_t.so (nwchem_amd/csrc, hand-written HIP for gfx950).
Public C-ABI: include/tce_t.h.  Python here is a thin ctypes layer.
"""
from ._lib import build, lib  # noqa: F401
from .job import (BF16X3, BF16X6, FP32, FP64, TILING_ORBITAL, TILING_REFERENCE, TILING_SUBTILE,  # noqa: F401
                  TriplesBatch, TriplesJob, init)
from .server import TriplesServer  # noqa: F401
from .tiles import Blocks, TileModel, synth_blocks, synth_dense, tile_ranges  # noqa: F401

__all__ = ["build", "lib", "init", "TriplesJob", "TriplesBatch", "TriplesServer", "FP64", "FP32", "BF16X3", "BF16X6", "TileModel", "Blocks", "synth_blocks",
           "synth_dense", "tile_ranges"]
