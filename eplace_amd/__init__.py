"""MI355X-native ePlace global-placement engine.

HIP kernels for gfx950 behind the C ABI in include/eplace_amd.h
(libeplace_amd.so); this package is the host-side mirror of the reference's
EPlacer_2D / EplaceNesterovOpt call surface plus design I/O.
"""
from .design import Design, read_bookshelf, synthetic, write_bookshelf, write_pl  # noqa: F401
from .placer import CGP, FILLERONLY, MGP, EPlacer2D, EplaceNesterovOpt  # noqa: F401
from .qplace import quadratic_placement  # noqa: F401

__all__ = ["quadratic_placement", "Design", "read_bookshelf", "synthetic", "write_bookshelf", "write_pl", "EPlacer2D",
           "EplaceNesterovOpt", "MGP", "FILLERONLY", "CGP"]
