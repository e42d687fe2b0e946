"""flam_amd: MI355X recursive skeletonization factorization (FLAM's rskelf).

Public surface (mirrors the reference's rskelf/rskelf_sv/rskelf_mv):

    from flam_amd import rskelf, rskelf_sv, rskelf_mv, LaplaceDL2D, ProxyRing
"""
from .rskelf import (FlamError, Factor, LaplaceDL2D, ProxyRing, kernel_matvec, rskelf,
                     rskelf_mv, rskelf_partial_info, rskelf_partial_mv, rskelf_partial_sv,
                     rskelf_sv)

__all__ = ["FlamError", "Factor", "LaplaceDL2D", "ProxyRing", "kernel_matvec", "rskelf",
           "rskelf_mv", "rskelf_partial_info", "rskelf_partial_mv", "rskelf_partial_sv",
           "rskelf_sv"]
