"""flam_amd: MI355X recursive skeletonization factorization (FLAM's rskelf).

Public surface (mirrors the reference's rskelf/rskelf_sv/rskelf_mv):

    from flam_amd import rskelf, rskelf_sv, rskelf_mv, LaplaceDL2D, ProxyRing

plus rskelf_batch (many independent problems in one factorization).
"""
from .rskelf import (Batch, FlamError, Factor, HelmholtzCFIE2D, LaplaceDL2D, LaplaceSL3D, ProxyPoints, ProxyRing,
                     kernel_matvec, kernel_matvec_batch, rskelf, rskelf_batch, rskelf_cholmv,
                     rskelf_cholsv, rskelf_logdet, rskelf_mv, rskelf_partial_info,
                     rskelf_partial_mv, rskelf_partial_sv, rskelf_sv, release_cache, stack_problems,
                     rskelf_import, export_all)

__all__ = ["Batch", "FlamError", "Factor", "HelmholtzCFIE2D", "LaplaceDL2D", "LaplaceSL3D", "ProxyPoints",
           "ProxyRing", "kernel_matvec",
           "kernel_matvec_batch", "rskelf", "rskelf_batch", "rskelf_cholmv", "rskelf_cholsv",
           "rskelf_logdet", "rskelf_mv", "rskelf_partial_info",
           "rskelf_partial_mv", "rskelf_partial_sv", "rskelf_sv", "release_cache", "stack_problems",
           "rskelf_import", "export_all"]
