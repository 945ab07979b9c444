"""MI355X (gfx950) per-step numerical core of NRG Ljubljana: batched eigensolver, grouped
recalculation products, density-matrix rotations and FDM spectral binning, as HIP kernels behind
the C ABI of include/nrgcore.h."""
from .core import (Context, NrgError, lib, LIB_PATH, NRG_SYEV_FP64, NRG_SYEV_MIXED, NRG_GEMM_FP64,
                   NRG_GEMM_OZAKI, Gemm3Term, FdmPair, BinGrid, HamTerm, NRG_HAM_BLOCK, NRG_HAM_DIAG,
                   NRG_HAM_IDENT)

__all__ = ["Context", "NrgError", "lib", "LIB_PATH", "NRG_SYEV_FP64", "NRG_SYEV_MIXED", "NRG_GEMM_FP64",
           "NRG_GEMM_OZAKI", "Gemm3Term", "FdmPair", "BinGrid", "HamTerm", "NRG_HAM_BLOCK", "NRG_HAM_DIAG",
           "NRG_HAM_IDENT"]
