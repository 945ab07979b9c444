"""MI355X (gfx950) per-step numerical core of NRG Ljubljana: batched eigensolver, grouped
recalculation products, density-matrix rotations and FDM spectral binning, as HIP kernels behind
the C ABI of include/nrgcore.h; a whole NRG calculation on top of them (include/nrgrun.h)."""
from .core import (Context, NrgError, lib, LIB_PATH, NRG_SYEV_FP64, NRG_SYEV_MIXED, NRG_GEMM_FP64,
                   NRG_GEMM_OZAKI, Gemm3Term, FdmPair, BinGrid, HamTerm, NRG_HAM_BLOCK, NRG_HAM_DIAG,
                   NRG_HAM_IDENT, run_nrg, runlib, RUN_LIB_PATH, TABLES_DIR)

__all__ = ["Context", "NrgError", "lib", "LIB_PATH", "NRG_SYEV_FP64", "NRG_SYEV_MIXED", "NRG_GEMM_FP64",
           "NRG_GEMM_OZAKI", "Gemm3Term", "FdmPair", "BinGrid", "HamTerm", "NRG_HAM_BLOCK", "NRG_HAM_DIAG",
           "NRG_HAM_IDENT", "run_nrg", "runlib", "RUN_LIB_PATH", "TABLES_DIR"]
