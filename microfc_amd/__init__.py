"""MI355X-native implementation of MicroFC's simulation right-hand side.

The compute path is the HIP library libmfc_rhs.so (C ABI in include/mfc_rhs.h);
this package holds its host-side mirror of the reference's interfaces
(rhs.RhsModule), the case-file handling (case) and initial conditions (ic).
"""
from .rhs import (EXPORTED, MfcError, Params, PREC_FP32_BULK, PREC_FP32_LEARNED, PREC_FP64, PREC_FP64_FAST,
                  RhsModule, load_library, populate_grid_buffers, uniform_cell_boundaries)

__all__ = ["EXPORTED", "MfcError", "Params", "PREC_FP32_BULK", "PREC_FP32_LEARNED", "PREC_FP64", "PREC_FP64_FAST",
           "RhsModule",
           "load_library", "populate_grid_buffers", "uniform_cell_boundaries"]
