// Sweep kernels for num_fluids = 7, plain FP64 form (see mfc_sweep.hpp).
#include "mfc_launch.hpp"

namespace mfc {
MFC_SWEEP_LAUNCH_F64(7)
} // namespace mfc
