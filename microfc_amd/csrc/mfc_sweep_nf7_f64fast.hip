// Sweep kernels for num_fluids = 7, contracted FP64 form (see
// mfc_device.hpp and mfc_sweep.hpp); built with FMA contraction.
#define MFC_FORM f64fast
#define MFC_FAST_F64 1
#include "mfc_launch.hpp"

namespace mfc {
MFC_SWEEP_LAUNCH_F64FAST(7)
} // namespace mfc
