// Sweep kernels for num_fluids = 1, FP32-bulk and learned forms (see
// mfc_sweep.hpp); built with FMA contraction.
#define MFC_FORM f32
#include "mfc_launch.hpp"

namespace mfc {
MFC_SWEEP_LAUNCH_F32(1)
} // namespace mfc
