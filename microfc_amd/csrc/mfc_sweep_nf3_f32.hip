// Sweep kernels for num_fluids = 3, FP32-bulk and learned forms (see
// mfc_sweep.hpp); built with FMA contraction.
#define MFC_FORM f32
#include "mfc_launch.hpp"

namespace mfc {
MFC_SWEEP_LAUNCH_F32(3)
} // namespace mfc
