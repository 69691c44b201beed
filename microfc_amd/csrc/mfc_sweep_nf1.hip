// Sweep kernels for num_fluids = 1 (see mfc_sweep.hpp).
#include "mfc_launch.hpp"

namespace mfc {
MFC_SWEEP_LAUNCH_IMPL(1)
} // namespace mfc
