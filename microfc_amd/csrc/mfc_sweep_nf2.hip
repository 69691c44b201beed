// Sweep kernels for num_fluids = 2 (see mfc_sweep.hpp).
#include "mfc_launch.hpp"

namespace mfc {
MFC_SWEEP_LAUNCH_IMPL(2)
} // namespace mfc
