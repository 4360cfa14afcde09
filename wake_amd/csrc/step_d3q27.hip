// step_d3q27.hip -- D3Q27 instantiation of the fused step kernels.
#include "step_launch.hpp"

namespace wake_hip {
namespace detail {
WAKE_DEFINE_LATTICE(d3q27, D3Q27)
}  // namespace detail
}  // namespace wake_hip
