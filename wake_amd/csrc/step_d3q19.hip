// step_d3q19.hip -- D3Q19 instantiation of the fused step kernels.
#include "step_launch.hpp"

namespace wake_hip {
namespace detail {
WAKE_DEFINE_LATTICE(d3q19, D3Q19)
}  // namespace detail
}  // namespace wake_hip
