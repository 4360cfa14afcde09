// step_d2q9.hip -- D2Q9 instantiation of the fused step kernels.
#include "step_launch.hpp"

namespace wake_hip {
namespace detail {
WAKE_DEFINE_LATTICE(d2q9, D2Q9)
}  // namespace detail
}  // namespace wake_hip
