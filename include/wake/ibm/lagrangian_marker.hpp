// lagrangian_marker.hpp -- one immersed-boundary marker
// (include/wake/ibm/lagrangian_marker.hpp:7-14 in the reference).
#pragma once

#include "wake/core/types.hpp"

namespace wake {

struct LagrangianMarker {
  Vec3 pos{0.0f, 0.0f, 0.0f};         // position
  Vec3 target_vel{0.0f, 0.0f, 0.0f};  // prescribed boundary velocity
  Vec3 fluid_vel{0.0f, 0.0f, 0.0f};   // interpolated fluid velocity
  Vec3 force{0.0f, 0.0f, 0.0f};       // restoring force times ds
  real_t ds{1.0f};                    // arc length / area weight
};

}  // namespace wake
