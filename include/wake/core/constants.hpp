// constants.hpp -- lattice constants of the wake C++ API (values as in the
// reference's include/wake/core/constants.hpp).
#pragma once

#include "wake/core/types.hpp"

namespace wake {

inline constexpr real_t CS = 0.5773502691896258f;
inline constexpr real_t CS2 = 1.0f / 3.0f;
inline constexpr real_t INV_CS2 = 3.0f;
inline constexpr real_t INV_2_CS4 = 4.5f;
inline constexpr real_t INV_2_CS2 = 1.5f;
inline constexpr real_t RHO_0 = 1.0f;
inline constexpr real_t DEFAULT_SMAGORINSKY_CONSTANT = 0.14f;
inline constexpr real_t PI = 3.14159265358979323846f;
inline constexpr real_t TWO_PI = 6.28318530717958647692f;

}  // namespace wake
