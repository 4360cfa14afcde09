// momentum_exchange.hpp -- force coefficients from momentum-exchange forces.
//
// compute_aerodynamic_coefficients keeps the reference's signature and
// normalisation (include/wake/forces/momentum_exchange.hpp:20-37): C = F /
// (1/2 rho u^2 A), C_m = T / (1/2 rho u^2 A c), zero when the dynamic pressure
// or the area is below 1e-8.  The forces themselves come from
// LatticeGrid*::momentum_exchange(), reduced on the GPU in FP64.
#pragma once

#include "wake/core/constants.hpp"
#include "wake/core/types.hpp"

namespace wake {

struct AerodynamicForces {
  real_t fx{0.0f}, fy{0.0f}, fz{0.0f}, torque_z{0.0f};
  real_t cd{0.0f}, cl{0.0f}, cm{0.0f};
};

inline AerodynamicForces compute_aerodynamic_coefficients(real_t fx, real_t fy, real_t fz, real_t torque_z,
                                                          real_t rho_inf, real_t u_inf, real_t characteristic_area,
                                                          real_t chord_length = 1.0f) {
  AerodynamicForces out;
  out.fx = fx;
  out.fy = fy;
  out.fz = fz;
  out.torque_z = torque_z;
  const real_t q = 0.5f * rho_inf * u_inf * u_inf;
  if (q > 1e-8f && characteristic_area > 1e-8f) {
    out.cd = fx / (q * characteristic_area);
    out.cl = fy / (q * characteristic_area);
    out.cm = torque_z / (q * characteristic_area * chord_length);
  }
  return out;
}

}  // namespace wake
