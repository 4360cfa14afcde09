// field_analytics.hpp -- vorticity and Q-criterion of a grid's velocity
// field (the reference's FieldAnalytics, include/wake/analytics/
// field_analytics.hpp, src/analytics/field_analytics.cpp:18-55), computed on
// the GPU from the device-resident fields and copied into the buffers.
#pragma once

#include "wake/core/memory.hpp"
#include "wake/core/types.hpp"
#include "wake/grid/lattice_grid_2d.hpp"
#include "wake/grid/lattice_grid_3d.hpp"

namespace wake {

class FieldAnalytics {
 public:
  static void compute_vorticity(const LatticeGrid2D& grid, AlignedBuffer<real_t>& vorticity_out);
  static void compute_3d_features(const LatticeGrid3D& grid, AlignedBuffer<real_t>& omega_x_out,
                                  AlignedBuffer<real_t>& omega_y_out, AlignedBuffer<real_t>& omega_z_out,
                                  AlignedBuffer<real_t>& q_criterion_out);
  static real_t compute_max_vorticity_2d(const AlignedBuffer<real_t>& vorticity);
};

}  // namespace wake
