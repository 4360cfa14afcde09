// immersed_boundary.hpp -- direct-forcing immersed boundary on a 2D grid.
//
// The interface of include/wake/ibm/immersed_boundary.hpp:9-50 in the
// reference.  update() interpolates the velocity at the markers
// (ibm_interpolate_velocity_2d), sets the restoring force
// (U_target - U_marker) / dt and spreads f ds onto the Eulerian body-force
// arrays (ibm_spread_force_2d); both kernels run on the GPU.  The Eulerian
// arrays are host arrays of nx * ny floats, as in the reference.
#pragma once

#include <cstddef>
#include <vector>

#include "wake/core/memory.hpp"
#include "wake/core/types.hpp"
#include "wake/ibm/lagrangian_marker.hpp"

namespace wake {

class ImmersedBoundarySolver {
 public:
  ImmersedBoundarySolver(int nx, int ny);

  void add_circle(real_t cx, real_t cy, real_t radius, int num_points);
  void add_polygon(const std::vector<Vec2>& points);

  void update(const real_t* euler_ux, const real_t* euler_uy, real_t* euler_gx, real_t* euler_gy, real_t dt = 1.0f);

  // force on the body: minus the sum of the markers' force on the fluid
  Vec3 compute_net_force() const;

  std::size_t num_markers() const noexcept { return markers_.size(); }
  const std::vector<LagrangianMarker>& markers() const noexcept { return markers_; }

 private:
  void sync_buffers();

  int nx_{0};
  int ny_{0};
  std::vector<LagrangianMarker> markers_;
  AlignedBuffer<real_t> marker_x_, marker_y_, marker_ux_, marker_uy_, marker_fx_, marker_fy_, marker_ds_;
};

}  // namespace wake
