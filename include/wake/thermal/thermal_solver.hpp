// thermal_solver.hpp -- the reference's D2Q9 temperature solver on the GPU.
//
// The interface of include/wake/thermal/thermal_solver.hpp:9-51 in the
// reference.  One step (src/thermal/thermal_solver.cpp:83-123): BGK of the
// temperature populations with the fluid grid's current velocity, the
// isothermal south / north walls (anti-bounce-back on the post-collision
// populations), pull streaming with the fluid grid's cell types, and the
// per-node Boussinesq acceleration beta (T - T_ref) g.  The populations
// stay in HBM; temperature_data() is a host mirror refreshed on first read
// after a step.  The reference calls its streaming kernel with two
// arguments missing (the wall velocity); here it streams with a wall
// velocity of zero.
#pragma once

#include <memory>

#include "wake/core/constants.hpp"
#include "wake/core/types.hpp"
#include "wake/grid/lattice_grid_2d.hpp"

namespace wake_hip {
class ThermalD2Q9;
}

namespace wake {

class ThermalSolver2D {
 public:
  ThermalSolver2D(int nx, int ny, real_t tau_thermal = 0.8f, real_t t_ref = 0.5f, real_t beta = 0.05f);
  ~ThermalSolver2D();
  ThermalSolver2D(const ThermalSolver2D& o);
  ThermalSolver2D& operator=(const ThermalSolver2D& o);
  ThermalSolver2D(ThermalSolver2D&&) noexcept;
  ThermalSolver2D& operator=(ThermalSolver2D&&) noexcept;

  void initialize(real_t t_init = 0.5f);
  void step(LatticeGrid2D& fluid_grid, real_t g_gravity = -0.001f);

  void set_isothermal_wall_south(real_t t_hot);
  void set_isothermal_wall_north(real_t t_cold);

  const real_t* temperature_data() const noexcept;
  real_t temperature(int x, int y) const;
  // per-node Boussinesq acceleration of the last step (host mirror)
  const real_t* buoyancy_data() const noexcept;

  real_t compute_mean_temperature() const;
  real_t compute_nusselt_number() const;

  wake_hip::ThermalD2Q9& engine() const { return *impl_; }

 private:
  std::unique_ptr<wake_hip::ThermalD2Q9> impl_;
};

}  // namespace wake
