// lattice_grid_2d.hpp -- D2Q9 grid on one MI355X.
//
// Public API of the reference's LatticeGrid2D (include/wake/grid/
// lattice_grid_2d.hpp:14-113); step() streams with walls in y exactly like
// the reference.  set_periodic(true) selects fully periodic streaming, which
// the reference grid cannot express.  Populations stay in HBM; the *_data()
// pointers are host mirrors refreshed on first access after a step.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "wake/core/constants.hpp"
#include "wake/core/types.hpp"

namespace wake_hip {
class Domain;
}

namespace wake {

class LatticeGrid2D {
 public:
  LatticeGrid2D(int nx, int ny, real_t relaxation_time = 0.6f, PopulationStorage storage = PopulationStorage::F32);
  ~LatticeGrid2D();
  LatticeGrid2D(const LatticeGrid2D&) = delete;
  LatticeGrid2D& operator=(const LatticeGrid2D&) = delete;
  LatticeGrid2D(LatticeGrid2D&&) noexcept;
  LatticeGrid2D& operator=(LatticeGrid2D&&) noexcept;

  void initialize(real_t rho_init = RHO_0, real_t ux_init = 0.0f, real_t uy_init = 0.0f);
  void initialize_custom(const std::function<void(int x, int y, real_t& rho, real_t& ux, real_t& uy)>& init_fn);
  void step(CollisionType collision = CollisionType::MRT, real_t gx = 0.0f, real_t gy = 0.0f, bool use_les = false,
            real_t smag_const = DEFAULT_SMAGORINSKY_CONSTANT);
  void update_macroscopic();

  void set_cell_type(int x, int y, CellType type);
  CellType get_cell_type(int x, int y) const;
  void set_solid_box(int x0, int y0, int x1, int y1);
  void set_solid_circle(real_t cx, real_t cy, real_t radius);
  void set_inflow_west(real_t u_in_x, real_t u_in_y = 0.0f);
  void set_outflow_east(real_t rho_out = RHO_0);
  void set_moving_wall_north(real_t u_wall_x);

  int nx() const noexcept { return nx_; }
  int ny() const noexcept { return ny_; }
  std::size_t total_nodes() const noexcept;
  uint64_t current_step() const noexcept;
  real_t relaxation_time() const noexcept { return tau_; }
  void set_relaxation_time(real_t tau) noexcept;
  real_t kinematic_viscosity() const noexcept { return CS2 * (tau_ - 0.5f); }

  const real_t* rho_data() const;
  const real_t* ux_data() const;
  const real_t* uy_data() const;
  const real_t* speed_data() const;
  const uint8_t* cell_type_data() const noexcept;
  const real_t* f_src_data() const;
  const real_t* f_dst_data() const;

  real_t rho(int x, int y) const { return rho_data()[static_cast<std::size_t>(y) * nx_ + x]; }
  real_t ux(int x, int y) const { return ux_data()[static_cast<std::size_t>(y) * nx_ + x]; }
  real_t uy(int x, int y) const { return uy_data()[static_cast<std::size_t>(y) * nx_ + x]; }
  real_t speed(int x, int y) const { return speed_data()[static_cast<std::size_t>(y) * nx_ + x]; }

  real_t compute_total_mass() const;
  real_t compute_total_kinetic_energy() const;
  real_t compute_max_speed() const;

  // ---- beyond the reference API
  void set_periodic(bool periodic);
  void set_inflow_parabolic(real_t u_max);
  void set_convective_outflow_east(real_t u_conv);
  void set_les_coupled(bool coupled);
  void add_noise(real_t amplitude, uint64_t seed);
  // (Fx, Fy, torque_z) on the post-stream state (karman_vortex.cpp:107-114).
  Vec3 momentum_exchange(real_t cx, real_t cy, real_t r_max) const;
  // From the current step on, keep momentum_exchange(cx, cy, r_max) of every
  // step, computed inside the step launch itself (no per-step call).
  void enable_force_history(real_t cx, real_t cy, real_t r_max);
  // Forces of steps *first_step .. current_step() (Fx, Fy, torque_z).
  std::vector<Vec3> force_history(uint64_t* first_step = nullptr) const;
  float run(int steps, CollisionType collision = CollisionType::MRT, real_t gx = 0.0f, real_t gy = 0.0f,
            bool use_les = false, real_t smag_const = DEFAULT_SMAGORINSKY_CONSTANT);
  void check_finite() const;
  wake_hip::Domain& domain() const { return *domain_; }

 private:
  int nx_, ny_;
  real_t tau_;
  std::unique_ptr<wake_hip::Domain> domain_;
};

}  // namespace wake
