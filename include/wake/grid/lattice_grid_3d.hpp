// lattice_grid_3d.hpp -- D3Q19 (or D3Q27) grid on one MI355X.
//
// Public API of the reference's LatticeGrid3D (include/wake/grid/
// lattice_grid_3d.hpp:21-123): same constructor, initialisation, step(),
// cell-type setters, curved-link builder, accessors and diagnostics, so code
// written against the reference compiles and behaves the same.  The
// populations live in HBM and one fused HIP kernel per step replaces the
// ISPC kernel sequence; the *_data() pointers are host mirrors that are
// refreshed from the device on first access after a step.  Extra methods
// cover what the reference lacks (D3Q27, FP16/BF16 storage, per-axis
// periodicity, 3D open boundaries, LES-coupled collision, device momentum
// exchange, timed multi-step runs).
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

struct CurvedBoundaryLink {
  int fluid_index;
  int q_dir;
  real_t delta;
  Vec3 wall_velocity;
};

class LatticeGrid3D {
 public:
  LatticeGrid3D(int nx, int ny, int nz, real_t relaxation_time = 0.6f, int lattice = 19,
                PopulationStorage storage = PopulationStorage::F32);
  ~LatticeGrid3D();
  LatticeGrid3D(const LatticeGrid3D&) = delete;
  LatticeGrid3D& operator=(const LatticeGrid3D&) = delete;
  LatticeGrid3D(LatticeGrid3D&&) noexcept;
  LatticeGrid3D& operator=(LatticeGrid3D&&) noexcept;

  void initialize(real_t rho_init = RHO_0, real_t ux_init = 0.0f, real_t uy_init = 0.0f, real_t uz_init = 0.0f);
  void initialize_custom(
      const std::function<void(int x, int y, int z, real_t& rho, real_t& ux, real_t& uy, real_t& uz)>& init_fn);

  // One time step: collision, streaming, boundaries (lattice_grid_3d.cpp:301-380).
  void step(CollisionType collision = CollisionType::MRT, real_t gx = 0.0f, real_t gy = 0.0f, real_t gz = 0.0f,
            bool use_les = false, real_t smag_const = DEFAULT_SMAGORINSKY_CONSTANT);
  void update_macroscopic();

  void set_periodic(bool periodic);
  bool is_periodic() const noexcept;
  void set_cell_type(int x, int y, int z, CellType type);
  CellType get_cell_type(int x, int y, int z) const;
  void set_solid_box(int x0, int y0, int z0, int x1, int y1, int z1);
  void set_solid_sphere(real_t cx, real_t cy, real_t cz, real_t radius);
  void set_solid_cylinder_z(real_t cx, real_t cy, real_t radius);
  void set_moving_wall_top(real_t u_wall_x, real_t u_wall_z = 0.0f);
  void build_curved_sphere_links(real_t cx, real_t cy, real_t cz, real_t radius);

  int nx() const noexcept { return nx_; }
  int ny() const noexcept { return ny_; }
  int nz() const noexcept { return nz_; }
  std::size_t total_nodes() const noexcept;
  uint64_t current_step() const noexcept;
  real_t relaxation_time() const noexcept { return tau_; }
  void set_relaxation_time(real_t tau) noexcept;
  real_t kinematic_viscosity() const noexcept { return CS2 * (tau_ - 0.5f); }

  const real_t* rho_data() const;
  const real_t* ux_data() const;
  const real_t* uy_data() const;
  const real_t* uz_data() const;
  const real_t* speed_data() const;
  const uint8_t* cell_type_data() const noexcept;
  const real_t* f_src_data() const;

  real_t rho(int x, int y, int z) const { return rho_data()[idx(x, y, z)]; }
  real_t ux(int x, int y, int z) const { return ux_data()[idx(x, y, z)]; }
  real_t uy(int x, int y, int z) const { return uy_data()[idx(x, y, z)]; }
  real_t uz(int x, int y, int z) const { return uz_data()[idx(x, y, z)]; }
  real_t speed(int x, int y, int z) const { return speed_data()[idx(x, y, z)]; }

  // Fluid-node sums reduced on the device in FP64.
  real_t compute_total_mass() const;
  real_t compute_total_kinetic_energy() const;
  real_t compute_max_speed() const;

  // ---- beyond the reference API
  void set_periodic_axes(bool per_y, bool per_z);
  void set_inflow_west(real_t u_in_x, real_t u_in_y = 0.0f);
  void set_inflow_turbulence(real_t amplitude, uint64_t seed);
  void set_convective_outflow_east(real_t u_conv);
  void set_les_coupled(bool coupled);
  void add_noise(real_t amplitude, uint64_t seed);
  // (Fx, Fy, Fz) on the post-stream state, links within r_max of the centre.
  Vec3 momentum_exchange(real_t cx, real_t cy, real_t cz, real_t r_max) const;
  // From the current step on, keep momentum_exchange(cx, cy, cz, r_max) of
  // every step, computed inside the step launch itself (no per-step call).
  void enable_force_history(real_t cx, real_t cy, real_t cz, real_t r_max);
  // Forces of steps *first_step .. current_step() (3D: Fx, Fy, Fz).
  std::vector<Vec3> force_history(uint64_t* first_step = nullptr) const;
  // `steps` steps; returns the device milliseconds (hip events) of the launches.
  float run(int steps, CollisionType collision = CollisionType::MRT, real_t gx = 0.0f, real_t gy = 0.0f,
            real_t gz = 0.0f, bool use_les = false, real_t smag_const = DEFAULT_SMAGORINSKY_CONSTANT);
  // Throws NumericalDivergenceException if a density or velocity is not finite.
  void check_finite() const;
  wake_hip::Domain& domain() const { return *domain_; }

 private:
  std::size_t idx(int x, int y, int z) const {
    return (static_cast<std::size_t>(z) * ny_ + y) * nx_ + x;
  }
  int nx_, ny_, nz_;
  real_t tau_;
  std::unique_ptr<wake_hip::Domain> domain_;
};

}  // namespace wake
