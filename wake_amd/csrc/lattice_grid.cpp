// lattice_grid.cpp -- LatticeGrid2D / LatticeGrid3D over wake_hip::Domain.
#include <cmath>
#include <string>
#include <vector>

#include "domain.hpp"
#include "wake/core/error.hpp"
#include "wake/grid/lattice_grid_2d.hpp"
#include "wake/grid/lattice_grid_3d.hpp"

namespace wake {

namespace {
// GPU runtime failures surface as the reference's exception type.
template <class F>
auto guarded(F&& f) -> decltype(f()) {
  try {
    return f();
  } catch (const wake_hip::HipError& e) {
    throw WakeException(e.what());
  }
}

// Counted on the device over the current density and speed (no download).
void finite_or_throw(wake_hip::Domain& d) {
  long long first = -1;
  const long long n = guarded([&] { return d.count_nonfinite(&first); });
  if (n > 0)
    throw NumericalDivergenceException("non-finite density or velocity at node " + std::to_string(first) + " (" +
                                       std::to_string(n) + " nodes)");
}
std::vector<Vec3> to_vec3(const std::vector<double>& rows) {
  std::vector<Vec3> v(rows.size() / 3);
  for (std::size_t i = 0; i < v.size(); ++i)
    v[i] = Vec3(static_cast<real_t>(rows[3 * i]), static_cast<real_t>(rows[3 * i + 1]),
                static_cast<real_t>(rows[3 * i + 2]));
  return v;
}
}  // namespace

// ------------------------------------------------------------------ 3D
LatticeGrid3D::LatticeGrid3D(int nx, int ny, int nz, real_t relaxation_time, int lattice, PopulationStorage storage)
    : nx_(nx), ny_(ny), nz_(nz), tau_(relaxation_time) {
  domain_ = guarded([&] {
    return std::make_unique<wake_hip::Domain>(3, lattice, nx, ny, nz, relaxation_time, static_cast<int>(storage));
  });
}
LatticeGrid3D::~LatticeGrid3D() = default;
LatticeGrid3D::LatticeGrid3D(LatticeGrid3D&&) noexcept = default;
LatticeGrid3D& LatticeGrid3D::operator=(LatticeGrid3D&&) noexcept = default;

void LatticeGrid3D::initialize(real_t rho, real_t ux, real_t uy, real_t uz) {
  guarded([&] { domain_->initialize(rho, ux, uy, uz); });
}

void LatticeGrid3D::initialize_custom(
    const std::function<void(int, int, int, real_t&, real_t&, real_t&, real_t&)>& init_fn) {
  const std::size_t n = total_nodes();
  std::vector<float> r(n), a(n), b(n), c(n);
  for (int z = 0; z < nz_; ++z)
    for (int y = 0; y < ny_; ++y)
      for (int x = 0; x < nx_; ++x) {
        const std::size_t i = idx(x, y, z);
        real_t lr = RHO_0, lx = 0.0f, ly = 0.0f, lz = 0.0f;
        init_fn(x, y, z, lr, lx, ly, lz);
        r[i] = lr;
        a[i] = lx;
        b[i] = ly;
        c[i] = lz;
      }
  guarded([&] { domain_->initialize_fields(r.data(), a.data(), b.data(), c.data()); });
}

void LatticeGrid3D::step(CollisionType collision, real_t gx, real_t gy, real_t gz, bool use_les, real_t smag) {
  guarded([&] { domain_->step(static_cast<int>(collision), gx, gy, gz, use_les, smag); });
}

float LatticeGrid3D::run(int steps, CollisionType collision, real_t gx, real_t gy, real_t gz, bool use_les,
                         real_t smag) {
  return guarded([&] { return domain_->run(steps, static_cast<int>(collision), gx, gy, gz, use_les, smag, true); });
}

void LatticeGrid3D::update_macroscopic() { guarded([&] { domain_->update_macroscopic(); }); }
void LatticeGrid3D::set_periodic(bool p) { guarded([&] { domain_->set_periodic(p); }); }
bool LatticeGrid3D::is_periodic() const noexcept { return domain_->is_periodic(); }
void LatticeGrid3D::set_cell_type(int x, int y, int z, CellType t) {
  guarded([&] { domain_->set_cell_type(x, y, z, static_cast<uint8_t>(t)); });
}
CellType LatticeGrid3D::get_cell_type(int x, int y, int z) const {
  return static_cast<CellType>(domain_->get_cell_type(x, y, z));
}
void LatticeGrid3D::set_solid_box(int x0, int y0, int z0, int x1, int y1, int z1) {
  guarded([&] { domain_->set_solid_box(x0, y0, z0, x1, y1, z1); });
}
void LatticeGrid3D::set_solid_sphere(real_t cx, real_t cy, real_t cz, real_t r) {
  guarded([&] { domain_->set_solid_sphere(cx, cy, cz, r); });
}
void LatticeGrid3D::set_solid_cylinder_z(real_t cx, real_t cy, real_t r) {
  guarded([&] { domain_->set_solid_cylinder_z(cx, cy, r); });
}
void LatticeGrid3D::set_moving_wall_top(real_t uwx, real_t uwz) {
  guarded([&] { domain_->set_moving_wall_top(uwx, uwz); });
}
void LatticeGrid3D::build_curved_sphere_links(real_t cx, real_t cy, real_t cz, real_t r) {
  guarded([&] { domain_->build_curved_sphere_links(cx, cy, cz, r); });
}
std::size_t LatticeGrid3D::total_nodes() const noexcept { return domain_->total_nodes(); }
uint64_t LatticeGrid3D::current_step() const noexcept { return domain_->current_step(); }
void LatticeGrid3D::set_relaxation_time(real_t tau) noexcept {
  tau_ = tau;
  domain_->set_tau(tau);
}
const real_t* LatticeGrid3D::rho_data() const { return guarded([&] { return domain_->rho(); }); }
const real_t* LatticeGrid3D::ux_data() const { return guarded([&] { return domain_->ux(); }); }
const real_t* LatticeGrid3D::uy_data() const { return guarded([&] { return domain_->uy(); }); }
const real_t* LatticeGrid3D::uz_data() const { return guarded([&] { return domain_->uz(); }); }
const real_t* LatticeGrid3D::speed_data() const { return guarded([&] { return domain_->speed(); }); }
const uint8_t* LatticeGrid3D::cell_type_data() const noexcept { return domain_->cell_type(); }
const real_t* LatticeGrid3D::f_src_data() const { return guarded([&] { return domain_->f_src(); }); }
real_t LatticeGrid3D::compute_total_mass() const {
  return static_cast<real_t>(guarded([&] { return domain_->total_mass(); }));
}
real_t LatticeGrid3D::compute_total_kinetic_energy() const {
  return static_cast<real_t>(guarded([&] { return domain_->total_kinetic_energy(); }));
}
real_t LatticeGrid3D::compute_max_speed() const {
  return static_cast<real_t>(guarded([&] { return domain_->max_speed(); }));
}
void LatticeGrid3D::set_periodic_axes(bool py, bool pz) { guarded([&] { domain_->set_periodic_axes(py, pz); }); }
void LatticeGrid3D::set_inflow_west(real_t ux, real_t uy) { guarded([&] { domain_->set_inflow_west(ux, uy); }); }
void LatticeGrid3D::set_inflow_turbulence(real_t a, uint64_t seed) {
  guarded([&] { domain_->set_inflow_turbulence(a, seed); });
}
void LatticeGrid3D::set_convective_outflow_east(real_t u) {
  guarded([&] { domain_->set_convective_outflow_east(u); });
}
void LatticeGrid3D::set_les_coupled(bool c) { domain_->set_les_coupled(c); }
void LatticeGrid3D::add_noise(real_t a, uint64_t seed) { guarded([&] { domain_->add_noise(a, seed); }); }
Vec3 LatticeGrid3D::momentum_exchange(real_t cx, real_t cy, real_t cz, real_t r_max) const {
  double f[3];
  guarded([&] { domain_->momentum_exchange(cx, cy, cz, r_max, f); });
  return Vec3(static_cast<real_t>(f[0]), static_cast<real_t>(f[1]), static_cast<real_t>(f[2]));
}
void LatticeGrid3D::enable_force_history(real_t cx, real_t cy, real_t cz, real_t r_max) {
  guarded([&] { domain_->set_force_probe(cx, cy, cz, r_max); });
}
std::vector<Vec3> LatticeGrid3D::force_history(uint64_t* first_step) const {
  return guarded([&] { return to_vec3(domain_->force_history(first_step)); });
}
void LatticeGrid3D::check_finite() const { finite_or_throw(*domain_); }

// ------------------------------------------------------------------ 2D
LatticeGrid2D::LatticeGrid2D(int nx, int ny, real_t relaxation_time, PopulationStorage storage)
    : nx_(nx), ny_(ny), tau_(relaxation_time) {
  domain_ = guarded([&] {
    return std::make_unique<wake_hip::Domain>(2, 9, nx, ny, 1, relaxation_time, static_cast<int>(storage));
  });
}
LatticeGrid2D::~LatticeGrid2D() = default;
LatticeGrid2D::LatticeGrid2D(LatticeGrid2D&&) noexcept = default;
LatticeGrid2D& LatticeGrid2D::operator=(LatticeGrid2D&&) noexcept = default;

void LatticeGrid2D::initialize(real_t rho, real_t ux, real_t uy) {
  guarded([&] { domain_->initialize(rho, ux, uy, 0.0f); });
}
void LatticeGrid2D::initialize_custom(const std::function<void(int, int, real_t&, real_t&, real_t&)>& init_fn) {
  const std::size_t n = total_nodes();
  std::vector<float> r(n), a(n), b(n);
  for (int y = 0; y < ny_; ++y)
    for (int x = 0; x < nx_; ++x) {
      const std::size_t i = static_cast<std::size_t>(y) * nx_ + x;
      real_t lr = RHO_0, lx = 0.0f, ly = 0.0f;
      init_fn(x, y, lr, lx, ly);
      r[i] = lr;
      a[i] = lx;
      b[i] = ly;
    }
  guarded([&] { domain_->initialize_fields(r.data(), a.data(), b.data(), nullptr); });
}
void LatticeGrid2D::step(CollisionType collision, real_t gx, real_t gy, bool use_les, real_t smag) {
  guarded([&] { domain_->step(static_cast<int>(collision), gx, gy, 0.0f, use_les, smag); });
}
float LatticeGrid2D::run(int steps, CollisionType collision, real_t gx, real_t gy, bool use_les, real_t smag) {
  return guarded([&] { return domain_->run(steps, static_cast<int>(collision), gx, gy, 0.0f, use_les, smag, true); });
}
void LatticeGrid2D::update_macroscopic() { guarded([&] { domain_->update_macroscopic(); }); }
void LatticeGrid2D::set_cell_type(int x, int y, CellType t) {
  guarded([&] { domain_->set_cell_type(x, y, 0, static_cast<uint8_t>(t)); });
}
CellType LatticeGrid2D::get_cell_type(int x, int y) const {
  return static_cast<CellType>(domain_->get_cell_type(x, y, 0));
}
void LatticeGrid2D::set_solid_box(int x0, int y0, int x1, int y1) {
  guarded([&] { domain_->set_solid_box(x0, y0, 0, x1, y1, 0); });
}
void LatticeGrid2D::set_solid_circle(real_t cx, real_t cy, real_t r) {
  guarded([&] { domain_->set_solid_circle(cx, cy, r); });
}
void LatticeGrid2D::set_inflow_west(real_t ux, real_t uy) { guarded([&] { domain_->set_inflow_west(ux, uy); }); }
void LatticeGrid2D::set_outflow_east(real_t rho_out) { guarded([&] { domain_->set_outflow_east(rho_out); }); }
void LatticeGrid2D::set_moving_wall_north(real_t uwx) { guarded([&] { domain_->set_moving_wall_top(uwx, 0.0f); }); }
std::size_t LatticeGrid2D::total_nodes() const noexcept { return domain_->total_nodes(); }
uint64_t LatticeGrid2D::current_step() const noexcept { return domain_->current_step(); }
void LatticeGrid2D::set_relaxation_time(real_t tau) noexcept {
  tau_ = tau;
  domain_->set_tau(tau);
}
const real_t* LatticeGrid2D::rho_data() const { return guarded([&] { return domain_->rho(); }); }
const real_t* LatticeGrid2D::ux_data() const { return guarded([&] { return domain_->ux(); }); }
const real_t* LatticeGrid2D::uy_data() const { return guarded([&] { return domain_->uy(); }); }
const real_t* LatticeGrid2D::speed_data() const { return guarded([&] { return domain_->speed(); }); }
const uint8_t* LatticeGrid2D::cell_type_data() const noexcept { return domain_->cell_type(); }
const real_t* LatticeGrid2D::f_src_data() const { return guarded([&] { return domain_->f_src(); }); }
const real_t* LatticeGrid2D::f_dst_data() const { return guarded([&] { return domain_->f_dst(); }); }
real_t LatticeGrid2D::compute_total_mass() const {
  return static_cast<real_t>(guarded([&] { return domain_->total_mass(); }));
}
real_t LatticeGrid2D::compute_total_kinetic_energy() const {
  return static_cast<real_t>(guarded([&] { return domain_->total_kinetic_energy(); }));
}
real_t LatticeGrid2D::compute_max_speed() const {
  return static_cast<real_t>(guarded([&] { return domain_->max_speed(); }));
}
void LatticeGrid2D::set_periodic(bool p) { guarded([&] { domain_->set_periodic(p); }); }
void LatticeGrid2D::set_inflow_parabolic(real_t u) { guarded([&] { domain_->set_inflow_parabolic(u); }); }
void LatticeGrid2D::set_convective_outflow_east(real_t u) {
  guarded([&] { domain_->set_convective_outflow_east(u); });
}
void LatticeGrid2D::set_les_coupled(bool c) { domain_->set_les_coupled(c); }
void LatticeGrid2D::add_noise(real_t a, uint64_t seed) { guarded([&] { domain_->add_noise(a, seed); }); }
Vec3 LatticeGrid2D::momentum_exchange(real_t cx, real_t cy, real_t r_max) const {
  double f[3];
  guarded([&] { domain_->momentum_exchange(cx, cy, 0.0f, r_max, f); });
  return Vec3(static_cast<real_t>(f[0]), static_cast<real_t>(f[1]), static_cast<real_t>(f[2]));
}
void LatticeGrid2D::enable_force_history(real_t cx, real_t cy, real_t r_max) {
  guarded([&] { domain_->set_force_probe(cx, cy, 0.0f, r_max); });
}
std::vector<Vec3> LatticeGrid2D::force_history(uint64_t* first_step) const {
  return guarded([&] { return to_vec3(domain_->force_history(first_step)); });
}
void LatticeGrid2D::check_finite() const { finite_or_throw(*domain_); }

}  // namespace wake
