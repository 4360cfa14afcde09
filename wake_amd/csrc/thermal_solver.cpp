// thermal_solver.cpp -- the reference's D2Q9 temperature solver on the GPU
// (src/thermal/thermal_solver.cpp:49-148), its C++ class and its C ABI.
#include "thermal_d2q9.hpp"

#include "capi_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "lattice.hpp"
#include "wake/core/error.hpp"
#include "wake/thermal/thermal_solver.hpp"
#include "wake_hip.h"

namespace wake_hip {

namespace {
// Resets the runtime's last error as check() does (domain.cpp).
void ck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw HipError(std::string("thermal solver: ") + what + ": " + hipGetErrorString(e));
  }
}
}  // namespace

ThermalD2Q9::ThermalD2Q9(int nx, int ny, float tau_thermal, float t_ref, float beta)
    : nx_(nx), ny_(ny), N_(static_cast<size_t>(nx) * ny), tau_(tau_thermal), t_ref_(t_ref), beta_(beta) {
  if (nx <= 0 || ny <= 0) throw HipError("thermal solver: grid dimensions must be positive");
  try {
    ck(hipMalloc(&g_src_, 9 * N_ * sizeof(float)), "hipMalloc");
    ck(hipMalloc(&g_dst_, 9 * N_ * sizeof(float)), "hipMalloc");
    ck(hipMalloc(&temp_, N_ * sizeof(float)), "hipMalloc");
    ck(hipMalloc(&gy_, N_ * sizeof(float)), "hipMalloc");
    initialize(t_ref);
  } catch (...) {
    release();  // the destructor does not run for a constructor that throws
    throw;
  }
}

ThermalD2Q9::ThermalD2Q9(const ThermalD2Q9& o) : ThermalD2Q9(o.nx_, o.ny_, o.tau_, o.t_ref_, o.beta_) {
  south_ = o.south_;
  north_ = o.north_;
  t_hot_ = o.t_hot_;
  t_cold_ = o.t_cold_;
  ck(hipDeviceSynchronize(), "sync");
  ck(hipMemcpy(g_src_, o.g_src_, 9 * N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
  ck(hipMemcpy(g_dst_, o.g_dst_, 9 * N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
  ck(hipMemcpy(temp_, o.temp_, N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
  ck(hipMemcpy(gy_, o.gy_, N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
}

// Waits are device-wide: the fluid domain whose stream ran the last step
// may already be gone (its stream handle with it).
ThermalD2Q9::~ThermalD2Q9() { release(); }

void ThermalD2Q9::release() noexcept {
  (void)hipDeviceSynchronize();
  for (float** p : {&g_src_, &g_dst_, &temp_, &gy_}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  (void)hipGetLastError();
}

// thermal_solver.cpp:49-60: g = w_q T in both buffers, T = t_init, no force.
void ThermalD2Q9::initialize(float t_init) {
  ck(hipDeviceSynchronize(), "sync");
  std::vector<float> g(9 * N_);
  for (int q = 0; q < 9; ++q) std::fill(g.begin() + q * N_, g.begin() + (q + 1) * N_, D2Q9::w[q] * t_init);
  ck(hipMemcpy(g_src_, g.data(), g.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
  ck(hipMemcpy(g_dst_, g.data(), g.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
  std::vector<float> t(N_, t_init), z(N_, 0.0f);
  ck(hipMemcpy(temp_, t.data(), N_ * sizeof(float), hipMemcpyHostToDevice), "upload");
  ck(hipMemcpy(gy_, z.data(), N_ * sizeof(float), hipMemcpyHostToDevice), "upload");
  h_valid_[0] = h_valid_[1] = false;
}

// thermal_solver.cpp:83-123
void ThermalD2Q9::step(Domain& fluid, float g_gravity) {
  if (fluid.dim() != 2 || fluid.nx() != nx_ || fluid.ny() != ny_)
    throw HipError("thermal solver: the fluid grid must be 2D and of the same size");
  const float* ux = fluid.device_field(1);
  const float* uy = fluid.device_field(2);
  const uint8_t* ct = fluid.device_cell_types();
  stream_ = fluid.stream();
  const int n = static_cast<int>(N_);
  ck(launch_thermal_collide_d2q9(g_src_, g_dst_, ct, ux, uy, temp_, tau_, n, stream_), "collide");
  if (south_) ck(launch_thermal_isothermal_wall(g_dst_, nx_, ny_, 0, t_hot_, stream_), "south wall");
  if (north_) ck(launch_thermal_isothermal_wall(g_dst_, nx_, ny_, ny_ - 1, t_cold_, stream_), "north wall");
  ck(launch_stream_walls_d2q9(g_dst_, g_src_, ct, nx_, ny_, 0.0f, 0.0f, stream_), "stream");
  // thermal_solver.cpp:114-122 clears the force and adds the buoyancy; one
  // kernel writes 0 + force (same bits)
  ck(launch_thermal_buoyancy_2d(temp_, gy_, t_ref_, beta_, g_gravity, n, false, stream_), "buoyancy");
  h_valid_[0] = h_valid_[1] = false;
}

const float* ThermalD2Q9::host(int which) {
  if (!h_valid_[which]) {
    h_[which].resize(N_);
    ck(hipDeviceSynchronize(), "sync");
    ck(hipMemcpy(h_[which].data(), which == 0 ? temp_ : gy_, N_ * sizeof(float), hipMemcpyDeviceToHost), "d2h");
    h_valid_[which] = true;
  }
  return h_[which].data();
}

// thermal_solver.cpp:126-131
double ThermalD2Q9::mean_temperature() {
  const float* t = host(0);
  double s = 0.0;
  for (size_t i = 0; i < N_; ++i) s += t[i];
  return s / static_cast<double>(N_);
}

// thermal_solver.cpp:133-146: bottom-wall gradient over the conduction one
double ThermalD2Q9::nusselt() {
  if (std::fabs(t_hot_ - t_cold_) < 1e-6f || ny_ <= 1) return 1.0;
  const float* t = host(0);
  double grad = 0.0;
  for (int x = 0; x < nx_; ++x) grad += t[x] - t[nx_ + x];
  return (grad / nx_) / ((t_hot_ - t_cold_) / (ny_ - 1));
}

}  // namespace wake_hip

namespace wake {

namespace {
template <class F>
auto guarded(F&& f) -> decltype(f()) {
  try {
    return f();
  } catch (const wake_hip::HipError& e) {
    throw WakeException(e.what());
  }
}
}  // namespace

ThermalSolver2D::ThermalSolver2D(int nx, int ny, real_t tau_thermal, real_t t_ref, real_t beta)
    : impl_(guarded([&] { return std::make_unique<wake_hip::ThermalD2Q9>(nx, ny, tau_thermal, t_ref, beta); })) {}
ThermalSolver2D::~ThermalSolver2D() = default;
ThermalSolver2D::ThermalSolver2D(const ThermalSolver2D& o)
    : impl_(guarded([&] { return std::make_unique<wake_hip::ThermalD2Q9>(*o.impl_); })) {}
ThermalSolver2D& ThermalSolver2D::operator=(const ThermalSolver2D& o) {
  if (this != &o) impl_ = guarded([&] { return std::make_unique<wake_hip::ThermalD2Q9>(*o.impl_); });
  return *this;
}
ThermalSolver2D::ThermalSolver2D(ThermalSolver2D&&) noexcept = default;
ThermalSolver2D& ThermalSolver2D::operator=(ThermalSolver2D&&) noexcept = default;

void ThermalSolver2D::initialize(real_t t_init) {
  guarded([&] { impl_->initialize(t_init); });
}
void ThermalSolver2D::step(LatticeGrid2D& fluid_grid, real_t g_gravity) {
  guarded([&] { impl_->step(fluid_grid.domain(), g_gravity); });
}
void ThermalSolver2D::set_isothermal_wall_south(real_t t_hot) { impl_->set_wall_south(t_hot); }
void ThermalSolver2D::set_isothermal_wall_north(real_t t_cold) { impl_->set_wall_north(t_cold); }
const real_t* ThermalSolver2D::temperature_data() const noexcept {
  try {
    return impl_->host(0);
  } catch (...) {
    return nullptr;
  }
}
const real_t* ThermalSolver2D::buoyancy_data() const noexcept {
  try {
    return impl_->host(1);
  } catch (...) {
    return nullptr;
  }
}
real_t ThermalSolver2D::temperature(int x, int y) const {
  return guarded([&] { return impl_->host(0)[static_cast<size_t>(y) * impl_->nx() + x]; });
}
real_t ThermalSolver2D::compute_mean_temperature() const {
  return guarded([&] { return static_cast<real_t>(impl_->mean_temperature()); });
}
real_t ThermalSolver2D::compute_nusselt_number() const {
  return guarded([&] { return static_cast<real_t>(impl_->nusselt()); });
}

}  // namespace wake

// ------------------------------------------------------------------ C ABI
struct wake_thermal2d {
  wake_hip::ThermalD2Q9 t;
};

namespace {
template <class F>
int tguard(wake_thermal2d* t, F&& f) {
  if (!t) return wake_hip::set_last_error("wake: null thermal solver handle");
  try {
    f(t->t);
    return 0;
  } catch (const std::exception& e) {
    return wake_hip::set_last_error(e.what());
  }
}
}  // namespace

extern "C" {
int wake_thermal2d_create(int nx, int ny, float tau_thermal, float t_ref, float beta, wake_thermal2d** out) {
  if (!out) return wake_hip::set_last_error("wake: null output handle");
  try {
    *out = new wake_thermal2d{wake_hip::ThermalD2Q9(nx, ny, tau_thermal, t_ref, beta)};
    return 0;
  } catch (const std::exception& e) {
    *out = nullptr;
    return wake_hip::set_last_error(e.what());
  }
}
int wake_thermal2d_destroy(wake_thermal2d* t) {
  delete t;
  return 0;
}
int wake_thermal2d_initialize(wake_thermal2d* t, float t_init) {
  return tguard(t, [&](wake_hip::ThermalD2Q9& s) { s.initialize(t_init); });
}
int wake_thermal2d_set_walls(wake_thermal2d* t, int south, float t_hot, int north, float t_cold) {
  return tguard(t, [&](wake_hip::ThermalD2Q9& s) {
    if (south) s.set_wall_south(t_hot);
    if (north) s.set_wall_north(t_cold);
  });
}
int wake_thermal2d_step(wake_thermal2d* t, wake_domain* fluid, float g_gravity) {
  return tguard(t, [&](wake_hip::ThermalD2Q9& s) {
    if (!fluid || !fluid->d) throw wake_hip::HipError("thermal step: null fluid domain");
    s.step(*fluid->d, g_gravity);
  });
}
int wake_thermal2d_read(wake_thermal2d* t, int which, float* out) {
  return tguard(t, [&](wake_hip::ThermalD2Q9& s) {
    if ((which != 0 && which != 1) || !out) throw wake_hip::HipError("thermal read: field 0 or 1 into a buffer");
    std::memcpy(out, s.host(which), static_cast<size_t>(s.nx()) * s.ny() * sizeof(float));
  });
}
int wake_thermal2d_stats(wake_thermal2d* t, double out[2]) {
  return tguard(t, [&](wake_hip::ThermalD2Q9& s) {
    if (!out) throw wake_hip::HipError("thermal stats: null output");
    out[0] = s.mean_temperature();
    out[1] = s.nusselt();
  });
}
}  // extern "C"
