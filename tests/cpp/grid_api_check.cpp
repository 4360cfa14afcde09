// grid_api_check.cpp -- exercises the C++ drop-in API (LatticeGrid2D/3D,
// SimulationRunner, compute_aerodynamic_coefficients) the way the
// reference's own validation tests use it, on the GPU.  Exit status 0 means
// every check passed.
#include <cmath>
#include <cstdio>
#include <memory>

#include "wake/core/error.hpp"
#include "wake/forces/momentum_exchange.hpp"
#include "wake/grid/lattice_grid_2d.hpp"
#include "wake/grid/lattice_grid_3d.hpp"
#include "wake/runner/simulation_runner.hpp"

using namespace wake;

static int failures = 0;
#define EXPECT(cond, ...)                 \
  do {                                    \
    if (!(cond)) {                        \
      std::printf("FAIL: " __VA_ARGS__);  \
      std::printf("\n");                  \
      ++failures;                         \
    }                                     \
  } while (0)

// 3D Taylor-Green energy decay, as tests/validation/test_taylor_green_decay.cpp
static void taylor_green_3d() {
  const int L = 32;
  const real_t tau = 0.8f, nu = CS2 * (tau - 0.5f), u0 = 0.04f, k = TWO_PI / static_cast<real_t>(L);
  LatticeGrid3D grid(L, L, L, tau);
  grid.set_periodic(true);
  grid.initialize_custom([&](int x, int y, int z, real_t& rho, real_t& ux, real_t& uy, real_t& uz) {
    const real_t kx = k * x, ky = k * y, kz = k * z;
    ux = u0 * std::sin(kx) * std::cos(ky) * std::cos(kz);
    uy = -u0 * std::cos(kx) * std::sin(ky) * std::cos(kz);
    uz = 0.0f;
    const real_t p = (RHO_0 * u0 * u0 / 16.0f) * (std::cos(2.0f * kx) + std::cos(2.0f * ky)) * (std::cos(2.0f * kz) + 2.0f);
    rho = RHO_0 + p / CS2;
  });
  const real_t E0 = grid.compute_total_kinetic_energy();
  for (int s = 1; s <= 100; ++s) {
    grid.step(CollisionType::MRT);
    if (s % 20 == 0) {
      const real_t e = grid.compute_total_kinetic_energy() / E0;
      const real_t a = std::exp(-6.0f * k * k * nu * static_cast<real_t>(s));
      EXPECT(std::fabs(e - a) <= 0.025f, "TG3D step %d: E/E0 %g vs %g", s, e, a);
    }
  }
  EXPECT(grid.current_step() == 100, "step counter");
  std::printf("taylor_green_3d: E/E0(100) = %g\n", grid.compute_total_kinetic_energy() / E0);
}

// 2D Poiseuille channel, as tests/validation/test_poiseuille_convergence.cpp
static void poiseuille_2d() {
  const int NX = 8, NY = 33;
  const real_t tau = 0.8f, nu = (1.0f / 3.0f) * (tau - 0.5f), gx = 1e-4f, H = static_cast<real_t>(NY - 2);
  LatticeGrid2D grid(NX, NY, tau);
  for (int x = 0; x < NX; ++x) {
    grid.set_cell_type(x, 0, CellType::Solid);
    grid.set_cell_type(x, NY - 1, CellType::Solid);
  }
  for (int s = 0; s < 7000; ++s) grid.step(CollisionType::TRT, gx, 0.0f);
  real_t num = 0, den = 0;
  for (int y = 1; y < NY - 1; ++y) {
    real_t u = 0;
    for (int x = 0; x < NX; ++x) u += grid.ux(x, y);
    u /= NX;
    const real_t d = static_cast<real_t>(y) - 0.5f;
    const real_t ua = (gx / (2.0f * nu)) * d * (H - d);
    num += (u - ua) * (u - ua);
    den += ua * ua;
  }
  const real_t err = std::sqrt(num / den);
  EXPECT(err <= 0.015f, "Poiseuille rel L2 %g", err);
  std::printf("poiseuille_2d: rel L2 %g\n", err);
}

// Sphere wake forces (sphere_3d_wake.cpp:64-114 call sequence, small grid)
static void sphere_forces() {
  const int nx = 64, ny = 32, nz = 32;
  const real_t R = 6.0f, u = 0.05f, cx = nx * 0.25f, cy = ny * 0.5f, cz = nz * 0.5f;
  LatticeGrid3D grid(nx, ny, nz, 3.0f * (u * 2.0f * R / 200.0f) + 0.5f);
  grid.set_periodic(true);
  grid.initialize(RHO_0, u, 0.0f, 0.0f);
  grid.set_solid_sphere(cx, cy, cz, R);
  grid.build_curved_sphere_links(cx, cy, cz, R);
  EXPECT(!grid.is_periodic(), "sphere clears periodicity");
  for (int s = 0; s < 200; ++s) grid.step(CollisionType::MRT, 0.0f, 0.0f, 0.0f, true, 0.12f);
  const Vec3 F = grid.momentum_exchange(cx, cy, cz, 2.0f * R);
  const AerodynamicForces c = compute_aerodynamic_coefficients(F.x, F.y, F.z, 0.0f, RHO_0, u, PI * R * R, 2.0f * R);
  EXPECT(std::isfinite(c.cd) && c.cd > 0.0f, "C_D %g", c.cd);
  EXPECT(std::fabs(c.cl) < 0.05f * c.cd, "C_L %g vs C_D %g", c.cl, c.cd);
  grid.check_finite();
  std::printf("sphere_forces: C_D %g C_L %g\n", c.cd, c.cl);
}

// Runner: callbacks once per step, MLUPS average, max_steps respected.
class GridScenario : public Scenario {
 public:
  GridScenario() : grid_(32, 16, 0.7f) { meta_.name = "poiseuille"; }
  const ScenarioMetadata& metadata() const noexcept override { return meta_; }
  void setup() override {
    for (int x = 0; x < 32; ++x) {
      grid_.set_cell_type(x, 0, CellType::Solid);
      grid_.set_cell_type(x, 15, CellType::Solid);
    }
  }
  void step() override {
    grid_.step(CollisionType::BGK, 1e-5f, 0.0f);
    mlups_ = 1.0;
  }
  uint64_t current_step() const noexcept override { return grid_.current_step(); }
  real_t current_time() const noexcept override { return static_cast<real_t>(grid_.current_step()); }
  real_t current_mlups() const noexcept override { return static_cast<real_t>(mlups_); }
  void render_terminal(std::ostream&) override {}
  void export_frame(const std::string&, int) override { ++frames; }
  void export_telemetry(const std::string&) override { ++telemetry; }
  bool is_finished() const noexcept override { return false; }
  int frames = 0, telemetry = 0;

 private:
  ScenarioMetadata meta_;
  LatticeGrid2D grid_;
  double mlups_ = 0.0;
};

static void runner() {
  auto sc = std::make_shared<GridScenario>();
  RunnerConfig cfg;
  cfg.max_steps = 30;
  cfg.export_interval = 10;
  cfg.live_terminal_render = false;
  cfg.target_fps = 0.0;
  cfg.export_dir = "/tmp/wake_runner_check";
  SimulationRunner r(sc, cfg);
  int calls = 0;
  r.set_step_callback([&](uint64_t, double) { ++calls; });
  r.run();
  EXPECT(calls == 30, "runner callbacks %d", calls);
  EXPECT(r.current_step() == 30, "runner steps %llu", static_cast<unsigned long long>(r.current_step()));
  EXPECT(std::fabs(r.average_mlups() - 1.0) < 1e-9, "runner mlups");
  EXPECT(sc->frames == 4 && sc->telemetry == 1, "exports %d %d", sc->frames, sc->telemetry);
  std::printf("runner: %d callbacks\n", calls);
}

int main() {
  try {
    taylor_green_3d();
    poiseuille_2d();
    sphere_forces();
    runner();
  } catch (const WakeException& e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 2;
  }
  std::printf(failures ? "grid_api_check: %d failure(s)\n" : "grid_api_check: all passed\n", failures);
  return failures ? 1 : 0;
}
