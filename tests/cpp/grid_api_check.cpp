// grid_api_check.cpp -- exercises the C++ drop-in API (LatticeGrid2D/3D,
// SimulationRunner, compute_aerodynamic_coefficients) the way the
// reference's own validation tests use it, on the GPU.  Exit status 0 means
// every check passed.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <memory>
#include <string>
#include <vector>

#include "wake_hip.h"
#include "wake/analytics/field_analytics.hpp"
#include "wake/core/error.hpp"
#include "wake/export/csv_telemetry.hpp"
#include "wake/export/vtk_exporter.hpp"
#include "wake/forces/aerodynamic_telemetry.hpp"
#include "wake/forces/momentum_exchange.hpp"
#include "wake/forces/strouhal_analyzer.hpp"
#include "wake/geometry/naca_airfoil.hpp"
#include "wake/geometry/obj_loader.hpp"
#include "wake/geometry/primitives.hpp"
#include "wake/ibm/immersed_boundary.hpp"
#include "wake/thermal/thermal_solver.hpp"
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

  // The reference's scenarios call the kernel export on the grid's own
  // arrays (sphere_3d_wake.cpp:98-105).  The grid's f_src mirror is reduced
  // from its device copy; a host copy of the same bytes is uploaded.  Both
  // must give the same bits.
  const size_t n19 = static_cast<size_t>(19) * nx * ny * nz;
  float f1[3], f2[3];
  const real_t* mirror = grid.f_src_data();  // the readback happens here, outside the timing
  const std::vector<float> copy(mirror, mirror + n19);
  auto t0 = std::chrono::steady_clock::now();
  d3q19_compute_momentum_exchange_forces(mirror, grid.cell_type_data(), nx, ny, nz, cx, cy, cz, 2.0f * R, &f1[0],
                                         &f1[1], &f1[2]);
  auto t1 = std::chrono::steady_clock::now();
  auto t2 = std::chrono::steady_clock::now();
  d3q19_compute_momentum_exchange_forces(copy.data(), grid.cell_type_data(), nx, ny, nz, cx, cy, cz, 2.0f * R, &f2[0],
                                         &f2[1], &f2[2]);
  auto t3 = std::chrono::steady_clock::now();
  EXPECT(std::memcmp(f1, f2, sizeof(f1)) == 0, "MEM export on the grid mirror (%g %g %g) vs a host copy (%g %g %g)",
         f1[0], f1[1], f1[2], f2[0], f2[1], f2[2]);
  EXPECT(f1[0] > 0.0f, "MEM export drag %g", f1[0]);
  std::printf("MEM export: grid mirror %.2f ms (resident), host copy %.2f ms (uploaded)\n",
              std::chrono::duration<double, std::milli>(t1 - t0).count(),
              std::chrono::duration<double, std::milli>(t3 - t2).count());
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

// 2D primitives and NACA 0012, as tests/unit/test_geometry.cpp:10-101
static void geometry() {
  LatticeGrid2D g2(64, 64, 0.02f);
  g2.initialize();
  GeometryPrimitives::rasterize_circle(g2, 32.0f, 32.0f, 10.0f);
  EXPECT(g2.get_cell_type(32, 32) == CellType::Solid, "circle centre");
  EXPECT(g2.get_cell_type(0, 0) == CellType::Fluid, "circle far corner");
  GeometryPrimitives::rasterize_rectangle(g2, 5.0f, 5.0f, 15.0f, 15.0f);
  EXPECT(g2.get_cell_type(10, 10) == CellType::Solid, "rectangle");
  GeometryPrimitives::rasterize_rotated_plate(g2, 45.0f, 45.0f, 20.0f, 4.0f, 45.0f);
  EXPECT(g2.get_cell_type(45, 45) == CellType::Solid, "rotated plate");

  NacaAirfoil naca("0012", 50.0f, 5.0f);
  EXPECT(naca.max_thickness() == 0.12f && naca.max_camber() == 0.0f, "NACA 0012 parameters");
  EXPECT(!naca.polygon().empty(), "NACA polygon");
  LatticeGrid2D a2(128, 64, 0.02f);
  a2.initialize();
  naca.rasterize(a2, 30.0f, 32.0f);
  EXPECT(a2.get_cell_type(42, 32) == CellType::Solid, "NACA quarter chord");
  EXPECT(a2.get_cell_type(0, 0) == CellType::Fluid, "NACA far corner");
  // the rasterized cells are exactly the ones contains_point accepts
  int mism = 0, solid = 0;
  for (int y = 0; y < 64; ++y)
    for (int x = 0; x < 128; ++x) {
      const bool in = naca.contains_point(static_cast<real_t>(x), static_cast<real_t>(y), 30.0f, 32.0f);
      mism += in != (a2.get_cell_type(x, y) == CellType::Solid);
      solid += in;
    }
  EXPECT(mism == 0 && solid > 50, "NACA raster vs contains_point: %d mismatches, %d solid", mism, solid);
  LatticeGrid3D a3(64, 64, 32, 0.02f);
  a3.initialize();
  naca.voxelize_extruded(a3, 15.0f, 32.0f, 8, 24);
  EXPECT(a3.get_cell_type(25, 32, 16) == CellType::Solid, "extruded wing");
  EXPECT(a3.get_cell_type(25, 32, 0) == CellType::Fluid && a3.get_cell_type(25, 32, 31) == CellType::Fluid,
         "extrusion span");

  // axis-aligned ellipsoid: the FP32 test (x-cx)^2/rx^2 + ... <= 1 node by node
  LatticeGrid3D e3(40, 32, 24, 0.6f);
  e3.initialize();
  GeometryPrimitives::voxelize_ellipsoid(e3, 20.0f, 16.0f, 12.0f, 9.0f, 6.0f, 4.0f);
  int emism = 0;
  for (int z = 0; z < 24; ++z)
    for (int y = 0; y < 32; ++y)
      for (int x = 0; x < 40; ++x) {
        const double dx = (x - 20.0) / 9.0, dy = (y - 16.0) / 6.0, dz = (z - 12.0) / 4.0;
        const double r = dx * dx + dy * dy + dz * dz;
        if (std::fabs(r - 1.0) < 1e-5) continue;  // FP32 ties decide these
        emism += (r <= 1.0) != (e3.get_cell_type(x, y, z) == CellType::Solid);
      }
  EXPECT(emism == 0, "ellipsoid: %d mismatches", emism);
  std::printf("geometry: NACA 0012 %d solid cells\n", solid);
}

// OBJ cube parsed, fitted into [10, 22]^3 and voxelized with its interior,
// as tests/unit/test_geometry.cpp:29-72; the cell types are then exactly
// the cells 10..22 on every axis (floor/ceil cell boxes, touching overlaps).
static void obj_mesh() {
  const std::string cube =
      "v -1.0 -1.0 -1.0\nv  1.0 -1.0 -1.0\nv  1.0  1.0 -1.0\nv -1.0  1.0 -1.0\n"
      "v -1.0 -1.0  1.0\nv  1.0 -1.0  1.0\nv  1.0  1.0  1.0\nv -1.0  1.0  1.0\n"
      "f 1 2 3 4\nf 5 8 7 6\nf 1 5 6 2\nf 2 6 7 3\nf 3 7 8 4\nf 5 1 4 8\n";
  ObjLoader loader;
  EXPECT(loader.load_from_string(cube), "OBJ parse");
  EXPECT(loader.vertex_count() == 8 && loader.triangle_count() == 12, "OBJ counts %zu %zu", loader.vertex_count(),
         loader.triangle_count());
  LatticeGrid3D grid(32, 32, 32, 0.02f);
  grid.initialize();
  loader.fit_into_box(Vec3(10.0f, 10.0f, 10.0f), Vec3(22.0f, 22.0f, 22.0f));
  Vec3 lo, hi;
  loader.get_bounds(lo, hi);
  EXPECT(lo.x == 10.0f && hi.z == 22.0f, "fit_into_box bounds");
  loader.voxelize_into_grid(grid, true);
  EXPECT(grid.get_cell_type(16, 16, 16) == CellType::Solid, "cube centre");
  EXPECT(grid.get_cell_type(2, 2, 2) == CellType::Fluid && grid.get_cell_type(30, 30, 30) == CellType::Fluid,
         "cube far corners");
  int mism = 0;
  auto in = [](int i) { return i >= 10 && i <= 22; };
  for (int z = 0; z < 32; ++z)
    for (int y = 0; y < 32; ++y)
      for (int x = 0; x < 32; ++x)
        mism += (in(x) && in(y) && in(z)) != (grid.get_cell_type(x, y, z) == CellType::Solid);
  EXPECT(mism == 0, "voxelized cube: %d cells off the 10..22 block", mism);
  EXPECT(!grid.is_periodic(), "a Solid write clears the periodic mode");
  std::printf("obj_mesh: 12 triangles, %d mismatches\n", mism);
}

static std::string slurp(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}

// VtkExporter, the .pvd manifest and the telemetry CSV / JSON writers, in
// the formats of src/export/vtk_exporter.cpp and csv_telemetry.cpp.
static void exports() {
  const std::string dir = "/tmp/wake_export_check";
  std::filesystem::create_directories(dir);
  const int nx = 24, ny = 16, nz = 12;
  const std::size_t N = static_cast<std::size_t>(nx) * ny * nz;
  LatticeGrid3D g(nx, ny, nz, 0.6f);
  g.initialize(1.0f, 0.03f, 0.0f, 0.0f);
  g.set_solid_sphere(8.0f, 8.0f, 6.0f, 3.0f);
  for (int s = 0; s < 5; ++s) g.step(CollisionType::MRT);
  AlignedBuffer<real_t> wx, wy, wz, q;
  FieldAnalytics::compute_3d_features(g, wx, wy, wz, q);
  std::vector<Vec3> vort(N);
  for (std::size_t i = 0; i < N; ++i) vort[i] = Vec3(wx[i], wy[i], wz[i]);
  EXPECT(VtkExporter::export_vti_3d(dir + "/a.vti", g, q.data(), vort.data()), "binary vti");
  const std::string bin = slurp(dir + "/a.vti");
  // arrays: velocity 3, speed, density, Q, vorticity 3 floats, cell type 1 byte
  const std::size_t payload = 6 * 8 + N * (4 * (3 + 1 + 1 + 1 + 3) + 1);
  const std::size_t at = bin.find("<AppendedData encoding=\"raw\">\n   _");
  EXPECT(at != std::string::npos && bin.size() > payload && bin.size() - payload < at + 200, "vti size %zu",
         bin.size());
  float first_speed = 0.0f;  // Speed block follows the Velocity block
  std::memcpy(&first_speed, bin.data() + bin.find('_', at) + 1 + 8 + 12 * N + 8, sizeof(float));
  EXPECT(first_speed == g.speed(0, 0, 0), "binary Speed[0] %g vs %g", first_speed, g.speed(0, 0, 0));
  EXPECT(VtkExporter::export_vti_3d(dir + "/a_ascii.vti", g, nullptr, nullptr, VtkFormat::Ascii), "ascii vti");
  const std::string txt = slurp(dir + "/a_ascii.vti");
  EXPECT(txt.find("<DataArray type=\"UInt8\" Name=\"CellType\" NumberOfComponents=\"1\" format=\"ascii\">") !=
             std::string::npos,
         "ascii CellType array");
  EXPECT(VtkExporter::write_pvd_manifest(dir + "/run.pvd", "frame", 0, 20, 10, 0.5f), "pvd");
  const std::string pvd = slurp(dir + "/run.pvd");
  EXPECT(pvd.find("<DataSet timestep=\"5\" group=\"\" part=\"0\" file=\"frame_000010.vti\"/>") != std::string::npos,
         "pvd entry");

  AerodynamicTelemetry tel(3.0f, 1.0f);
  const real_t cds[3] = {1.0f, 2.0f, 3.0f}, cls[3] = {0.5f, -0.5f, 0.0f};
  TelemetryExporter ex;
  for (int k = 0; k < 3; ++k) {
    AerodynamicForces f;
    f.cd = cds[k];
    f.cl = cls[k];
    f.fx = 0.25f * k;
    tel.record(10 * k, 0.5f * k, f);
    ex.record_aero({static_cast<uint64_t>(10 * k), 0.5f * k, f});
    ex.record_performance({static_cast<uint64_t>(10 * k), 0.4, 30000.0 + k, 0.05f, 1.0f, 0.001f});
  }
  EXPECT(tel.compute_mean_cd() == 2.0f && tel.compute_mean_cd(1) == 2.5f, "mean cd");
  EXPECT(std::fabs(tel.compute_rms_cl() - std::sqrt(0.5f / 3.0f)) < 1e-6f, "cl std %g", tel.compute_rms_cl());
  tel.export_csv(dir + "/aero.csv");
  EXPECT(slurp(dir + "/aero.csv").rfind("step,time,fx,fy,fz,cd,cl,cm\n0,0.000000,0.000000,", 0) == 0, "aero csv");
  EXPECT(ex.write_aero_csv(dir + "/aero2.csv") && slurp(dir + "/aero2.csv") == slurp(dir + "/aero.csv"),
         "exporter aero csv");
  EXPECT(ex.write_performance_csv(dir + "/perf.csv") &&
             slurp(dir + "/perf.csv").find("20,0.400000,30002.000000,0.050000,1.000000,0.001000\n") !=
                 std::string::npos,
         "performance csv");
  EXPECT(ex.write_summary_json(dir + "/summary.json", "sphere", 0.2f, 0.01f), "json");
  const std::string js = slurp(dir + "/summary.json");
  EXPECT(js.find("\"mean_drag_coefficient\": 2,") != std::string::npos &&
             js.find("\"average_mlups\": 30001") != std::string::npos &&
             js.find("\"strouhal_number\": 0.2") != std::string::npos,
         "summary json:\n%s", js.c_str());
  // a name with a quote and a backslash stays one JSON string; a NaN is null
  EXPECT(ex.write_summary_json(dir + "/summary2.json", "run \"a\\b\"\n", std::nanf(""), 0.01f), "json 2");
  const std::string js2 = slurp(dir + "/summary2.json");
  EXPECT(js2.find("\"simulation\": \"run \\\"a\\\\b\\\"\\u000a\",") != std::string::npos &&
             js2.find("\"strouhal_number\": null") != std::string::npos,
         "escaped json:\n%s", js2.c_str());
  EXPECT(!VtkExporter::export_vti_3d("/nonexistent_dir/x.vti", g), "unwritable path reports false");
  std::printf("exports: %zu-byte binary vti, %zu-byte ascii vti\n", bin.size(), txt.size());
}

// ThermalSolver2D driven like thermal_convection.cpp:52-71: hot south wall,
// cold north wall, fluid stepped with a buoyancy from the mean temperature;
// a copy continues exactly like the original.
static void thermal_solver() {
  const int nx = 64, ny = 24;
  LatticeGrid2D fluid(nx, ny, 0.56f);
  fluid.initialize(RHO_0, 0.0f, 0.0f);
  ThermalSolver2D th(nx, ny, 0.53f, 0.5f, 0.05f);
  th.initialize(0.5f);
  th.set_isothermal_wall_south(1.0f);
  th.set_isothermal_wall_north(0.0f);
  for (int s = 0; s < 50; ++s) {
    th.step(fluid, -0.001f);
    fluid.step(CollisionType::TRT, 0.0f, -0.0005f * 0.05f * (th.compute_mean_temperature() - 0.5f));
  }
  ThermalSolver2D copy = th;
  LatticeGrid2D fluid2(nx, ny, 0.56f);
  fluid2.initialize_custom([&](int x, int y, real_t& r, real_t& ux, real_t& uy) {
    r = fluid.rho(x, y);
    ux = fluid.ux(x, y);
    uy = fluid.uy(x, y);
  });
  th.step(fluid, -0.001f);
  copy.step(fluid2, -0.001f);
  EXPECT(std::memcmp(th.temperature_data(), copy.temperature_data(), sizeof(float) * nx * ny) == 0,
         "copied solver diverges");
  const real_t tm = th.compute_mean_temperature(), nu = th.compute_nusselt_number();
  EXPECT(tm > 0.3f && tm < 0.7f && std::isfinite(nu) && nu > 0.0f, "mean T %g, Nu %g", tm, nu);
  EXPECT(th.temperature(nx / 2, 1) > th.temperature(nx / 2, ny - 2), "hot below cold");
  ThermalSolver2D moved = std::move(copy);
  EXPECT(moved.temperature_data() != nullptr && moved.buoyancy_data() != nullptr, "moved solver");
  std::printf("thermal_solver: mean T %g, Nu %g\n", tm, nu);
}

// A cylinder of markers in a uniform stream: the interpolated velocity is
// the stream (partition of unity), the body force is U sum ds, and the
// spread force lands within two cells of the circle.
static void immersed_boundary() {
  const int nx = 64, ny = 48;
  const std::size_t N = static_cast<std::size_t>(nx) * ny;
  std::vector<real_t> ux(N, 0.05f), uy(N, 0.0f), gx(N, 0.0f), gy(N, 0.0f);
  ImmersedBoundarySolver ibm(nx, ny);
  ibm.add_circle(30.0f, 24.0f, 8.0f, 64);
  ibm.add_polygon({Vec2(50.0f, 10.0f), Vec2(56.0f, 10.0f), Vec2(56.0f, 16.0f)});
  EXPECT(ibm.num_markers() == 67, "markers %zu", ibm.num_markers());
  ibm.update(ux.data(), uy.data(), gx.data(), gy.data(), 1.0f);
  real_t worst = 0.0f, sum_ds = 0.0f;
  for (const LagrangianMarker& m : ibm.markers()) {
    worst = std::fmax(worst, std::fabs(m.fluid_vel.x - 0.05f) + std::fabs(m.fluid_vel.y));
    sum_ds += m.ds;
  }
  EXPECT(worst < 1e-6f, "interpolated stream off by %g", worst);
  const Vec3 f = ibm.compute_net_force();
  EXPECT(std::fabs(f.x - 0.05f * sum_ds) < 1e-4f * sum_ds && std::fabs(f.y) < 1e-6f, "net force %g %g", f.x, f.y);
  double far = 0.0, total = 0.0;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < 46; ++x) {
      const double r = std::hypot(x - 30.0, y - 24.0), v = gx[static_cast<std::size_t>(y) * nx + x];
      total += v;
      if (r > 8.0 + 4.0 || r < 8.0 - 4.0) far += std::fabs(v);  // stencil reach: 2 in x, 3 in y
    }
  EXPECT(total < 0.0 && far == 0.0, "spread force total %g, %g outside the band", total, far);
  std::printf("immersed_boundary: net force %g (U sum ds %g)\n", f.x, 0.05f * sum_ds);
}

// rigid rotation curl = 2 omega; pure shear Q = 0 and w_z = -gamma, as
// tests/unit/test_field_analytics.cpp:8-80
static void analytics() {
  const int NX = 16, NY = 16;
  const real_t om = 0.05f;
  LatticeGrid2D g2(NX, NY);
  g2.initialize_custom([&](int x, int y, real_t& rho, real_t& ux, real_t& uy) {
    rho = 1.0f;
    ux = -om * (static_cast<real_t>(y) - 7.5f);
    uy = om * (static_cast<real_t>(x) - 7.5f);
  });
  AlignedBuffer<real_t> w;
  FieldAnalytics::compute_vorticity(g2, w);
  EXPECT(w.size() == static_cast<std::size_t>(NX * NY), "vorticity size");
  real_t worst = 0.0f;
  for (int y = 2; y < NY - 2; ++y)
    for (int x = 2; x < NX - 2; ++x) worst = std::fmax(worst, std::fabs(w[y * NX + x] - 2.0f * om));
  EXPECT(worst <= 1e-5f, "rigid rotation curl off by %g", worst);
  // the maximum includes the periodic seam, where the rotation field jumps
  real_t wmax = 0.0f;
  for (std::size_t i = 0; i < w.size(); ++i) wmax = std::fmax(wmax, std::fabs(w[i]));
  EXPECT(FieldAnalytics::compute_max_vorticity_2d(w) == wmax && wmax >= 2.0f * om, "max vorticity %g vs %g",
         FieldAnalytics::compute_max_vorticity_2d(w), wmax);

  const int L = 16;
  const real_t gam = 0.04f;
  LatticeGrid3D g3(L, L, L);
  g3.initialize_custom([&](int, int y, int, real_t& rho, real_t& ux, real_t& uy, real_t& uz) {
    rho = 1.0f;
    ux = gam * static_cast<real_t>(y);
    uy = uz = 0.0f;
  });
  AlignedBuffer<real_t> wx, wy, wz, q;
  FieldAnalytics::compute_3d_features(g3, wx, wy, wz, q);
  real_t wq = 0.0f, ww = 0.0f;
  for (int z = 2; z < L - 2; ++z)
    for (int y = 2; y < L - 2; ++y)
      for (int x = 2; x < L - 2; ++x) {
        const int i = (z * L + y) * L + x;
        wq = std::fmax(wq, std::fabs(q[i]));
        ww = std::fmax(ww, std::fabs(wz[i] + gam) + std::fabs(wx[i]) + std::fabs(wy[i]));
      }
  EXPECT(wq <= 1e-5f && ww <= 1e-5f, "pure shear: |Q| %g, vorticity off by %g", wq, ww);
  std::printf("analytics: curl error %g, shear Q %g\n", worst, wq);
}

// 0.025 Hz sinusoid plus a 0.1 Hz harmonic over 512 samples, as
// tests/unit/test_forces_momentum_exchange.cpp:70-100
static void strouhal() {
  std::vector<real_t> sig(512);
  for (std::size_t i = 0; i < sig.size(); ++i) {
    const real_t t = static_cast<real_t>(i);
    sig[i] = 0.35f * std::sin(TWO_PI * 0.025f * t) + 0.02f * std::cos(TWO_PI * 0.1f * t);
  }
  StrouhalAnalyzer an;
  const real_t f = an.analyze_frequency(sig, 1.0f);
  EXPECT(std::fabs(f - 0.025f) <= 0.003f, "Strouhal frequency %g", f);
  EXPECT(std::fabs(f - 13.0f / 512.0f) < 1e-7f, "peak bin %g (13/512 expected)", f);
  const real_t st = an.compute_strouhal_number(f, 20.0f, 0.1f);
  EXPECT(std::fabs(st - f * 200.0f) < 1e-5f, "St %g", st);
  EXPECT(an.frequencies().size() == 256 && an.power_spectrum().size() == 256, "spectrum size %zu",
         an.frequencies().size());
  std::printf("strouhal: f = %g, St = %g\n", f, st);
}

// The reference's sphere-wake scenario (sphere_3d_wake.cpp:24-132) on this
// API, driven by SimulationRunner: free stream, sphere with BFL links, MRT
// with the reference's LES switch (Re >= 500, Cs 0.12), the forces of every
// step (here from the per-step force history instead of a momentum-exchange
// call after each step), telemetry every 10 steps; the aero CSV written at
// the end holds every step's forces and coefficients.
class SphereWakeScenario : public Scenario {
 public:
  SphereWakeScenario(int nx, int ny, int nz, real_t re, real_t u, real_t r)
      : nx_(nx), ny_(ny), nz_(nz), re_(re), u_(u), r_(r), grid_(nx, ny, nz, 0.02f) {
    meta_.name = "sphere_3d_turbulent_wake";
    meta_.dimension = SimulationDimension::Dim3D;
    meta_.reynolds_number = re;
    cx_ = static_cast<real_t>(nx) * 0.25f;
    cy_ = static_cast<real_t>(ny) * 0.50f;
    cz_ = static_cast<real_t>(nz) * 0.50f;
    const real_t nu = (u * 2.0f * r) / re;
    tau_ = 3.0f * nu + 0.5f;
    grid_.set_relaxation_time(tau_);
  }
  const ScenarioMetadata& metadata() const noexcept override { return meta_; }
  void setup() override {
    grid_.set_relaxation_time(tau_);
    grid_.set_periodic(true);
    grid_.initialize(RHO_0, u_, 0.0f, 0.0f);
    grid_.set_solid_sphere(cx_, cy_, cz_, r_);
    grid_.build_curved_sphere_links(cx_, cy_, cz_, r_);
    grid_.enable_force_history(cx_, cy_, cz_, 2.0f * r_);
  }
  void step() override {
    const auto t0 = std::chrono::steady_clock::now();
    grid_.step(CollisionType::MRT, 0.0f, 0.0f, 0.0f, re_ >= 500.0f, 0.12f);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms > 1e-6) mlups_ = (static_cast<double>(nx_) * ny_ * nz_ / 1e6) / (ms * 1e-3);
    const uint64_t cur = grid_.current_step();
    if (cur % 10 == 0) {
      uint64_t first = 0;
      const std::vector<Vec3> f = grid_.force_history(&first);
      const real_t area = PI * r_ * r_;
      for (uint64_t k = recorded_; k <= cur; ++k) {
        if (k < first || k - first >= f.size()) continue;
        const Vec3& v = f[k - first];
        const AerodynamicForces c = compute_aerodynamic_coefficients(v.x, v.y, v.z, 0.0f, RHO_0, u_, area, 2.0f * r_);
        telemetry_.record_aero({k, static_cast<real_t>(k), c});
      }
      recorded_ = cur + 1;
      PerformanceMetric perf;
      perf.step = cur;
      perf.step_time_ms = ms;
      perf.mlups = mlups_;
      perf.max_speed = grid_.compute_max_speed();
      perf.total_mass = grid_.compute_total_mass();
      perf.kinetic_energy = grid_.compute_total_kinetic_energy();
      telemetry_.record_performance(perf);
    }
  }
  uint64_t current_step() const noexcept override { return grid_.current_step(); }
  real_t current_time() const noexcept override { return static_cast<real_t>(grid_.current_step()); }
  real_t current_mlups() const noexcept override { return static_cast<real_t>(mlups_); }
  void render_terminal(std::ostream&) override {}
  void export_frame(const std::string&, int) override {}
  void export_telemetry(const std::string& dir) override {
    EXPECT(telemetry_.write_aero_csv(dir + "/sphere_aero.csv"), "aero csv");
    EXPECT(telemetry_.write_performance_csv(dir + "/sphere_perf.csv"), "perf csv");
  }
  bool is_finished() const noexcept override { return false; }

 private:
  int nx_, ny_, nz_;
  real_t re_, u_, r_, tau_, cx_, cy_, cz_;
  LatticeGrid3D grid_;
  ScenarioMetadata meta_;
  TelemetryExporter telemetry_;
  uint64_t recorded_ = 0;
  double mlups_ = 0.0;
};

// grid_api_check sphere_runner <dir> <steps> <re>: the scenario above, 64x32x32, R 6, u 0.05.
static int sphere_runner(const char* dir, int steps, real_t re) {
  auto sc = std::make_shared<SphereWakeScenario>(64, 32, 32, re, 0.05f, 6.0f);
  RunnerConfig cfg;
  cfg.max_steps = static_cast<uint64_t>(steps);
  cfg.export_interval = 0;
  cfg.live_terminal_render = false;
  cfg.target_fps = 0.0;
  cfg.export_dir = dir;
  SimulationRunner r(sc, cfg);
  r.run();
  EXPECT(r.current_step() == static_cast<uint64_t>(steps), "runner steps");
  std::printf(failures ? "sphere_runner: %d failure(s)\n" : "sphere_runner: done\n", failures);
  return failures ? 1 : 0;
}

int main(int argc, char** argv) {
  if (argc == 5 && std::strcmp(argv[1], "sphere_runner") == 0) {
    try {
      return sphere_runner(argv[2], std::atoi(argv[3]), static_cast<real_t>(std::atof(argv[4])));
    } catch (const WakeException& e) {
      std::printf("FAIL: exception %s\n", e.what());
      return 2;
    }
  }
  try {
    geometry();
    obj_mesh();
    analytics();
    exports();
    thermal_solver();
    immersed_boundary();
    strouhal();
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
