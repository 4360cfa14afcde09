// host_api.cpp -- host-side C++ API around the device-resident grid:
// obstacle rasterisers, the NACA airfoil, the OBJ mesh loader, field
// analytics and the Strouhal analyser (include/wake/geometry,
// include/wake/analytics, include/wake/forces/strouhal_analyzer.hpp).
// Analytic shapes are rasterised on the host in FP32 (bit-identical flags);
// meshes are voxelized and analytics computed on the GPU.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "domain.hpp"
#include "wake/analytics/field_analytics.hpp"
#include "wake/core/constants.hpp"
#include "wake/core/error.hpp"
#include "wake/forces/strouhal_analyzer.hpp"
#include "wake/geometry/naca_airfoil.hpp"
#include "wake/geometry/obj_loader.hpp"
#include "wake/geometry/primitives.hpp"
#include "wake/ibm/immersed_boundary.hpp"
#include "wake_hip.h"

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

// ------------------------------------------------------- immersed boundary
// immersed_boundary.cpp:22-128; the two kernels are the ISPC-named exports.
namespace {
void ibm_check(const char* what) {
  if (wake_hip_clear_error() != 0) throw WakeException(std::string(what) + ": " + wake_hip_last_error());
}
}  // namespace

ImmersedBoundarySolver::ImmersedBoundarySolver(int nx, int ny) : nx_(nx), ny_(ny) {}

// immersed_boundary.cpp:25-39: num_points markers on the circle, ds = R dtheta
void ImmersedBoundarySolver::add_circle(real_t cx, real_t cy, real_t radius, int num_points) {
  if (num_points < 3) return;
  const real_t dth = TWO_PI / static_cast<real_t>(num_points);
  for (int i = 0; i < num_points; ++i) {
    const real_t th = static_cast<real_t>(i) * dth;
    LagrangianMarker m;
    m.pos = Vec3(cx + radius * std::cos(th), cy + radius * std::sin(th), 0.0f);
    m.ds = radius * dth;
    markers_.push_back(m);
  }
  sync_buffers();
}

// immersed_boundary.cpp:41-57: one marker per vertex, ds = length of the
// edge to the next vertex (closed polygon)
void ImmersedBoundarySolver::add_polygon(const std::vector<Vec2>& points) {
  if (points.size() < 2) return;
  for (std::size_t i = 0; i < points.size(); ++i) {
    const Vec2 d = points[(i + 1) % points.size()] - points[i];
    LagrangianMarker m;
    m.pos = Vec3(points[i].x, points[i].y, 0.0f);
    m.ds = std::sqrt(d.norm_sq());
    markers_.push_back(m);
  }
  sync_buffers();
}

void ImmersedBoundarySolver::sync_buffers() {
  const std::size_t n = markers_.size();
  for (auto* b : {&marker_x_, &marker_y_, &marker_ux_, &marker_uy_, &marker_fx_, &marker_fy_, &marker_ds_}) b->resize(n);
  for (std::size_t i = 0; i < n; ++i) {
    marker_x_[i] = markers_[i].pos.x;
    marker_y_[i] = markers_[i].pos.y;
    marker_ds_[i] = markers_[i].ds;
  }
}

void ImmersedBoundarySolver::update(const real_t* euler_ux, const real_t* euler_uy, real_t* euler_gx,
                                    real_t* euler_gy, real_t dt) {
  const int n = static_cast<int>(markers_.size());
  if (n == 0) return;
  (void)wake_hip_clear_error();
  ibm_interpolate_velocity_2d(euler_ux, euler_uy, marker_x_.data(), marker_y_.data(), marker_ux_.data(),
                              marker_uy_.data(), nx_, ny_, n);
  ibm_check("ibm_interpolate_velocity_2d");
  const real_t inv_dt = 1.0f / dt;
  for (int i = 0; i < n; ++i) {
    LagrangianMarker& m = markers_[i];
    m.fluid_vel.x = marker_ux_[i];
    m.fluid_vel.y = marker_uy_[i];
    marker_fx_[i] = (m.target_vel.x - marker_ux_[i]) * inv_dt;
    marker_fy_[i] = (m.target_vel.y - marker_uy_[i]) * inv_dt;
    m.force.x = marker_fx_[i] * m.ds;
    m.force.y = marker_fy_[i] * m.ds;
  }
  ibm_spread_force_2d(marker_x_.data(), marker_y_.data(), marker_fx_.data(), marker_fy_.data(), marker_ds_.data(),
                      euler_gx, euler_gy, nx_, ny_, n);
  ibm_check("ibm_spread_force_2d");
}

Vec3 ImmersedBoundarySolver::compute_net_force() const {
  Vec3 f(0.0f, 0.0f, 0.0f);
  for (const LagrangianMarker& m : markers_) {
    f.x -= m.force.x;
    f.y -= m.force.y;
  }
  return f;
}

// ------------------------------------------------------------- OBJ meshes
// obj_loader.cpp:11-19
bool ObjLoader::load_file(const std::string& filepath) {
  std::ifstream in(filepath);
  if (!in) return false;
  std::ostringstream text;
  text << in.rdbuf();
  return load_from_string(text.str());
}

// obj_loader.cpp:21-83: `v x y z` vertices; `f` faces of v, v/vt, v/vt/vn or
// v//vn tokens, 1-based (negative: from the end; 0 is left as 0),
// triangulated as a fan; triangles with an index out of range are dropped.
// A face token that is not an integer throws std::invalid_argument (from
// std::stoi), as the reference's parser does (obj_loader.cpp:50-52).
bool ObjLoader::load_from_string(const std::string& obj_data) {
  vertices_.clear();
  triangles_.clear();
  std::istringstream lines(obj_data);
  std::string line;
  std::vector<int> face;
  while (std::getline(lines, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::string tag;
    ls >> tag;
    if (tag == "v") {
      real_t x, y, z;
      if (ls >> x >> y >> z) vertices_.emplace_back(x, y, z);
    } else if (tag == "f") {
      face.clear();
      std::string tok;
      while (ls >> tok) {
        int idx = std::stoi(tok.substr(0, tok.find('/')));
        if (idx > 0) idx -= 1;
        else if (idx < 0) idx += static_cast<int>(vertices_.size());
        face.push_back(idx);
      }
      const int nv = static_cast<int>(vertices_.size());
      auto ok = [nv](int i) { return i >= 0 && i < nv; };
      for (std::size_t k = 1; k + 1 < face.size(); ++k)
        if (ok(face[0]) && ok(face[k]) && ok(face[k + 1]))
          triangles_.push_back({vertices_[face[0]], vertices_[face[k]], vertices_[face[k + 1]]});
    }
  }
  return !triangles_.empty();
}

template <class F>
void ObjLoader::map_points(F&& f) {
  for (auto& v : vertices_) v = f(v);
  for (auto& t : triangles_) {
    t.v0 = f(t.v0);
    t.v1 = f(t.v1);
    t.v2 = f(t.v2);
  }
}

// obj_loader.cpp:85-162
void ObjLoader::translate(const Vec3& offset) {
  map_points([&](const Vec3& p) { return p + offset; });
}
void ObjLoader::scale(real_t s) { scale(Vec3(s, s, s)); }
void ObjLoader::scale(const Vec3& s) {
  map_points([&](const Vec3& p) { return Vec3(p.x * s.x, p.y * s.y, p.z * s.z); });
}
void ObjLoader::rotate_x(real_t angle_deg) {
  const real_t a = angle_deg * (PI / 180.0f), c = std::cos(a), s = std::sin(a);
  map_points([&](const Vec3& p) { return Vec3(p.x, p.y * c - p.z * s, p.y * s + p.z * c); });
}
void ObjLoader::rotate_y(real_t angle_deg) {
  const real_t a = angle_deg * (PI / 180.0f), c = std::cos(a), s = std::sin(a);
  map_points([&](const Vec3& p) { return Vec3(p.x * c + p.z * s, p.y, -p.x * s + p.z * c); });
}
void ObjLoader::rotate_z(real_t angle_deg) {
  const real_t a = angle_deg * (PI / 180.0f), c = std::cos(a), s = std::sin(a);
  map_points([&](const Vec3& p) { return Vec3(p.x * c - p.y * s, p.x * s + p.y * c, p.z); });
}

// obj_loader.cpp:164-183
void ObjLoader::get_bounds(Vec3& min_out, Vec3& max_out) const {
  if (vertices_.empty()) {
    min_out = max_out = Vec3(0.0f, 0.0f, 0.0f);
    return;
  }
  min_out = max_out = vertices_.front();
  for (const Vec3& v : vertices_) {
    min_out = Vec3(std::min(min_out.x, v.x), std::min(min_out.y, v.y), std::min(min_out.z, v.z));
    max_out = Vec3(std::max(max_out.x, v.x), std::max(max_out.y, v.y), std::max(max_out.z, v.z));
  }
}

// obj_loader.cpp:185-206: a mesh flat along some axis is left unchanged.
void ObjLoader::fit_into_box(const Vec3& target_min, const Vec3& target_max) {
  Vec3 lo, hi;
  get_bounds(lo, hi);
  const Vec3 size = hi - lo, target = target_max - target_min;
  if (size.x < 1e-8f || size.y < 1e-8f || size.z < 1e-8f) return;
  const real_t s = std::min({target.x / size.x, target.y / size.y, target.z / size.z});
  translate(lo * -1.0f);
  scale(s);
  translate(target_min + (target - size * s) * 0.5f);
}

// obj_loader.cpp:308-400 on the device (LatticeGrid3D's domain).
void ObjLoader::voxelize_into_grid(LatticeGrid3D& grid, bool fill_interior) const {
  std::vector<float> tris;
  tris.reserve(9 * triangles_.size());
  for (const Triangle& t : triangles_)
    for (const Vec3* v : {&t.v0, &t.v1, &t.v2}) tris.insert(tris.end(), {v->x, v->y, v->z});
  guarded([&] {
    grid.domain().voxelize_mesh(tris.data(), static_cast<long long>(triangles_.size()), fill_interior);
  });
}

// ------------------------------------------------------------- primitives
// primitives.cpp:7-13
void GeometryPrimitives::rasterize_circle(LatticeGrid2D& grid, real_t cx, real_t cy, real_t radius) {
  grid.set_solid_circle(cx, cy, radius);
}

void GeometryPrimitives::rasterize_rectangle(LatticeGrid2D& grid, real_t x0, real_t y0, real_t x1, real_t y1) {
  grid.set_solid_box(static_cast<int>(x0), static_cast<int>(y0), static_cast<int>(x1), static_cast<int>(y1));
}

// primitives.cpp:15-44: cells whose offset, turned into the plate frame,
// lies within half the length and half the thickness.
void GeometryPrimitives::rasterize_rotated_plate(LatticeGrid2D& grid, real_t cx, real_t cy, real_t length,
                                                 real_t thickness, real_t angle_deg) {
  const real_t rad = angle_deg * (PI / 180.0f);
  const real_t cos_a = std::cos(rad);
  const real_t sin_a = std::sin(rad);
  const real_t half_len = 0.5f * length;
  const real_t half_thick = 0.5f * thickness;
  const int min_x = std::max(0, static_cast<int>(cx - half_len - thickness));
  const int max_x = std::min(grid.nx() - 1, static_cast<int>(cx + half_len + thickness));
  const int min_y = std::max(0, static_cast<int>(cy - half_len - thickness));
  const int max_y = std::min(grid.ny() - 1, static_cast<int>(cy + half_len + thickness));
  for (int y = min_y; y <= max_y; ++y)
    for (int x = min_x; x <= max_x; ++x) {
      const real_t dx = static_cast<real_t>(x) - cx;
      const real_t dy = static_cast<real_t>(y) - cy;
      const real_t lx = dx * cos_a + dy * sin_a;
      const real_t ly = -dx * sin_a + dy * cos_a;
      if (std::abs(lx) <= half_len && std::abs(ly) <= half_thick) grid.set_cell_type(x, y, CellType::Solid);
    }
}

// primitives.cpp:46-53: circles on a square lattice of pitch `spacing`.
void GeometryPrimitives::rasterize_porous_array(LatticeGrid2D& grid, real_t start_x, real_t end_x, real_t spacing,
                                                real_t obstacle_radius) {
  if (!(spacing > 0.0f)) throw InvalidConfigurationException("porous array spacing must be positive");
  for (real_t x = start_x; x <= end_x; x += spacing)
    for (real_t y = spacing; y < static_cast<real_t>(grid.ny()) - spacing; y += spacing)
      rasterize_circle(grid, x, y, obstacle_radius);
}

void GeometryPrimitives::voxelize_sphere(LatticeGrid3D& grid, real_t cx, real_t cy, real_t cz, real_t radius) {
  grid.set_solid_sphere(cx, cy, cz, radius);
}

void GeometryPrimitives::voxelize_cylinder_z(LatticeGrid3D& grid, real_t cx, real_t cy, real_t radius) {
  grid.set_solid_cylinder_z(cx, cy, radius);
}

void GeometryPrimitives::voxelize_box(LatticeGrid3D& grid, real_t x0, real_t y0, real_t z0, real_t x1, real_t y1,
                                      real_t z1) {
  grid.set_solid_box(static_cast<int>(x0), static_cast<int>(y0), static_cast<int>(z0), static_cast<int>(x1),
                     static_cast<int>(y1), static_cast<int>(z1));
}

void GeometryPrimitives::voxelize_ellipsoid(LatticeGrid3D& grid, real_t cx, real_t cy, real_t cz, real_t rx,
                                            real_t ry, real_t rz) {
  guarded([&] { grid.domain().set_solid_ellipsoid(cx, cy, cz, rx, ry, rz, nullptr); });
}

void GeometryPrimitives::voxelize_ellipsoid(LatticeGrid3D& grid, real_t cx, real_t cy, real_t cz, real_t rx,
                                            real_t ry, real_t rz, const real_t axes[9]) {
  guarded([&] { grid.domain().set_solid_ellipsoid(cx, cy, cz, rx, ry, rz, axes); });
}

// ------------------------------------------------------------------- NACA
NacaAirfoil::NacaAirfoil(const std::string& code, real_t chord, real_t angle_of_attack_deg, std::size_t num_points) {
  configure(code, chord, angle_of_attack_deg, num_points);
}

// naca_airfoil.cpp:17-41: m = digit 1 / 100, p = digit 2 / 10, t = digits 3-4 / 100.
void NacaAirfoil::configure(const std::string& code, real_t chord, real_t angle_of_attack_deg,
                            std::size_t num_points) {
  if (code.size() != 4 || !std::all_of(code.begin(), code.end(), [](char c) { return c >= '0' && c <= '9'; }))
    throw std::invalid_argument("NACA code must be 4 digits, e.g. 0012");
  code_ = code;
  chord_ = chord;
  aoa_deg_ = angle_of_attack_deg;
  camber_ = static_cast<real_t>(code[0] - '0') / 100.0f;
  camber_pos_ = static_cast<real_t>(code[1] - '0') / 10.0f;
  thickness_ = static_cast<real_t>(std::stoi(code.substr(2, 2))) / 100.0f;
  generate_coordinates(num_points);
}

int NacaAirfoil::digits() const { return std::stoi(code_); }

// naca_airfoil.cpp:43-110: cosine-spaced stations x = (1 - cos(i pi / n)) / 2,
// closed-trailing-edge thickness polynomial, camber line and its slope,
// upper/lower points normal to the camber line, turned by -aoa.
void NacaAirfoil::generate_coordinates(std::size_t num_points) {
  polygon_.clear();
  if (num_points == 0) return;
  const real_t alpha_rad = -aoa_deg_ * (PI / 180.0f);
  const real_t cos_a = std::cos(alpha_rad);
  const real_t sin_a = std::sin(alpha_rad);
  std::vector<AirfoilPoint> upper, lower;
  upper.reserve(num_points + 1);
  lower.reserve(num_points + 1);
  for (std::size_t i = 0; i <= num_points; ++i) {
    const real_t beta = static_cast<real_t>(i) * PI / static_cast<real_t>(num_points);
    const real_t x = 0.5f * (1.0f - std::cos(beta));
    const real_t yt = 5.0f * thickness_ *
                      (0.2969f * std::sqrt(x) - 0.1260f * x - 0.3516f * x * x + 0.2843f * x * x * x -
                       0.1036f * x * x * x * x);
    real_t yc = 0.0f, dyc_dx = 0.0f;
    if (camber_pos_ > 1e-6f) {
      if (x < camber_pos_) {
        yc = (camber_ / (camber_pos_ * camber_pos_)) * (2.0f * camber_pos_ * x - x * x);
        dyc_dx = (2.0f * camber_ / (camber_pos_ * camber_pos_)) * (camber_pos_ - x);
      } else {
        const real_t omp = 1.0f - camber_pos_;
        yc = (camber_ / (omp * omp)) * ((1.0f - 2.0f * camber_pos_) + 2.0f * camber_pos_ * x - x * x);
        dyc_dx = (2.0f * camber_ / (omp * omp)) * (camber_pos_ - x);
      }
    }
    const real_t theta = std::atan(dyc_dx);
    const real_t cos_th = std::cos(theta), sin_th = std::sin(theta);
    const real_t xu = (x - yt * sin_th) * chord_, yu = (yc + yt * cos_th) * chord_;
    const real_t xl = (x + yt * sin_th) * chord_, yl = (yc - yt * cos_th) * chord_;
    upper.push_back({xu * cos_a - yu * sin_a, xu * sin_a + yu * cos_a});
    lower.push_back({xl * cos_a - yl * sin_a, xl * sin_a + yl * cos_a});
  }
  polygon_.reserve(2 * num_points + 1);
  for (std::size_t i = 0; i <= num_points; ++i) polygon_.push_back(upper[i]);
  for (std::size_t i = num_points; i > 0; --i) polygon_.push_back(lower[i]);
}

// naca_airfoil.cpp:113-150 (the same FP32 statements as the grid's NACA
// voxelizer, which rasterize / voxelize_extruded call).
bool NacaAirfoil::contains_point(real_t x, real_t y, real_t lead_x, real_t lead_y) const {
  const real_t dx = x - lead_x;
  const real_t dy = y - lead_y;
  const real_t alpha_rad = -aoa_deg_ * (PI / 180.0f);
  const real_t cos_a = std::cos(-alpha_rad);
  const real_t sin_a = std::sin(-alpha_rad);
  const real_t x_chord = dx * cos_a - dy * sin_a;
  const real_t y_chord = dx * sin_a + dy * cos_a;
  if (x_chord < 0.0f || x_chord > chord_) return false;
  const real_t xc = x_chord / chord_;
  const real_t yt = 5.0f * thickness_ *
                    (0.2969f * std::sqrt(xc) - 0.1260f * xc - 0.3516f * xc * xc + 0.2843f * xc * xc * xc -
                     0.1036f * xc * xc * xc * xc) *
                    chord_;
  real_t yc = 0.0f;
  if (camber_pos_ > 1e-6f) {
    if (xc < camber_pos_) {
      yc = (camber_ / (camber_pos_ * camber_pos_)) * (2.0f * camber_pos_ * xc - xc * xc) * chord_;
    } else {
      const real_t omp = 1.0f - camber_pos_;
      yc = (camber_ / (omp * omp)) * ((1.0f - 2.0f * camber_pos_) + 2.0f * camber_pos_ * xc - xc * xc) * chord_;
    }
  }
  return std::abs(y_chord - yc) <= yt;
}

void NacaAirfoil::rasterize(LatticeGrid2D& grid, real_t lead_x, real_t lead_y) const {
  guarded([&] { grid.domain().set_solid_naca(digits(), chord_, aoa_deg_, lead_x, lead_y, 0, 0); });
}

void NacaAirfoil::voxelize_extruded(LatticeGrid3D& grid, real_t lead_x, real_t lead_y, int z_start, int z_end) const {
  guarded([&] { grid.domain().set_solid_naca(digits(), chord_, aoa_deg_, lead_x, lead_y, z_start, z_end); });
}

// -------------------------------------------------------------- analytics
void FieldAnalytics::compute_vorticity(const LatticeGrid2D& grid, AlignedBuffer<real_t>& vorticity_out) {
  const std::size_t n = grid.total_nodes();
  if (vorticity_out.size() != n) vorticity_out.resize(n);
  guarded([&] { std::memcpy(vorticity_out.data(), grid.domain().analytics(2), n * sizeof(real_t)); });
}

void FieldAnalytics::compute_3d_features(const LatticeGrid3D& grid, AlignedBuffer<real_t>& omega_x_out,
                                         AlignedBuffer<real_t>& omega_y_out, AlignedBuffer<real_t>& omega_z_out,
                                         AlignedBuffer<real_t>& q_criterion_out) {
  const std::size_t n = grid.total_nodes();
  AlignedBuffer<real_t>* outs[4] = {&omega_x_out, &omega_y_out, &omega_z_out, &q_criterion_out};
  guarded([&] {
    for (int k = 0; k < 4; ++k) {
      if (outs[k]->size() != n) outs[k]->resize(n);
      std::memcpy(outs[k]->data(), grid.domain().analytics(k), n * sizeof(real_t));
    }
  });
}

// field_analytics.cpp:57-63
real_t FieldAnalytics::compute_max_vorticity_2d(const AlignedBuffer<real_t>& vorticity) {
  real_t m = 0.0f;
  for (std::size_t i = 0; i < vorticity.size(); ++i) m = std::max(m, std::abs(vorticity[i]));
  return m;
}

// --------------------------------------------------------------- Strouhal
namespace {
// In-place radix-2 decimation-in-time transform (n a power of two):
// bit-reversal reordering, then butterflies whose twiddle is advanced by
// repeated multiplication, in FP32 complex arithmetic.
void fft_radix2(std::vector<std::complex<real_t>>& a) {
  const std::size_t n = a.size();
  for (std::size_t i = 1, j = 0; i < n; ++i) {
    std::size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (std::size_t len = 2; len <= n; len <<= 1) {
    const real_t ang = -TWO_PI / static_cast<real_t>(len);
    const std::complex<real_t> wl(std::cos(ang), std::sin(ang));
    for (std::size_t i = 0; i < n; i += len) {
      std::complex<real_t> w(1.0f, 0.0f);
      for (std::size_t k = 0; k < len / 2; ++k) {
        const std::complex<real_t> u = a[i + k];
        const std::complex<real_t> v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
        w *= wl;
      }
    }
  }
}
}  // namespace

real_t StrouhalAnalyzer::analyze_frequency(const std::vector<real_t>& signal, real_t dt) {
  if (signal.size() < 16 || dt <= 0.0f) return 0.0f;
  real_t sum = 0.0f;
  for (real_t v : signal) sum += v;
  const real_t mean = sum / static_cast<real_t>(signal.size());
  std::size_t n = 1;
  while (n < signal.size()) n <<= 1;
  n = std::min<std::size_t>(n, 4096);
  const std::size_t m = std::min(n, signal.size());
  std::vector<std::complex<real_t>> a(n, {0.0f, 0.0f});
  for (std::size_t i = 0; i < m; ++i) {
    const real_t win = 0.5f * (1.0f - std::cos(TWO_PI * static_cast<real_t>(i) / static_cast<real_t>(m - 1)));
    a[i] = {(signal[i] - mean) * win, 0.0f};
  }
  fft_radix2(a);
  const std::size_t half = n / 2;
  frequencies_.assign(half, 0.0f);
  power_spectrum_.assign(half, 0.0f);
  const real_t df = 1.0f / (static_cast<real_t>(n) * dt);
  std::size_t peak = 1;
  real_t best = 0.0f;
  for (std::size_t k = 1; k < half; ++k) {
    frequencies_[k] = static_cast<real_t>(k) * df;
    const real_t p = std::norm(a[k]);
    power_spectrum_[k] = p;
    if (p > best) {
      best = p;
      peak = k;
    }
  }
  return half > 1 ? frequencies_[peak] : 0.0f;
}

real_t StrouhalAnalyzer::compute_strouhal_number(real_t f_shed, real_t diameter, real_t u_inf) {
  if (u_inf <= 1e-8f) return 0.0f;
  return f_shed * diameter / u_inf;
}

}  // namespace wake
