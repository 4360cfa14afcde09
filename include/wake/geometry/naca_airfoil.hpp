// naca_airfoil.hpp -- NACA 4-digit airfoil (the reference's NacaAirfoil,
// include/wake/geometry/naca_airfoil.hpp, src/geometry/naca_airfoil.cpp):
// cosine-spaced contour polygon, point test in the chord frame, 2D
// rasterisation and spanwise extrusion with the reference's FP32
// arithmetic (bit-identical flags).
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "wake/core/types.hpp"
#include "wake/grid/lattice_grid_2d.hpp"
#include "wake/grid/lattice_grid_3d.hpp"

namespace wake {

struct AirfoilPoint {
  real_t x, y;
};

class NacaAirfoil {
 public:
  explicit NacaAirfoil(const std::string& code = "0012", real_t chord = 100.0f, real_t angle_of_attack_deg = 0.0f,
                       std::size_t num_points = 200);
  void configure(const std::string& code, real_t chord, real_t angle_of_attack_deg, std::size_t num_points = 200);

  // Leading edge -> upper surface -> trailing edge -> lower surface.
  const std::vector<AirfoilPoint>& polygon() const { return polygon_; }
  bool contains_point(real_t x, real_t y, real_t lead_x, real_t lead_y) const;
  void rasterize(LatticeGrid2D& grid, real_t lead_x, real_t lead_y) const;
  void voxelize_extruded(LatticeGrid3D& grid, real_t lead_x, real_t lead_y, int z_start, int z_end) const;

  real_t chord() const noexcept { return chord_; }
  real_t angle_of_attack() const noexcept { return aoa_deg_; }
  real_t max_thickness() const noexcept { return thickness_; }
  real_t max_camber() const noexcept { return camber_; }

 private:
  void generate_coordinates(std::size_t num_points);
  int digits() const;

  std::string code_;
  real_t chord_{100.0f};
  real_t aoa_deg_{0.0f};
  real_t camber_{0.0f};
  real_t camber_pos_{0.0f};
  real_t thickness_{0.12f};
  std::vector<AirfoilPoint> polygon_;
};

}  // namespace wake
