// obj_loader.hpp -- Wavefront OBJ meshes voxelized into a LatticeGrid3D.
//
// The interface of include/wake/geometry/obj_loader.hpp:9-56 in the
// reference: parse (`v`, `f` with fan triangulation, 1-based or negative
// indices), transform on the host, then voxelize_into_grid, which runs on
// the device (surface cells by a triangle/box overlap test, enclosed cells
// by a walk from the outer faces) and leaves the grid's cell types equal to
// the reference CPU voxelizer's.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "wake/core/types.hpp"
#include "wake/grid/lattice_grid_3d.hpp"

namespace wake {

struct Triangle {
  Vec3 v0;
  Vec3 v1;
  Vec3 v2;
};

class ObjLoader {
 public:
  ObjLoader() = default;

  bool load_file(const std::string& filepath);
  bool load_from_string(const std::string& obj_data);

  std::size_t vertex_count() const { return vertices_.size(); }
  std::size_t triangle_count() const { return triangles_.size(); }
  const std::vector<Vec3>& vertices() const { return vertices_; }
  const std::vector<Triangle>& triangles() const { return triangles_; }

  void translate(const Vec3& offset);
  void scale(real_t s);
  void scale(const Vec3& s);
  void rotate_x(real_t angle_deg);
  void rotate_y(real_t angle_deg);
  void rotate_z(real_t angle_deg);

  void get_bounds(Vec3& min_out, Vec3& max_out) const;
  // Uniform scale into [target_min, target_max], centred in the slack.
  void fit_into_box(const Vec3& target_min, const Vec3& target_max);

  // Cells whose unit box overlaps a triangle become Solid; with
  // fill_interior, so do the cells the outer faces cannot reach.
  void voxelize_into_grid(LatticeGrid3D& grid, bool fill_interior = true) const;

 private:
  template <class F>
  void map_points(F&& f);

  std::vector<Vec3> vertices_;
  std::vector<Triangle> triangles_;
};

}  // namespace wake
