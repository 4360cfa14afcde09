// vtk_exporter.hpp -- grid fields to VTK XML ImageData (.vti) files.
//
// The interface of include/wake/export/vtk_exporter.hpp:9-36 in the
// reference.  The point arrays and their order are the reference's:
// Velocity (3 components; uz = 0 in 2D), Speed, Density, then the optional
// Vorticity (2D scalar) or Q_Criterion and Vorticity (3D vector), then
// CellType (UInt8).  The reference writes them as ASCII text; here the
// default is raw little-endian binary in an appended block (UInt64 sizes),
// which ParaView reads the same way and which is exact and many times
// smaller and faster at 256^3.  VtkFormat::Ascii reproduces the reference's
// text layout (one line per x row, default float formatting).
#pragma once

#include <string>

#include "wake/core/types.hpp"
#include "wake/grid/lattice_grid_2d.hpp"
#include "wake/grid/lattice_grid_3d.hpp"

namespace wake {

enum class VtkFormat { Binary, Ascii };

class VtkExporter {
 public:
  static bool export_vti_2d(const std::string& filepath, const LatticeGrid2D& grid, const real_t* vorticity = nullptr,
                            VtkFormat format = VtkFormat::Binary);
  static bool export_vti_3d(const std::string& filepath, const LatticeGrid3D& grid,
                            const real_t* q_criterion = nullptr, const Vec3* vorticity = nullptr,
                            VtkFormat format = VtkFormat::Binary);
  // ParaView .pvd collection of file_prefix_NNNNNN.vti for
  // start_step, start_step + step_interval, ... <= end_step (time = step * dt)
  static bool write_pvd_manifest(const std::string& manifest_path, const std::string& file_prefix, int start_step,
                                 int end_step, int step_interval, real_t dt = 1.0f);
};

}  // namespace wake
