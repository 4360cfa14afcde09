// export.cpp -- on-disk formats: VTK XML ImageData of the device fields,
// aerodynamic / performance telemetry as CSV and a JSON summary.
//
// src/export/vtk_exporter.cpp:8-223, src/export/csv_telemetry.cpp:8-127 and
// src/forces/aerodynamic_telemetry.cpp:8-62.  The fields come from the
// domain's host mirrors (one read-back kernel and one copy per field after
// a step); the .vti is written either as the reference's ASCII text or as
// one raw little-endian appended block per array.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "domain.hpp"
#include "wake/core/error.hpp"
#include "wake/export/csv_telemetry.hpp"
#include "wake/export/vtk_exporter.hpp"
#include "wake/forces/aerodynamic_telemetry.hpp"

namespace wake_hip {

namespace {

struct PointArray {
  const char* type;  // "Float32" or "UInt8"
  const char* name;
  int ncomp;
  const void* data;  // ncomp values per point, interleaved
  size_t elem;       // bytes per value
};

void write_ascii(std::ostream& out, const PointArray& a, size_t npoints, int nx) {
  out << "        <DataArray type=\"" << a.type << "\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.ncomp
      << "\" format=\"ascii\">\n";
  for (size_t i = 0; i < npoints; ++i) {
    for (int c = 0; c < a.ncomp; ++c) {
      const size_t k = i * a.ncomp + c;
      if (a.elem == 1) out << static_cast<int>(static_cast<const uint8_t*>(a.data)[k]) << " ";
      else out << static_cast<const float*>(a.data)[k] << " ";
    }
    if ((i + 1) % static_cast<size_t>(nx) == 0) out << "\n";
  }
  out << "        </DataArray>\n";
}

}  // namespace

// vtk_exporter.cpp:8-191.  vort2d: one value per point (2D); qcrit: one per
// point and vort3d: three per point (3D).  Returns false when the file
// cannot be opened or written.
bool write_vti(Domain& d, const std::string& path, bool binary, const float* vort2d, const float* qcrit,
               const float* vort3d) {
  const int nx = d.nx(), ny = d.ny(), nz = d.nz();
  const size_t N = static_cast<size_t>(nx) * ny * nz;
  const bool three = d.dim() == 3;
  const float *ux = d.ux(), *uy = d.uy(), *uz = three ? d.uz() : nullptr;
  std::vector<float> vel(3 * N);
  for (size_t i = 0; i < N; ++i) {
    vel[3 * i] = ux[i];
    vel[3 * i + 1] = uy[i];
    vel[3 * i + 2] = three ? uz[i] : 0.0f;
  }
  std::vector<PointArray> arrays = {{"Float32", "Velocity", 3, vel.data(), 4},
                                    {"Float32", "Speed", 1, d.speed(), 4},
                                    {"Float32", "Density", 1, d.rho(), 4}};
  if (!three && vort2d) arrays.push_back({"Float32", "Vorticity", 1, vort2d, 4});
  if (three && qcrit) arrays.push_back({"Float32", "Q_Criterion", 1, qcrit, 4});
  if (three && vort3d) arrays.push_back({"Float32", "Vorticity", 3, vort3d, 4});
  arrays.push_back({"UInt8", "CellType", 1, d.cell_type(), 1});

  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) return false;
  const std::string ext = "0 " + std::to_string(nx - 1) + " 0 " + std::to_string(ny - 1) + " 0 " +
                          std::to_string(three ? nz - 1 : 0);
  out << "<?xml version=\"1.0\"?>\n";
  out << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
  out << "  <ImageData WholeExtent=\"" << ext << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
  out << "    <Piece Extent=\"" << ext << "\">\n";
  out << "      <PointData Vectors=\"Velocity\" Scalars=\"Speed\">\n";
  if (!binary) {
    for (const PointArray& a : arrays) write_ascii(out, a, N, nx);
  } else {
    uint64_t offset = 0;
    for (const PointArray& a : arrays) {
      out << "        <DataArray type=\"" << a.type << "\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.ncomp
          << "\" format=\"appended\" offset=\"" << offset << "\"/>\n";
      offset += sizeof(uint64_t) + N * a.ncomp * a.elem;
    }
  }
  out << "      </PointData>\n";
  out << "    </Piece>\n";
  out << "  </ImageData>\n";
  if (binary) {
    out << "  <AppendedData encoding=\"raw\">\n   _";
    for (const PointArray& a : arrays) {
      const uint64_t bytes = N * a.ncomp * a.elem;
      out.write(reinterpret_cast<const char*>(&bytes), sizeof(bytes));
      out.write(static_cast<const char*>(a.data), static_cast<std::streamsize>(bytes));
    }
    out << "\n  </AppendedData>\n";
  }
  out << "</VTKFile>\n";
  return static_cast<bool>(out);
}

}  // namespace wake_hip

namespace wake {

bool VtkExporter::export_vti_2d(const std::string& filepath, const LatticeGrid2D& grid, const real_t* vorticity,
                                VtkFormat format) {
  try {
    return wake_hip::write_vti(grid.domain(), filepath, format == VtkFormat::Binary, vorticity, nullptr, nullptr);
  } catch (const wake_hip::HipError& e) {
    throw WakeException(e.what());
  }
}

bool VtkExporter::export_vti_3d(const std::string& filepath, const LatticeGrid3D& grid, const real_t* q_criterion,
                                const Vec3* vorticity, VtkFormat format) {
  static_assert(sizeof(Vec3) == 3 * sizeof(float), "Vec3 is three packed floats");
  try {
    return wake_hip::write_vti(grid.domain(), filepath, format == VtkFormat::Binary, nullptr, q_criterion,
                              reinterpret_cast<const float*>(vorticity));
  } catch (const wake_hip::HipError& e) {
    throw WakeException(e.what());
  }
}

// vtk_exporter.cpp:193-221
bool VtkExporter::write_pvd_manifest(const std::string& manifest_path, const std::string& file_prefix, int start_step,
                                     int end_step, int step_interval, real_t dt) {
  std::ofstream out(manifest_path);
  if (!out.is_open()) return false;
  out << "<?xml version=\"1.0\"?>\n";
  out << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
  out << "  <Collection>\n";
  for (int s = start_step; step_interval > 0 && s <= end_step; s += step_interval) {
    std::ostringstream fn;
    fn << file_prefix << "_" << std::setw(6) << std::setfill('0') << s << ".vti";
    out << "    <DataSet timestep=\"" << static_cast<real_t>(s) * dt << "\" group=\"\" part=\"0\" file=\"" << fn.str()
        << "\"/>\n";
  }
  out << "  </Collection>\n";
  out << "</VTKFile>\n";
  return static_cast<bool>(out);
}

// ------------------------------------------------------ aerodynamic history
// aerodynamic_telemetry.cpp:8-62 (FP32 sums, as there)
AerodynamicTelemetry::AerodynamicTelemetry(real_t characteristic_area, real_t chord)
    : characteristic_area_(characteristic_area), chord_(chord) {}

void AerodynamicTelemetry::record(uint64_t step, real_t time, const AerodynamicForces& forces) {
  history_.push_back({step, time, forces});
}

real_t AerodynamicTelemetry::compute_mean_cd(std::size_t start_step) const {
  if (start_step >= history_.size()) return 0.0f;
  real_t sum = 0.0f;
  for (std::size_t i = start_step; i < history_.size(); ++i) sum += history_[i].forces.cd;
  return sum / static_cast<real_t>(history_.size() - start_step);
}

real_t AerodynamicTelemetry::compute_mean_cl(std::size_t start_step) const {
  if (start_step >= history_.size()) return 0.0f;
  real_t sum = 0.0f;
  for (std::size_t i = start_step; i < history_.size(); ++i) sum += history_[i].forces.cl;
  return sum / static_cast<real_t>(history_.size() - start_step);
}

real_t AerodynamicTelemetry::compute_rms_cl(std::size_t start_step) const {
  if (start_step >= history_.size()) return 0.0f;
  const real_t mean = compute_mean_cl(start_step);
  real_t sq = 0.0f;
  for (std::size_t i = start_step; i < history_.size(); ++i) {
    const real_t d = history_[i].forces.cl - mean;
    sq += d * d;
  }
  return std::sqrt(sq / static_cast<real_t>(history_.size() - start_step));
}

namespace {
void aero_rows(std::ostream& out, const std::vector<TelemetrySnapshot>& rows) {
  out << "step,time,fx,fy,fz,cd,cl,cm\n";
  out << std::fixed << std::setprecision(6);
  for (const TelemetrySnapshot& s : rows)
    out << s.step << "," << s.time << "," << s.forces.fx << "," << s.forces.fy << "," << s.forces.fz << ","
        << s.forces.cd << "," << s.forces.cl << "," << s.forces.cm << "\n";
}
}  // namespace

void AerodynamicTelemetry::export_csv(const std::string& filepath) const {
  std::ofstream out(filepath);
  if (out.is_open()) aero_rows(out, history_);
}

void AerodynamicTelemetry::clear() { history_.clear(); }

// ------------------------------------------------------ telemetry exporter
// csv_telemetry.cpp:8-127
void TelemetryExporter::record_aero(const TelemetrySnapshot& sample) { aero_samples_.push_back(sample); }
void TelemetryExporter::record_performance(const PerformanceMetric& perf) { perf_samples_.push_back(perf); }
void TelemetryExporter::clear() {
  aero_samples_.clear();
  perf_samples_.clear();
}

bool TelemetryExporter::write_aero_csv(const std::string& filepath) const {
  std::ofstream out(filepath);
  if (!out.is_open()) return false;
  aero_rows(out, aero_samples_);
  return static_cast<bool>(out);
}

bool TelemetryExporter::write_performance_csv(const std::string& filepath) const {
  std::ofstream out(filepath);
  if (!out.is_open()) return false;
  out << "step,step_time_ms,mlups,max_speed,total_mass,kinetic_energy\n";
  out << std::fixed << std::setprecision(6);
  for (const PerformanceMetric& p : perf_samples_)
    out << p.step << "," << p.step_time_ms << "," << p.mlups << "," << p.max_speed << "," << p.total_mass << ","
        << p.kinetic_energy << "\n";
  return static_cast<bool>(out);
}

namespace {
// A JSON string literal: quotes, backslashes and control characters escaped.
std::string json_string(const std::string& v) {
  std::string o = "\"";
  for (const unsigned char c : v) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += static_cast<char>(c);
    } else if (c < 0x20) {
      char buf[8];
      std::snprintf(buf, sizeof(buf), "\\u%04x", c);
      o += buf;
    } else {
      o += static_cast<char>(c);
    }
  }
  return o + "\"";
}
// A JSON number, or null for a NaN or an infinity (JSON has neither).
std::string json_number(double v) {
  if (!std::isfinite(v)) return "null";
  std::ostringstream o;
  o << v;
  return o.str();
}
}  // namespace

bool TelemetryExporter::write_summary_json(const std::string& filepath, const std::string& simulation_name,
                                           real_t strouhal_number, real_t shedding_frequency) const {
  std::ofstream out(filepath);
  if (!out.is_open()) return false;
  double mean_cd = 0.0, max_cd = 0.0, min_cd = 0.0, rms_cl = 0.0, avg_mlups = 0.0;
  if (!aero_samples_.empty()) {
    double sum_cd = 0.0, sum_cl2 = 0.0;
    max_cd = -1e9;
    min_cd = 1e9;
    for (const TelemetrySnapshot& s : aero_samples_) {
      sum_cd += s.forces.cd;
      sum_cl2 += s.forces.cl * s.forces.cl;
      max_cd = std::max<double>(max_cd, s.forces.cd);
      min_cd = std::min<double>(min_cd, s.forces.cd);
    }
    mean_cd = sum_cd / static_cast<double>(aero_samples_.size());
    rms_cl = std::sqrt(sum_cl2 / static_cast<double>(aero_samples_.size()));
  }
  if (!perf_samples_.empty()) {
    double sum = 0.0;
    for (const PerformanceMetric& p : perf_samples_) sum += p.mlups;
    avg_mlups = sum / static_cast<double>(perf_samples_.size());
  }
  out << "{\n";
  out << "  \"simulation\": " << json_string(simulation_name) << ",\n";
  out << "  \"total_samples\": " << aero_samples_.size() << ",\n";
  out << "  \"performance\": {\n";
  out << "    \"average_mlups\": " << json_number(avg_mlups) << "\n";
  out << "  },\n";
  out << "  \"aerodynamics\": {\n";
  out << "    \"mean_drag_coefficient\": " << json_number(mean_cd) << ",\n";
  out << "    \"max_drag_coefficient\": " << json_number(max_cd) << ",\n";
  out << "    \"min_drag_coefficient\": " << json_number(min_cd) << ",\n";
  out << "    \"rms_lift_coefficient\": " << json_number(rms_cl) << ",\n";
  out << "    \"shedding_frequency\": " << json_number(shedding_frequency) << ",\n";
  out << "    \"strouhal_number\": " << json_number(strouhal_number) << "\n";
  out << "  }\n";
  out << "}\n";
  return static_cast<bool>(out);
}

}  // namespace wake
