// csv_telemetry.hpp -- aerodynamic and performance histories to CSV / JSON.
//
// The interface of include/wake/export/csv_telemetry.hpp:9-55 in the
// reference, with its file formats: aero CSV columns step,time,fx,fy,fz,cd,
// cl,cm; performance CSV columns step,step_time_ms,mlups,max_speed,
// total_mass,kinetic_energy (fixed, 6 decimals); a JSON summary with the
// mean / max / min drag coefficient, the RMS lift coefficient about zero,
// the mean MLUPS and the given shedding frequency and Strouhal number.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "wake/core/types.hpp"
#include "wake/forces/aerodynamic_telemetry.hpp"

namespace wake {

struct PerformanceMetric {
  uint64_t step{0};
  double step_time_ms{0.0};
  double mlups{0.0};
  real_t max_speed{0.0f};
  real_t total_mass{0.0f};
  real_t kinetic_energy{0.0f};
};

class TelemetryExporter {
 public:
  TelemetryExporter() = default;

  void record_aero(const TelemetrySnapshot& sample);
  void record_performance(const PerformanceMetric& perf);

  bool write_aero_csv(const std::string& filepath) const;
  bool write_performance_csv(const std::string& filepath) const;
  bool write_summary_json(const std::string& filepath, const std::string& simulation_name,
                          real_t strouhal_number = 0.0f, real_t shedding_frequency = 0.0f) const;

  void clear();
  const std::vector<TelemetrySnapshot>& aero_samples() const { return aero_samples_; }
  const std::vector<PerformanceMetric>& perf_samples() const { return perf_samples_; }

 private:
  std::vector<TelemetrySnapshot> aero_samples_;
  std::vector<PerformanceMetric> perf_samples_;
};

}  // namespace wake
