// aerodynamic_telemetry.hpp -- force-coefficient history of a run.
//
// The interface of include/wake/forces/aerodynamic_telemetry.hpp:8-40 in the
// reference: record() one snapshot per sampled step; means over a trailing
// window; the CL statistic is its standard deviation about the window mean;
// CSV with the columns step,time,fx,fy,fz,cd,cl,cm.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "wake/forces/momentum_exchange.hpp"

namespace wake {

struct TelemetrySnapshot {
  uint64_t step;
  real_t time;
  AerodynamicForces forces;
};

class AerodynamicTelemetry {
 public:
  explicit AerodynamicTelemetry(real_t characteristic_area = 1.0f, real_t chord = 1.0f);

  void record(uint64_t step, real_t time, const AerodynamicForces& forces);

  // statistics over history()[start_step:]
  real_t compute_mean_cd(std::size_t start_step = 0) const;
  real_t compute_mean_cl(std::size_t start_step = 0) const;
  real_t compute_rms_cl(std::size_t start_step = 0) const;

  real_t characteristic_area() const noexcept { return characteristic_area_; }
  real_t chord() const noexcept { return chord_; }
  const std::vector<TelemetrySnapshot>& history() const noexcept { return history_; }
  std::size_t size() const noexcept { return history_.size(); }

  void export_csv(const std::string& filepath) const;
  void clear();

 private:
  real_t characteristic_area_{1.0f};
  real_t chord_{1.0f};
  std::vector<TelemetrySnapshot> history_;
};

}  // namespace wake
