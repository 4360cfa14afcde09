// thermal_d2q9.hpp -- device state of the reference's D2Q9 temperature
// solver (src/thermal/thermal_solver.cpp), driven by a fluid Domain.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "domain.hpp"

namespace wake_hip {

class ThermalD2Q9 {
 public:
  ThermalD2Q9(int nx, int ny, float tau_thermal, float t_ref, float beta);
  ThermalD2Q9(const ThermalD2Q9& o);
  ~ThermalD2Q9();
  ThermalD2Q9& operator=(const ThermalD2Q9&) = delete;

  void initialize(float t_init);
  void set_wall_south(float t_hot) {
    south_ = true;
    t_hot_ = t_hot;
  }
  void set_wall_north(float t_cold) {
    north_ = true;
    t_cold_ = t_cold;
  }
  // thermal_solver.cpp:83-123 with the fluid domain's velocity and cell types
  void step(Domain& fluid, float g_gravity);
  // host mirrors: 0 temperature, 1 buoyancy acceleration
  const float* host(int which);
  double mean_temperature();
  double nusselt();
  int nx() const { return nx_; }
  int ny() const { return ny_; }

 private:
  void release() noexcept;  // wait for the device, free every buffer
  int nx_, ny_;
  size_t N_;
  float tau_, t_ref_, beta_;
  bool south_ = false, north_ = false;
  float t_hot_ = 1.0f, t_cold_ = 0.0f;
  float *g_src_ = nullptr, *g_dst_ = nullptr, *temp_ = nullptr, *gy_ = nullptr;
  hipStream_t stream_ = nullptr;  // the fluid domain's stream, used only inside step()
  std::vector<float> h_[2];
  bool h_valid_[2] = {false, false};
};

}  // namespace wake_hip
