// strouhal_analyzer.hpp -- dominant frequency of a force history (the
// reference's StrouhalAnalyzer, include/wake/forces/strouhal_analyzer.hpp,
// src/forces/strouhal_analyzer.cpp:8-89): fewer than 16 samples or dt <= 0
// give 0; mean removed; padded to the next power of two capped at 4096 (only
// the first min(n, size) samples enter); Hann window; radix-2 FFT; the
// first maximum of |X_k|^2 over k = 1 .. n/2 - 1; St = f D / U (0 when
// U <= 1e-8).  Runs on the host, in FP32 as the reference does.
#pragma once

#include <vector>

#include "wake/core/types.hpp"

namespace wake {

class StrouhalAnalyzer {
 public:
  real_t analyze_frequency(const std::vector<real_t>& signal, real_t dt);
  real_t compute_strouhal_number(real_t f_shed, real_t diameter, real_t u_inf);
  const std::vector<real_t>& frequencies() const noexcept { return frequencies_; }
  const std::vector<real_t>& power_spectrum() const noexcept { return power_spectrum_; }

 private:
  std::vector<real_t> frequencies_;
  std::vector<real_t> power_spectrum_;
};

}  // namespace wake
