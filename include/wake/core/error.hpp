// error.hpp -- exceptions of the wake C++ API (the reference's
// include/wake/core/error.hpp:8-24 hierarchy).  GPU runtime failures surface
// as WakeException; a step that produces non-finite populations can be
// reported as NumericalDivergenceException by LatticeGrid*::check_finite().
#pragma once

#include <stdexcept>
#include <string>

namespace wake {

class WakeException : public std::runtime_error {
 public:
  explicit WakeException(const std::string& message) : std::runtime_error("wake: " + message) {}
};

class NumericalDivergenceException : public WakeException {
 public:
  explicit NumericalDivergenceException(const std::string& details)
      : WakeException("Numerical divergence detected: " + details) {}
};

class InvalidConfigurationException : public WakeException {
 public:
  explicit InvalidConfigurationException(const std::string& details)
      : WakeException("Invalid configuration: " + details) {}
};

}  // namespace wake
