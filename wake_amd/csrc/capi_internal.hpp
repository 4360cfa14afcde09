// capi_internal.hpp -- what the C ABI translation units share.
#pragma once

#include <string>

#include "domain.hpp"

struct wake_domain {
  wake_hip::Domain* d;
};

namespace wake_hip {
// Set the message of wake_hip_last_error() for this thread; returns 1.
int set_last_error(const std::string& msg);
}  // namespace wake_hip
