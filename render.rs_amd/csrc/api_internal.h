// Internal helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/rrhip.h"

namespace rrhip {

// Records `msg` as this thread's last error and returns `status`.
int fail(int status, const std::string& msg);
// Maps a HIP error to a status (and records where it happened).
int fail_hip(hipError_t e, const char* where);
// Resolves the current device's copy of the no-allocation staging array
// (once per device; the device must be current).
int noalloc_setup(int device);

}  // namespace rrhip

using rrhip::fail;
using rrhip::fail_hip;

#define RRHIP_HIP_TRY(expr)                              \
  do {                                                   \
    hipError_t rrhip_e_ = (expr);                        \
    if (rrhip_e_ != hipSuccess) return fail_hip(rrhip_e_, #expr); \
  } while (0)
