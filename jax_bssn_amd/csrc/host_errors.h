// Error reporting shared by the library's translation units: the message of
// the last failure of the calling thread (bssn_last_error_string) and the
// HIP status check of every launch.
#pragma once

#include <hip/hip_runtime.h>

#include "bssn_amd.h"

namespace bssn {

// Record a printf-style message for the calling thread and return code.
__attribute__((visibility("hidden"))) int fail(int code, const char *fmt, ...);
// Record the calling thread's message verbatim.
__attribute__((visibility("hidden"))) void set_error(const char *msg);

inline int check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) return fail(BSSN_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return BSSN_OK;
}

}  // namespace bssn
