// Library-wide entry points: version, thread-local last error, launch checks.
#include <stdarg.h>
#include <stdio.h>

#include "dcx_common.hpp"

namespace dcx {

namespace {
thread_local char g_last_error[512] = "";
}

dcx_status set_error(dcx_status code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

dcx_status check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DCX_ERR_HIP, "%s: launch failed: %s", what, hipGetErrorString(e));
  return DCX_OK;
}

}  // namespace dcx

extern "C" int dcx_version(void) { return 100; }
extern "C" const char* dcx_last_error(void) { return dcx::g_last_error; }
