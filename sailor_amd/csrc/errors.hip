// Thread-local error reporting for the C ABI.
#include <stdarg.h>

#include "common.h"

namespace {
thread_local char g_last_error[1024] = "";
}

namespace sailor {

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SAILOR_ERR_HIP, "%s: kernel launch failed: %s", what,
                     hipGetErrorString(e));
  return SAILOR_OK;
}

int device_slot(int* slot) {
  int dev = 0;
  SAILOR_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(SAILOR_ERR_ARG, "device %d: launch setup is recorded for devices 0-%d",
                     dev, kMaxDevices - 1);
  *slot = dev;
  return SAILOR_OK;
}

}  // namespace sailor

extern "C" const char* sailor_last_error(void) { return g_last_error; }

extern "C" int sailor_abi_version(void) { return 1; }
