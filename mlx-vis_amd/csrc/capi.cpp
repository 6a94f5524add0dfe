// Library-wide C ABI entry points: version, last error, device info, and the
// error/fault plumbing used by every kernel source.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "common.hpp"

namespace mv {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int fault_injected(const char* site) {
  const char* f = getenv("MIVIS_FAULT");
  if (!f || !*f) return 0;
  if (strcmp(f, "all") == 0) return 1;
  return strcmp(f, site) == 0;
}

}  // namespace mv

extern "C" {

int mv_abi_version(void) { return MV_ABI_VERSION; }

const char* mv_last_error(void) { return mv::g_last_error; }

int mv_device_info(int* cu_count, int* lds_per_block, int* wave_size) {
  MV_REQUIRE(cu_count && lds_per_block && wave_size, "mv_device_info: null output pointer");
  int dev = 0;
  MV_TRY_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  MV_TRY_HIP(hipGetDeviceProperties(&prop, dev));
  *cu_count = prop.multiProcessorCount;
  *lds_per_block = static_cast<int>(prop.sharedMemPerBlock);
  *wave_size = prop.warpSize;
  return MV_OK;
}

}  // extern "C"
