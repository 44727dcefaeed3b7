// Library housekeeping: version, per-thread error text, launch checks.
#include <stdarg.h>
#include <stdio.h>

#include "tq_common.h"

namespace tq {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(TQ_EHIP, "%s launch: %s", what, hipGetErrorString(e));
  return TQ_OK;
}

}  // namespace tq

extern "C" int tq_version(void) { return 10000; }
extern "C" const char* tq_last_error(void) { return tq::g_last_error.c_str(); }
extern "C" int tq_tile_rows(void) { return tq::kTileRows; }
