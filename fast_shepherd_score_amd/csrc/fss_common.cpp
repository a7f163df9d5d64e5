// Host-side helpers shared by the translation units of libfss_hip.so: the per-thread error text
// behind fss_last_error.
#include <stdarg.h>
#include <stdio.h>

#include <string>

#include "fss_common.h"

namespace fss {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

const char* last_error() { return g_last_error.c_str(); }

}  // namespace fss
