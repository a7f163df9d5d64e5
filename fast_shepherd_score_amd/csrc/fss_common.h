// Host-side helpers shared by the translation units of libfss_hip.so (not exported).
#pragma once
#include <string>

namespace fss {

// Records the calling thread's error text (returned by fss_last_error) and returns `code`.
__attribute__((visibility("hidden"))) int set_error(int code, const char* fmt, ...)
    __attribute__((format(printf, 2, 3)));
// The calling thread's last error text.
__attribute__((visibility("hidden"))) const char* last_error();

}  // namespace fss
