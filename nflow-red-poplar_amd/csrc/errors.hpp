// errors.hpp -- per-thread error message behind nflow_last_error().
#pragma once
#include <cstdarg>
#include <cstdio>
#include <string>

namespace nflow {

inline std::string& last_error_slot() {
  thread_local std::string s;
  return s;
}

inline int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
inline int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  last_error_slot() = buf;
  return code;
}

}  // namespace nflow

#define NFLOW_HIP(expr)                                                                                   \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return ::nflow::fail(NFLOW_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
