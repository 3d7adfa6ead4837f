// Error strings for calls without a context (thread-local) and a helper to
// format messages.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <string>

namespace lbamd {

inline std::string& global_error() {
  static thread_local std::string msg;
  return msg;
}

inline std::string vformat(const char* fmt, va_list ap) {
  char buf[1024];
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  return std::string(buf);
}

inline int fail_global(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  global_error() = vformat(fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace lbamd
