// status.cpp -- the calling thread's last error message (dgpu_last_error), host code
// shared by the HIP and the host-only sources.
#include "status.hpp"

#include <cstdarg>
#include <cstdio>
#include <string>

namespace {
thread_local std::string g_err;
}  // namespace

dgpu_status dgpu::fail(dgpu_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

extern "C" const char* dgpu_last_error(void) { return g_err.c_str(); }
