// errors.cpp -- per-thread error message of the C ABI.
#include <cstdarg>
#include <cstdio>

#include "sprgpu_internal.h"

namespace {
thread_local char g_msg[1024] = "";
}

int sprg_fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* sprg_last_error(void) { return g_msg; }
