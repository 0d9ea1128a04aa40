// Error reporting of the C ABI: a thread-local message for fw_last_error().
#include <stdarg.h>
#include <stdio.h>
#include "../../include/firewave.h"

static thread_local char g_msg[1024] = "";

int fw_set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* fw_last_error(void) { return g_msg; }
extern "C" int fw_version(void) { return 1; }
