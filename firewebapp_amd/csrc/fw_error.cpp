// Error reporting of the C ABI: a thread-local message for fw_last_error().
// Also the process-wide lock between graph recording and the library's
// synchronous legacy-stream work (graph.h).
#include <stdarg.h>
#include <stdio.h>
#include <shared_mutex>
#include "../../include/firewave.h"

namespace fw {
static std::shared_mutex g_record_mu;
void record_lock(bool exclusive) {
  if (exclusive) g_record_mu.lock();
  else g_record_mu.lock_shared();
}
void record_unlock(bool exclusive) {
  if (exclusive) g_record_mu.unlock();
  else g_record_mu.unlock_shared();
}
}  // namespace fw

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
