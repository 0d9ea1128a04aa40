// Optional HIP-event timing of selected kernels, for the benchmark's roofline
// figures.  While a kernel's timer is on, every stride-th launch of that
// kernel from the calling host thread records an event pair through the
// launch itself (hipExtLaunchKernelGGL), so a pair spans the dispatch's
// execution on its stream and not the queue gap before it.  Timers are per
// host thread: cases stepped concurrently from several threads keep their
// own.
#include <stdlib.h>

#include "fw_common.h"

namespace fw {

namespace {
struct KernelTimer {
  bool on = false;
  int used = 0, launches = 0;
  int stride = 1, seen = 0;
  static const int CAP = 4096;
  hipEvent_t* ev = nullptr;   // 2 * CAP events, created on first use
};
thread_local KernelTimer g_kt[FW_TIMER_KINDS];
// a worker thread launching on behalf of another thread uses that thread's
// timers (timer_adopt)
thread_local KernelTimer* g_kt_use = nullptr;
KernelTimer* timers() { return g_kt_use ? g_kt_use : g_kt; }
}  // namespace

void* timer_share() { return timers(); }
void timer_adopt(void* timers_of_another_thread) {
  g_kt_use = (KernelTimer*)timers_of_another_thread;
}

bool timer_on(int which) {
  return which >= 0 && which < FW_TIMER_KINDS && timers()[which].on;
}

KernelEvents timer_next(int which) {
  KernelEvents e;
  if (which < 0 || which >= FW_TIMER_KINDS) return e;
  KernelTimer& t = timers()[which];
  if (!t.on || t.used >= KernelTimer::CAP) return e;
  if ((t.seen++ % t.stride) != 0) return e;
  e.start = t.ev[2 * t.used];
  e.stop = t.ev[2 * t.used + 1];
  t.used += 1;
  t.launches += 1;
  return e;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_kernel_timer(int which, int stride) {
  FW_REQUIRE(which >= 0 && which < FW_TIMER_KINDS, "kernel_timer: unknown kernel %d", which);
  FW_REQUIRE(stride >= 0, "kernel_timer: negative stride");
  KernelTimer& t = g_kt[which];
  if (stride && !t.ev) {
    t.ev = (hipEvent_t*)calloc(2 * KernelTimer::CAP, sizeof(hipEvent_t));
    FW_REQUIRE(t.ev, "kernel_timer: out of host memory");
    for (int q = 0; q < 2 * KernelTimer::CAP; ++q) FW_CHECK_HIP(hipEventCreate(&t.ev[q]));
  }
  if (stride) {
    t.used = 0;
    t.launches = 0;
    t.seen = 0;
    t.stride = stride;
  }
  t.on = stride != 0;
  return FW_OK;
}

int fw_kernel_timer_read(int which, double* total_ms, int* launches) {
  FW_REQUIRE(which >= 0 && which < FW_TIMER_KINDS, "kernel_timer: unknown kernel %d", which);
  FW_REQUIRE(total_ms && launches, "kernel_timer_read: null argument");
  KernelTimer& t = g_kt[which];
  double tot = 0.0;
  for (int q = 0; q < t.used; ++q) {
    FW_CHECK_HIP(hipEventSynchronize(t.ev[2 * q + 1]));
    float ms = 0.f;
    FW_CHECK_HIP(hipEventElapsedTime(&ms, t.ev[2 * q], t.ev[2 * q + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = t.launches;
  return FW_OK;
}

int fw_spread_timer(int enable) { return fw_kernel_timer(FW_TIMER_SPECIES3, enable); }

int fw_spread_timer_read(double* total_ms, int* launches) {
  return fw_kernel_timer_read(FW_TIMER_SPECIES3, total_ms, launches);
}

}  // extern "C"
