// Cache of instantiated HIP graphs for the fixed-topology loops of the step.
//
// A loop whose launch sequence is fully determined by a small set of host
// values (its trip counts, flags and the device pointers it touches) is
// recorded once per distinct set by stream capture, instantiated, and
// replayed with one hipGraphLaunch afterwards.  The key is those values as
// 64-bit words; anything a kernel reads that changes between replays must
// come from device memory (the step scalar block), never from a captured
// by-value argument.  The captured sequences contain kernel launches only
// (no memset, copy or host synchronisation).  Recording happens on a private
// non-blocking stream (capture executes nothing, and a capturing stream that
// the legacy default stream synchronises with would make other threads'
// synchronous copies fail); the instantiated graph is then launched on the
// caller's stream, which must not be the legacy default stream.  Least
// recently used entries beyond `capacity` are destroyed.
#pragma once
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "fw_common.h"

namespace fw {

struct GraphCache {
  struct Entry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
    uint64_t used;
  };
  std::vector<Entry> entries;
  uint64_t clock = 0;
  int capacity = 24;
  long long hits = 0, captures = 0;
  hipStream_t cap = nullptr;   // the recording stream, created on first use

  ~GraphCache() {
    clear();
    if (cap) (void)hipStreamDestroy(cap);
  }

  void clear() {
    for (Entry& e : entries) (void)hipGraphExecDestroy(e.exec);
    entries.clear();
  }

  // Launch the graph recorded for key on s, recording it first with
  // record(cs) (which issues the launches to the stream cs it is given and
  // returns an FW_* code).
  template <typename F>
  int run(const std::vector<uint64_t>& key, hipStream_t s, F&& record) {
    ++clock;
    for (Entry& e : entries)
      if (e.key == key) {
        e.used = clock;
        ++hits;
        FW_CHECK_HIP(hipGraphLaunch(e.exec, s));
        return FW_OK;
      }
    hipGraphExec_t ex = nullptr;
    {
      RecordScope lock;   // no legacy-stream copies elsewhere meanwhile
      if (!cap) FW_CHECK_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      FW_CHECK_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
      const int rc = record(cap);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(cap, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ec != hipSuccess || !g)
        return fw_set_error(FW_E_HIP, "graph capture: %s", hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess)
        return fw_set_error(FW_E_HIP, "graph instantiate: %s", hipGetErrorString(ei));
    }
    ++captures;
    if ((int)entries.size() >= capacity) {
      size_t lru = 0;
      for (size_t q = 1; q < entries.size(); ++q)
        if (entries[q].used < entries[lru].used) lru = q;
      (void)hipGraphExecDestroy(entries[lru].exec);
      entries.erase(entries.begin() + (long)lru);
    }
    entries.push_back(Entry{key, ex, clock});
    FW_CHECK_HIP(hipGraphLaunch(ex, s));
    return FW_OK;
  }
};

// Whether the PCG and DOM loops replay their iterations from graphs
// (FIREWAVE_LOOP_GRAPHS=1; read at each solve).  Off by default: at
// 256 x 64 x 48 a replayed PCG iteration (four kernels) or DOM source
// iteration measured slower than its direct launches (2.05 against
// 1.97 ms per step), the launch of a small graph costing more than the
// dispatches it saves.
inline bool loop_graphs_on() {
  const char* e = getenv("FIREWAVE_LOOP_GRAPHS");
  return e && e[0] == '1';
}

// Key words for pointers and doubles.
inline uint64_t kw(const void* p) { return (uint64_t)(uintptr_t)p; }
inline uint64_t kw(double v) {
  uint64_t u;
  memcpy(&u, &v, 8);
  return u;
}

}  // namespace fw
