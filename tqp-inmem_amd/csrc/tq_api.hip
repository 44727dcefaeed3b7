// Library housekeeping: version, per-thread error text, launch checks.
#include <stdarg.h>
#include <stdio.h>

#include <atomic>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

#include "tq_common.h"

namespace tq {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(TQ_EHIP, "%s launch: %s", what, hipGetErrorString(e));
  return TQ_OK;
}

// ------------------------------------------------------------------ kernel timing log
// A log of (kernel name, start event, stop event) per timed launch.  Event pairs are created on
// first use and reused after every read; launches beyond kTimingCap records run untimed and
// are counted as dropped.
namespace {
constexpr size_t kTimingCap = 1 << 16;
struct TimedLaunch {
  const char* name;
  hipEvent_t e0, e1;
};
std::atomic<bool> g_timing{false};
std::mutex g_timing_mu;
std::vector<TimedLaunch> g_log;     // first g_used entries are live
size_t g_used = 0;
int64_t g_dropped = 0;
}  // namespace

bool timing_on() { return g_timing.load(std::memory_order_relaxed); }

void timing_slot(const char* name, hipEvent_t* e0, hipEvent_t* e1) {
  std::lock_guard<std::mutex> g(g_timing_mu);
  if (g_used == g_log.size()) {
    if (g_log.size() >= kTimingCap) {
      ++g_dropped;
      return;
    }
    TimedLaunch t{nullptr, nullptr, nullptr};
    // timing-only events: no system-scope fence when stamped (the log is read after a full
    // device synchronisation), so a stamp adds no cache write-back to the kernel it brackets
    if (hipEventCreateWithFlags(&t.e0, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&t.e1, hipEventDisableSystemFence) != hipSuccess) {
      ++g_dropped;
      return;
    }
    g_log.push_back(t);
  }
  TimedLaunch& t = g_log[g_used++];
  t.name = name;
  *e0 = t.e0;
  *e1 = t.e1;
}

}  // namespace tq

extern "C" int tq_version(void) { return 10000; }
extern "C" const char* tq_last_error(void) { return tq::g_last_error.c_str(); }
extern "C" int tq_tile_rows(void) { return tq::kTileRows; }

extern "C" int tq_timing(int32_t on) {
  tq::g_timing.store(on != 0);
  return TQ_OK;
}

extern "C" int64_t tq_timing_count(void) {
  std::lock_guard<std::mutex> g(tq::g_timing_mu);
  return (int64_t)tq::g_used;
}

extern "C" int tq_timing_read(tq_ktime_t* out, int64_t cap, int64_t* n, int64_t* dropped) {
  std::lock_guard<std::mutex> g(tq::g_timing_mu);
  const int64_t m = (int64_t)tq::g_used < cap ? (int64_t)tq::g_used : cap;
  for (int64_t i = 0; i < m; ++i) {
    float ms = 0.0f;
    hipError_t e = hipEventElapsedTime(&ms, tq::g_log[i].e0, tq::g_log[i].e1);
    if (e != hipSuccess)
      return tq::set_error(TQ_EHIP, "timing read of %s: %s", tq::g_log[i].name,
                           hipGetErrorString(e));
    out[i].name = tq::g_log[i].name;
    out[i].ms = ms;
    out[i].pad = 0;
  }
  if (n) *n = m;
  if (dropped) *dropped = tq::g_dropped;
  tq::g_used = 0;
  tq::g_dropped = 0;
  return TQ_OK;
}

namespace tq {

int64_t resident_blocks(const void* kernel, size_t lds, int block) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, size_t, int>, int64_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 1024;
  }
  std::lock_guard<std::mutex> g(mu);
  const auto key = std::make_tuple(dev, kernel, lds, block);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, lds) != hipSuccess ||
      per < 1) {
    (void)hipGetLastError();
    per = 1;  // e.g. more LDS than a workgroup gets before it is allowed: one per CU
  }
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1) {
    (void)hipGetLastError();
    cus = 256;
  }
  return cache[key] = (int64_t)per * cus;
}

}  // namespace tq
