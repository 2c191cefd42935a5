// Host-side helpers shared by every translation unit of libjscc_hip.
#include "jscc_common.h"

#include <stdarg.h>

static thread_local char g_last_error[1024] = "";

void jscc_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

extern "C" const char* jscc_last_error(void) { return g_last_error; }

extern "C" int jscc_abi_version(void) { return 1; }

int jscc_num_cus() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cached[dev] = n;
  }
  return cached[dev];
}

int jscc_stream_cus(hipStream_t stream) {
  const int dev_cus = jscc_num_cus();
  if (stream == nullptr) return dev_cus;
  uint32_t mask[16] = {0};  // up to 512 compute units
  if (hipExtStreamGetCUMask(stream, 16, mask) != hipSuccess) {
    (void)hipGetLastError();  // not a masked stream: the whole device
    return dev_cus;
  }
  int n = 0;
  for (int i = 0; i < 16; ++i) n += __builtin_popcount(mask[i]);
  return (n > 0 && n < dev_cus) ? n : dev_cus;
}

// ---------------------------------------------------------------- kernel timer
#include <vector>

namespace {
struct ProfPair {
  hipEvent_t a, b;
  int tag;
};
thread_local bool g_prof_on = false;
thread_local std::vector<ProfPair> g_prof_pending;
thread_local std::vector<hipEvent_t> g_prof_pool;
thread_local hipEvent_t g_prof_open[JSCC_NUM_TAGS] = {};

hipEvent_t prof_event() {
  if (!g_prof_pool.empty()) {
    hipEvent_t e = g_prof_pool.back();
    g_prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool jscc_prof_on() { return g_prof_on; }

void jscc_prof_begin(int tag, hipStream_t s) {
  if (!g_prof_on || tag < 0 || tag >= JSCC_NUM_TAGS) return;
  hipEvent_t e = prof_event();
  if (e && hipEventRecord(e, s) == hipSuccess) g_prof_open[tag] = e;
}

void jscc_prof_end(int tag, hipStream_t s) {
  if (!g_prof_on || tag < 0 || tag >= JSCC_NUM_TAGS || !g_prof_open[tag]) return;
  hipEvent_t e = prof_event();
  if (e && hipEventRecord(e, s) == hipSuccess) g_prof_pending.push_back({g_prof_open[tag], e, tag});
  g_prof_open[tag] = nullptr;
}

extern "C" int jscc_prof_enable(int on) {
  g_prof_on = on != 0;
  return 0;
}

// Synchronises the recorded pairs of `tag`, returns their count and summed
// milliseconds, and recycles them.
extern "C" int jscc_prof_read(int tag, int* count, double* total_ms) {
  JSCC_REQUIRE(count && total_ms && tag >= 0 && tag < JSCC_NUM_TAGS, "jscc_prof_read: bad arguments");
  int n = 0;
  double ms = 0.0;
  std::vector<ProfPair> keep;
  for (const ProfPair& p : g_prof_pending) {
    if (p.tag != tag) {
      keep.push_back(p);
      continue;
    }
    JSCC_CHECK_HIP(hipEventSynchronize(p.b));
    float t = 0.0f;
    JSCC_CHECK_HIP(hipEventElapsedTime(&t, p.a, p.b));
    ms += t;
    ++n;
    g_prof_pool.push_back(p.a);
    g_prof_pool.push_back(p.b);
  }
  g_prof_pending.swap(keep);
  *count = n;
  *total_ms = ms;
  return 0;
}

extern "C" int jscc_device_cu_count(int* out) {
  if (!out) {
    jscc_set_error("jscc_device_cu_count: null output");
    return -2;
  }
  *out = jscc_num_cus();
  return 0;
}
