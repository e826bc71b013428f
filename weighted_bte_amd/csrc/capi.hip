// Library-wide C ABI: error reporting, version and device-memory helpers.
#include <stdarg.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <string>

#include "common.hpp"
#include "fp.hpp"

namespace bte {

static thread_local char g_last_error[512];

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

// Launch-shape knobs (bte_tune_set): index = knob id, with their allowed ranges.
struct TuneKnob {
  std::atomic<int64_t> value;
  int64_t lo, hi;
};
static TuneKnob g_tune[5] = {{{0}, 0, 0},
                             {{3}, 1, 16},
                             {{32768}, 0, (int64_t)1 << 30},
                             {{8}, 1, 64},
                             {{65536}, 1, (int64_t)1 << 24}};

int64_t tune_get(int knob) { return g_tune[knob].value.load(std::memory_order_relaxed); }

// Test support: the g_fail_at-th allocation from now made by the library's table caches fails as
// if the device were out of memory (bte_test_fail_alloc); 0 means none.
static std::atomic<int> g_fail_at{0};

bool test_alloc_should_fail() {
  int v = g_fail_at.load();
  while (v > 0)
    if (g_fail_at.compare_exchange_weak(v, v - 1)) return v == 1;
  return false;
}

static bool g_profile_on = false;
static std::mutex g_profile_mu;
static std::string g_profile_text;

PhaseTimer::PhaseTimer(hipStream_t s) : stream(s), on(g_profile_on) {
  for (int i = 0; i < 33; ++i) ev[i] = nullptr;
}

PhaseTimer::~PhaseTimer() {
  for (int i = 0; i < 33; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
}

void PhaseTimer::begin() {
  if (!on) return;
  if (hipEventCreate(&ev[0]) != hipSuccess) { on = false; return; }
  (void)hipEventRecord(ev[0], stream);
}

void PhaseTimer::mark(const char* name) {
  if (!on || n >= 32) return;
  if (hipEventCreate(&ev[n + 1]) != hipSuccess) { on = false; return; }
  (void)hipEventRecord(ev[n + 1], stream);
  names[n++] = name;
}

void PhaseTimer::publish() {
  if (!on) return;
  std::string t;
  char buf[96];
  for (int i = 0; i < n; ++i) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
    snprintf(buf, sizeof(buf), "%s%s=%.6f", i ? ";" : "", names[i], ms);
    t += buf;
  }
  std::lock_guard<std::mutex> lk(g_profile_mu);
  g_profile_text = t;
}

// Throughput probe for the roofline peak: 8 independent v_mad_u64_u32 + v_addc_co_u32 chains
// per lane (the same instruction pair the field multiplication issues per 32x32 product).
__global__ void __launch_bounds__(256) k_probe_mac(uint32_t* out, int iters, uint32_t seed) {
  uint32_t a = seed + threadIdx.x, b = seed ^ (blockIdx.x * 2654435761u);
  uint64_t acc[8];
  uint32_t hi[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    acc[k] = a + k;
    hi[k] = 0;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) mac(acc[k], hi[k], a + k, b);
  }
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) r ^= (uint32_t)acc[k] ^ (uint32_t)(acc[k] >> 32) ^ hi[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

}  // namespace bte

using namespace bte;

extern "C" int bte_profile_enable(int on) {
  g_profile_on = on != 0;
  return BTE_OK;
}

extern "C" int bte_profile_read(char* buf, size_t len) {
  if (!buf || !len) return set_error(BTE_EINVAL, "profile_read: empty buffer");
  std::lock_guard<std::mutex> lk(g_profile_mu);
  snprintf(buf, len, "%s", g_profile_text.c_str());
  return BTE_OK;
}

extern "C" int bte_probe_mac_rate(double* macs_per_second, void* stream) {
  if (!macs_per_second) return set_error(BTE_EINVAL, "probe_mac_rate: null out");
  hipStream_t s = (hipStream_t)stream;
  const int blocks = 4096, threads = 256, iters = 2048;
  uint32_t* out = nullptr;
  BTE_HIP(hipMallocAsync((void**)&out, (size_t)blocks * threads * 4, s));
  hipEvent_t e0, e1;
  BTE_HIP(hipEventCreate(&e0));
  BTE_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_probe_mac, dim3(blocks), dim3(threads), 0, s, out, iters, 1u);  // warm
  BTE_HIP(hipEventRecord(e0, s));
  hipLaunchKernelGGL(k_probe_mac, dim3(blocks), dim3(threads), 0, s, out, iters, 2u);
  BTE_HIP(hipEventRecord(e1, s));
  BTE_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  BTE_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFreeAsync(out, s);
  BTE_HIP(hipStreamSynchronize(s));
  *macs_per_second = (double)blocks * threads * iters * 8 / (ms * 1e-3);
  return BTE_OK;
}

extern "C" const char* bte_last_error(void) { return g_last_error; }

extern "C" int bte_test_fail_alloc(int nth) {
  if (nth < 0) return set_error(BTE_EINVAL, "test_fail_alloc: negative count");
  g_fail_at.store(nth);
  return BTE_OK;
}

extern "C" int bte_tune_set(int knob, int64_t value) {
  if (knob < BTE_TUNE_MSM_ACC_ROUNDS || knob > BTE_TUNE_MM_MIN_THREADS)
    return set_error(BTE_EINVAL, "tune_set: unknown knob %d", knob);
  if (value < g_tune[knob].lo || value > g_tune[knob].hi)
    return set_error(BTE_EINVAL, "tune_set: knob %d value %lld outside %lld..%lld", knob, (long long)value,
                     (long long)g_tune[knob].lo, (long long)g_tune[knob].hi);
  g_tune[knob].value.store(value);
  return BTE_OK;
}

extern "C" int bte_tune_get(int knob, int64_t* value) {
  if (knob < BTE_TUNE_MSM_ACC_ROUNDS || knob > BTE_TUNE_MM_MIN_THREADS || !value)
    return set_error(BTE_EINVAL, "tune_get: bad arguments");
  *value = tune_get(knob);
  return BTE_OK;
}

extern "C" const char* bte_version(void) { return "weighted_bte_amd 0.1 gfx950"; }

extern "C" int bte_device_alloc(void** ptr, size_t bytes) {
  if (!ptr) return set_error(BTE_EINVAL, "device_alloc: null out pointer");
  BTE_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return BTE_OK;
}

extern "C" int bte_device_free(void* ptr) {
  BTE_HIP(hipFree(ptr));
  return BTE_OK;
}

extern "C" int bte_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  BTE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  BTE_HIP(hipStreamSynchronize((hipStream_t)stream));
  return BTE_OK;
}

extern "C" int bte_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  BTE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  BTE_HIP(hipStreamSynchronize((hipStream_t)stream));
  return BTE_OK;
}

extern "C" int bte_stream_synchronize(void* stream) {
  BTE_HIP(hipStreamSynchronize((hipStream_t)stream));
  return BTE_OK;
}
