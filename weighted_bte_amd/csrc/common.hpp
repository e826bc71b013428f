// Shared runtime helpers for the HIP library: status codes, error propagation, launch sizing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/bte_hip.h"

namespace bteh {
struct Fr;
}  // namespace bteh

namespace bte {

// Record the first failure for bte_last_error() and return its status code.
int set_error(int code, const char* fmt, ...);

// Launch-shape knobs set through bte_tune_set (capi.hip); drivers read them when they size a
// launch, and plans keep the values they were created with.
int64_t tune_get(int knob);

// True when a failure was injected for the next cache allocation (bte_test_fail_next_allocs).
bool test_alloc_should_fail();

#define BTE_HIP(call)                                                                \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::bte::set_error(BTE_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                              hipGetErrorString(e_));                                \
  } while (0)

#define BTE_CHECK_LAUNCH() BTE_HIP(hipGetLastError())

// Propagate a nonzero status from a library call.
#define BTE_TRY_RC(x)      \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

// Optional per-phase timing with HIP events on the stream the kernels run on.  When enabled
// (bte_profile_enable), drivers mark phase boundaries; after their stream sync the elapsed times
// are published and bte_profile_read returns them as "name=ms;...".
struct PhaseTimer {
  hipStream_t stream;
  bool on;
  int n = 0;
  const char* names[32];
  hipEvent_t ev[33];
  explicit PhaseTimer(hipStream_t s);
  ~PhaseTimer();
  void begin();
  void mark(const char* name);  // closes the phase that started at the previous mark
  void publish();               // call after the stream is synchronised
};

// Batched MSMs with gathered points and scalars (msm.hip); c = 0 picks the window width.
int g1_msm_batched_gather(const void* points, size_t n_points, const uint32_t* point_index, const uint8_t* scalars,
                          const uint32_t* scalar_index, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                          int c, void* out, hipStream_t s);
int g2_msm_batched_gather(const void* points, size_t n_points, const uint32_t* point_index, const uint8_t* scalars,
                          const uint32_t* scalar_index, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                          int c, void* out, hipStream_t s);

// k = k1 + k2 lambda (lambda = |z|^2 - 1, the G1 endomorphism's eigenvalue) as k1[4] then k2[4]
// little-endian u32 words, the layout g1_glv_mul reads (fft.hip).
void glv_decompose_g1(const bteh::Fr& k, uint32_t out[8]);

// Lagrange-at-zero coefficients (interpolation.rs:31-140) of the selected points (HOST indices)
// of the size-nextpow2(W) radix-2 domain, as canonical scalars in DEVICE memory, each multiplied
// by scale_canonical (32 bytes, HOST) when it is not null (fr_ntt.hip).
int lagrange_at_zero_device(size_t W, const uint64_t* sel, size_t n, const uint8_t* scale_canonical, uint8_t* d_out,
                            hipStream_t s);

// Segmented cyclotomic prefix inverse transforms (fft.hip).
struct fp12;
int cyc_prefix_ifft_segmented(fp12* v, const int* log_sizes, const uint32_t* out_counts, size_t count,
                              hipStream_t s);

// Miller loops where element i of segment s (seg_off[s] <= i < seg_off[s+1]) uses line table
// i - seg_off[s] of tables[s]; tables and seg_off are device arrays (pairing.hip).
int miller_prepared_segmented(const void* const* tables, const uint64_t* seg_off, size_t nseg, const void* p,
                              size_t n, void* out, hipStream_t s);

// Multi-Miller loops over prepared tables (pair g: table pair_tables[g], point p[g]) with
// HOST segment offsets; pairs of a segment share accumulators in chunks (pairing.hip).
int multi_miller_prepared(const void* const* pair_tables, const void* p, size_t npairs, const uint64_t* seg_offsets,
                          size_t nseg, void* out, hipStream_t s);
// Same with segment s = pairs [lo[s], hi[s]) (HOST arrays; ranges may leave gaps).
int multi_miller_prepared_ranges(const void* const* pair_tables, const void* p, const uint64_t* lo,
                                 const uint64_t* hi, size_t nseg, void* out, hipStream_t s);
// The two halves of it: host chunking, then launches over device copies of the chunk arrays
// (part: one Fp12 per chunk) with no host work, for plans and graph capture.
#include <vector>
struct MMChunks {
  std::vector<uint64_t> lo, hi, seg;  // chunk ranges; chunks of segment s are seg[s] .. seg[s+1]
};
void mm_chunks(const uint64_t* lo, const uint64_t* hi, size_t nseg, MMChunks& out);
int mm_enqueue(const void* const* tables, const void* p, const uint64_t* d_lo, const uint64_t* d_hi, size_t nchunks,
               const uint64_t* d_seg, size_t nseg, void* part, void* out, hipStream_t s);

// Cyclotomic powers over K batches of B values with one exponent per slot i (element b B + i):
// canonical scalars (HOST, B x 32 bytes) or small signed integers (device, B) (fft.hip).
int cyc_pow_scalars_tiled(fp12* v, const uint8_t* host_scalars, size_t B, size_t K, hipStream_t s);
int cyc_pow_small_tiled(fp12* v, const int64_t* coefficients, size_t B, size_t K, hipStream_t s);

// Launch-only pieces for plans and graph capture (fft.hip): odd-power storage of the three-lane
// cyclotomic stages, the prefix inverse transform over `count` transforms with caller storage,
// and domain tables built ahead of time.
size_t cyc_tab_bytes(size_t n, size_t count, size_t three_lane_max);
int cyc_prefix_batched_enqueue(fp12* v, size_t n, size_t count, size_t out_count, uint32_t* tab, size_t three_lane_max,
                               hipStream_t s);
int fft_warm(size_t n, bool inverse, bool plans, hipStream_t s);

// Segment holding flat index g: off[f] <= g < off[f + 1], off[0] = 0, count segments.
__device__ inline uint32_t seg_of(const uint64_t* __restrict__ off, uint32_t count, uint64_t g) {
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

inline unsigned blocks_for(size_t n, unsigned threads) {
  return (unsigned)((n + threads - 1) / threads);
}

}  // namespace bte
