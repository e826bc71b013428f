// Pippenger multi-scalar multiplication over G1 and G2 for gfx950.
//
// Replaces blst_p{1,2}s_mult_pippenger as called by weighted_btx/src/blst_utils.rs:43-67
// (g1_multi_exp_affine_bytes), :88-99 (g2_multi_exp_affine_bytes) and :106-152
// (g2_multi_exp_affine_indexed / _raw): sum_i s_i * P_i over the low 255 bits of each
// 32-byte little-endian scalar; an empty input gives the identity.
//
// Pipeline (one stream, all device work except the last serial combination):
//   1. digits    signed c-bit window digits d in [-(2^(c-1)-1), 2^(c-1)] and a histogram of |d|
//                per (window, bucket): for c <= 16 one workgroup per (window, slice of points)
//                counting in LDS, else one thread per point with global counters
//   2. scan      exclusive scan of the histogram -> bucket offsets in one flat sorted array
//   3. scatter   counting-sort placement of (point index, sign) into bucket order (LDS cursors
//                per (window, slice) for c <= 16)
//   4. accumulate  the sorted array is cut into chunks of K entries, one thread per chunk; a
//                thread sums each run of equal bucket with mixed Jacobian+affine additions.  A
//                bucket wholly inside a chunk is written directly; pieces of buckets that
//                cross chunk boundaries go to per-chunk head/tail slots.  Chunking keeps the
//                work per thread bounded however skewed the buckets are (sparse scalars pile
//                onto power-of-two buckets).
//   5. fixup     a bucket split over two chunks of one workgroup is closed by the accumulation
//                lane holding its tail; other split buckets are queued: one thread each when they
//                cross a few chunks, a 64-thread workgroup and a tree for the sparse-scalar pile-ups
//   6. segsum    per window, segments of L buckets: running sums give
//                acc_j = sum_k (k+1) B_{jL+k} and run_j = sum_k B_{jL+k}
//   7. window sums  per window, workgroups tree-sum A = sum_j acc_j (in two halves) and, for
//                each bit b of the segment index, D_b = sum_{j: bit b of j} run_j, so that
//                T_w = A + L * sum_b 2^b D_b
//   8. host      sum_w 2^(c w) T_w by one Horner pass over at most 256 doublings; this serial
//                tail takes a few tenths of a millisecond on a host core and would cost
//                milliseconds in one GPU lane.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "curve.hpp"
#include "host_field.hpp"

namespace bte {

// Sorted entries per accumulation thread: K = max(MSM_CHUNK_MIN, ceil(E / tstar)), where tstar,
// the accumulation thread count, is set on the host to a whole number of full-occupancy rounds
// of the device so the last round is not left part-empty.  Kernels that need K recompute it from
// the entry count E = offsets[nb].
static constexpr uint32_t MSM_CHUNK_MIN = 16;
BTE_DEV uint32_t chunk_size(uint32_t E, uint32_t tstar) {
  const uint32_t k = (uint32_t)(((uint64_t)E + tstar - 1) / tstar);
  return k > MSM_CHUNK_MIN ? k : MSM_CHUNK_MIN;
}

template <class F> struct MsmTraits;
// Threads per window-sum workgroup: 512 keeps a workgroup to two wavefronts per SIMD, so the
// full Jacobian addition fits its registers without spilling (1024 spilled and was slower).
// ACC_WAVES: minimum wavefronts per SIMD asked of the accumulation kernel (G1 fits three with a
// few bytes of spill in its closing addition; G2 needs the whole register file).
template <> struct MsmTraits<fp> {
  using HF = bteh::Fp;
  static constexpr int REDUCE_THREADS = 512;
  static constexpr int ACC_WAVES = 3;
};
template <> struct MsmTraits<fp2> {
  using HF = bteh::Fp2;
  static constexpr int REDUCE_THREADS = 512;
  static constexpr int ACC_WAVES = 1;
};

struct MsmShape {
  int c;      // window bits
  int nw;     // windows, ceil(256 / c)
  uint32_t M; // buckets per window, 2^(c-1)
  int logL;   // log2 of buckets per reduction segment
  int logS;   // log2 of segments per window
};

static int g_window_override[2][64];  // [curve][log2 n] -> c (0 = default)
static std::mutex g_window_mu;

static int floor_log2(size_t n) {
  int lg = 0;
  while (lg < 63 && ((size_t)1 << (lg + 1)) <= n) ++lg;
  return lg;
}

// Built-in window width by log2 n.  Measured on MI355X (G1 and G2, 10% sparse scalars): widths
// 8, 10, 13 and 16 run well and the widths between them badly (G1 2^16: 1.69 ms at 10 against
// 2.03 at 12; 2^18: 2.59 at 13 against 3.86 at 14; 2^19: 3.84 at 16 against 11.3 at 15), so
// the table keeps to those four from 2^13 up; below that, log2 n - 4.
static int default_window_bits(int lg) {
  if (lg <= 12) return std::max(4, lg - 4);
  if (lg <= 14) return 8;
  if (lg <= 16) return 10;
  if (lg <= 18) return 13;
  return 16;
}

static MsmShape msm_shape(int curve, size_t n) {
  int lg = floor_log2(n);
  int c;
  {
    std::lock_guard<std::mutex> lk(g_window_mu);
    c = g_window_override[curve][lg];
  }
  if (c <= 0) c = default_window_bits(lg);
  MsmShape s;
  s.c = c;
  s.nw = (256 + c - 1) / c;
  s.M = 1u << (c - 1);
  s.logL = std::min(c - 1, 2);  // 4 buckets per segment thread: two wavefronts per SIMD in segsum
  s.logS = (c - 1) - s.logL;
  return s;
}

// ---------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_msm_digits(const uint32_t* __restrict__ scalars, size_t n, int c, int nw,
                             uint32_t M, int32_t* __restrict__ digits,
                             uint32_t* __restrict__ counts) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s[9];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = scalars[8 * i + k];
  s[7] &= 0x7fffffffu;  // nbits = 255
  s[8] = 0;
  const uint32_t mask = (1u << c) - 1u;
  int carry = 0;
  for (int w = 0; w < nw; ++w) {
    const int bit = w * c;
    const int word = bit >> 5, off = bit & 31;
    uint64_t two = (uint64_t)s[word] | ((uint64_t)s[word + 1] << 32);
    int d = (int)((two >> off) & mask) + carry;
    if (d > (int)M) {
      d -= (int)(1u << c);
      carry = 1;
    } else {
      carry = 0;
    }
    digits[(size_t)w * n + i] = d;
    if (d) atomicAdd(&counts[(size_t)w * M + (uint32_t)(d < 0 ? -d : d) - 1u], 1u);
  }
}

// Exclusive scan of counts[0..nb) into offsets[0..nb], offsets[nb] = total.  Three passes over
// 1024-element tiles (256 threads x 4).
__device__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* sh, uint32_t* total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    uint32_t t = tid >= off ? sh[tid - off] : 0u;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  uint32_t incl = sh[tid];
  *total = sh[255];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_tiles(const uint32_t* __restrict__ in, size_t nb, uint32_t* __restrict__ out,
                             uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t sh[256];
  size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = base + k < nb ? in[base + k] : 0u;
  uint32_t local = v[0] + v[1] + v[2] + v[3];
  uint32_t total;
  uint32_t ex = block_exclusive_scan_256(local, sh, &total);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < nb) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void k_scan_tile_sums(uint32_t* tile_sums, int ntiles, uint32_t* out_total) {
  __shared__ uint32_t sh[256];
  uint32_t carry = 0;
  for (int base = 0; base < ntiles; base += 1024) {
    int idx = base + threadIdx.x * 4;
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = idx + k < ntiles ? tile_sums[idx + k] : 0u;
    uint32_t total;
    uint32_t ex = block_exclusive_scan_256(v[0] + v[1] + v[2] + v[3], sh, &total);
    ex += carry;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (idx + k < ntiles) tile_sums[idx + k] = ex;
      ex += v[k];
    }
    carry += total;
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ void k_scan_add(uint32_t* out, size_t nb, const uint32_t* tile_sums) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < nb) out[i] += tile_sums[i / 1024];
}

__global__ void k_msm_scatter(const int32_t* __restrict__ digits, size_t n, int nw, uint32_t M,
                              const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursor,
                              uint32_t* __restrict__ sorted) {
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= n * (size_t)nw) return;
  const int d = digits[t];
  if (!d) return;
  const size_t w = t / n;
  const size_t i = t - w * n;
  const uint32_t b = (uint32_t)(w * M) + (uint32_t)(d < 0 ? -d : d) - 1u;
  const uint32_t pos = offsets[b] + atomicAdd(&cursor[b], 1u);
  sorted[pos] = (uint32_t)i | (d < 0 ? 0x80000000u : 0u);
}

// Counting sort with per-workgroup histograms in LDS (windows of up to 16 bits).  Workgroup
// (w, s) owns window w of the points in slice s: it counts that slice's digits in LDS and writes
// the M counts to hist[w][s][.]; a column scan turns them into per-slice starts inside each
// bucket; the scatter pass recomputes the digits and places entries through LDS cursors.  No
// global atomics, and the workgroups of one slice read the same scalars (with S a multiple of
// 8 they sit on one XCD and share its L2).
BTE_DEV int scalar_window_digit(const uint32_t* __restrict__ sc, int c, int w, uint32_t M) {
  uint32_t s[9];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = sc[k];
  s[7] &= 0x7fffffffu;  // nbits = 255
  s[8] = 0;
  const uint32_t mask = (1u << c) - 1u;
  int carry = 0, d = 0;
  for (int v = 0; v <= w; ++v) {  // the recoding of k_msm_digits up to window w
    const int bit = v * c;
    const int word = bit >> 5, off = bit & 31;
    const uint64_t two = (uint64_t)s[word] | ((uint64_t)s[word + 1] << 32);
    d = (int)((two >> off) & mask) + carry;
    carry = d > (int)M;
    if (carry) d -= (int)(1u << c);
  }
  return d;
}

template <int LDS_WORDS, int T>
__global__ void __launch_bounds__(T) k_msm_hist(const uint32_t* __restrict__ scalars, size_t n, int c, uint32_t M,
                                                uint32_t S, size_t P, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[LDS_WORDS];
  const uint32_t w = blockIdx.x / S, s = blockIdx.x - w * S;
  for (uint32_t b = threadIdx.x; b < M; b += T) h[b] = 0;
  __syncthreads();
  const size_t lo = (size_t)s * P, hi = lo + P < n ? lo + P : n;
  for (size_t i = lo + threadIdx.x; i < hi; i += T) {
    const int d = scalar_window_digit(scalars + 8 * i, c, (int)w, M);
    if (d) atomicAdd(&h[(uint32_t)(d < 0 ? -d : d) - 1u], 1u);
  }
  __syncthreads();
  uint32_t* out = hist + ((size_t)w * S + s) * M;
  for (uint32_t b = threadIdx.x; b < M; b += T) out[b] = h[b];
}

// hist[w][s][b] <- sum_{s' < s} hist[w][s'][b]; counts[w M + b] <- the column total.
__global__ void k_msm_hist_colscan(uint32_t* __restrict__ hist, int nw, uint32_t M, uint32_t S,
                                   uint32_t* __restrict__ counts) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= (size_t)nw * M) return;
  const size_t w = t / M, b = t - w * M;
  uint32_t* col = hist + w * S * M + b;
  uint32_t run = 0;
  for (uint32_t s = 0; s < S; ++s) {
    const uint32_t v = col[(size_t)s * M];
    col[(size_t)s * M] = run;
    run += v;
  }
  counts[t] = run;
}

template <int LDS_WORDS, int T>
__global__ void __launch_bounds__(T) k_msm_scatter_slices(const uint32_t* __restrict__ scalars, size_t n, int c,
                                                          uint32_t M, uint32_t S, size_t P,
                                                          const uint32_t* __restrict__ hist,
                                                          const uint32_t* __restrict__ offsets,
                                                          uint32_t* __restrict__ sorted) {
  __shared__ uint32_t cur[LDS_WORDS];
  const uint32_t w = blockIdx.x / S, s = blockIdx.x - w * S;
  const uint32_t* pre = hist + ((size_t)w * S + s) * M;
  const uint32_t* off = offsets + (size_t)w * M;
  for (uint32_t b = threadIdx.x; b < M; b += T) cur[b] = off[b] + pre[b];
  __syncthreads();
  const size_t lo = (size_t)s * P, hi = lo + P < n ? lo + P : n;
  for (size_t i = lo + threadIdx.x; i < hi; i += T) {
    const int d = scalar_window_digit(scalars + 8 * i, c, (int)w, M);
    if (d) {
      const uint32_t pos = atomicAdd(&cur[(uint32_t)(d < 0 ? -d : d) - 1u], 1u);
      sorted[pos] = (uint32_t)i | (d < 0 ? 0x80000000u : 0u);
    }
  }
}

// Accumulation workgroups are one wavefront: chunk t is lane t % 64 of workgroup t / 64.
static constexpr int ACC_BLOCK = 64;

template <class F>
__global__ void __launch_bounds__(ACC_BLOCK, MsmTraits<F>::ACC_WAVES)
    k_msm_accumulate(const aff<F>* __restrict__ points, uint32_t stride, const uint32_t* __restrict__ index,
                     size_t n_points,
                     const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ offsets, uint32_t nb,
                     uint32_t tstar, jac<F>* __restrict__ buckets, jac<F>* __restrict__ heads,
                     jac<F>* __restrict__ tails, int* __restrict__ err) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const uint32_t E = offsets[nb];
  const uint32_t K = chunk_size(E, tstar);
  const size_t start = t * K;
  const size_t stop = start + K < E ? start + K : E;
  jac<F> acc = jac_identity<F>();
  uint32_t b = 0;
  bool pending = false;  // last piece is a tail that the next lane's head completes
  if (start < E) {
    // Bucket holding `start`: offsets[lo] <= start < offsets[hi], narrowed to hi = lo + 1.
    uint32_t lo = 0, hi = nb;
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) >> 1;
      if (offsets[mid] <= start) lo = mid; else hi = mid;
    }
    b = lo;
    uint32_t bs = offsets[b], be = offsets[b + 1];
    auto emit = [&]() {
      if (bs < start) heads[t] = acc;
      else if (be > stop) tails[t] = acc;
      else buckets[b] = acc;
    };
    for (size_t pos = start; pos < stop; ++pos) {
      if (pos == be) {
        emit();
        acc = jac_identity<F>();
        do {
          ++b;
          bs = be;
          be = offsets[b + 1];
        } while (be == bs);
      }
      const uint32_t e = sorted[pos];
      uint32_t idx = e & 0x7fffffffu;
      if (index) {
        idx = index[idx];
        if (idx >= n_points) {
          *err = 1;
          continue;
        }
      }
      aff<F> q = *(const aff<F>*)((const char*)points + (size_t)idx * stride);
      if (e >> 31) q = aff_neg(q);
      acc = jac_add_aff(acc, q);
    }
    // A bucket that starts in this chunk and ends in the next one, held by the next lane of
    // this workgroup, is closed here; k_msm_fixup_queue leaves exactly these buckets alone.
    pending = bs >= start && be > stop && (be - 1) / K == t + 1 && threadIdx.x + 1 < ACC_BLOCK;
    if (!pending) emit();
  }
  __syncthreads();  // the workgroup's head pieces are written
  if (pending) buckets[b] = jac_add(acc, heads[t + 1]);
}

// Point records padded to a whole number of 128-byte lines (G1: 96 -> 128 bytes, G2: 192 -> 256),
// so each gather of the accumulation reads one or two full lines instead of straddling three.
template <class F> struct PadStride {
  static constexpr uint32_t value = (uint32_t)((sizeof(aff<F>) + 127) / 128 * 128);
};
template <class F>
__global__ void k_msm_pad_points(const aff<F>* __restrict__ in, size_t n, char* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) *(aff<F>*)(out + i * PadStride<F>::value) = in[i];
}

// Appends the threads with `want` set to a device queue with one atomic per workgroup of T
// threads; every thread of the workgroup must call it.  sh holds T / 64 + 1 words.
template <int T>
BTE_DEV uint32_t block_append(uint32_t* counter, bool want, uint32_t* sh) {
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  const uint64_t mask = __ballot(want);
  if (lane == 0) sh[wv] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int i = 0; i < T / 64; ++i) {
      const uint32_t c = sh[i];
      sh[i] = tot;
      tot += c;
    }
    sh[T / 64] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  const uint32_t r = sh[T / 64] + sh[wv] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  __syncthreads();  // sh is reused by the next call
  return r;
}

// Buckets that cross chunk boundaries and were not closed by k_msm_accumulate: those spanning at
// most FIX_SERIAL + 1 chunks go to the `mid` queue (one thread each), the rest (sparse-scalar
// pile-ups) to the `big` queue (one workgroup each).  Empty buckets are set to the identity.
static constexpr uint32_t FIX_SERIAL = 8;
static constexpr int QUEUE_BLOCK = 1024;

template <class F>
__global__ void __launch_bounds__(QUEUE_BLOCK) k_msm_fixup_queue(const uint32_t* __restrict__ offsets, uint32_t nb,
                                                                 uint32_t tstar, jac<F>* __restrict__ buckets,
                                                                 uint32_t* __restrict__ mid, uint32_t* __restrict__ nmid,
                                                                 uint32_t* __restrict__ big, uint32_t* __restrict__ nbig) {
  __shared__ uint32_t sh[QUEUE_BLOCK / 64 + 1];
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t K = chunk_size(offsets[nb], tstar);
  bool is_mid = false, is_big = false;
  if (b < nb) {
    const uint32_t bs = offsets[b], be = offsets[b + 1];
    if (bs == be) {
      buckets[b] = jac_identity<F>();
    } else {
      const uint32_t cs = bs / K, ce = (be - 1) / K;
      const bool closed = ce == cs || (ce == cs + 1 && cs / ACC_BLOCK == ce / ACC_BLOCK);
      is_big = ce - cs > FIX_SERIAL;
      is_mid = !closed && !is_big;
    }
  }
  const uint32_t qm = block_append<QUEUE_BLOCK>(nmid, is_mid, sh);
  const uint32_t qb = block_append<QUEUE_BLOCK>(nbig, is_big, sh);
  if (is_mid) mid[qm] = b;
  if (is_big) big[qb] = b;
}

// One launch for both queues so the two kinds overlap: workgroups below gbig take big buckets in
// turn (64 threads: strided partial sums over the pieces, then a tree in LDS); the others give
// each mid bucket one thread that adds its few pieces in order.
template <class F>
__global__ void __launch_bounds__(64) k_msm_fixup(const uint32_t* __restrict__ offsets, uint32_t nb, uint32_t tstar,
                                                  jac<F>* __restrict__ buckets, const jac<F>* __restrict__ heads,
                                                  const jac<F>* __restrict__ tails, const uint32_t* __restrict__ mid,
                                                  const uint32_t* __restrict__ nmid, const uint32_t* __restrict__ big,
                                                  const uint32_t* __restrict__ nbig, uint32_t gbig) {
  __shared__ jac<F> sh[64];
  const uint32_t K = chunk_size(offsets[nb], tstar);
  const int tid = threadIdx.x;
  if (blockIdx.x < gbig) {
    const uint32_t nq = *nbig;
    for (uint32_t qi = blockIdx.x; qi < nq; qi += gbig) {  // uniform per workgroup
      const uint32_t b = big[qi];
      const uint32_t bs = offsets[b], be = offsets[b + 1];
      const uint32_t cs = bs / K, ce = (be - 1) / K;
      jac<F> acc = tid == 0 ? tails[cs] : jac_identity<F>();
      for (uint32_t c = cs + 1 + tid; c <= ce; c += 64) acc = jac_add(acc, heads[c]);
      sh[tid] = acc;
      __syncthreads();
      for (int w = 32; w > 0; w >>= 1) {
        if (tid < w) sh[tid] = jac_add(sh[tid], sh[tid + w]);
        __syncthreads();
      }
      if (tid == 0) buckets[b] = sh[0];
      __syncthreads();
    }
  } else {
    const uint32_t nq = *nmid;
    const uint32_t stride = (gridDim.x - gbig) * 64u;
    for (uint32_t qi = (blockIdx.x - gbig) * 64u + tid; qi < nq; qi += stride) {
      const uint32_t b = mid[qi];
      const uint32_t bs = offsets[b], be = offsets[b + 1];
      const uint32_t cs = bs / K, ce = (be - 1) / K;
      jac<F> acc = tails[cs];
      for (uint32_t c = cs + 1; c <= ce; ++c) acc = jac_add(acc, heads[c]);
      buckets[b] = acc;
    }
  }
}

template <class F>
__global__ void __launch_bounds__(64) k_msm_segsum(const jac<F>* __restrict__ buckets, int nw, uint32_t M, int logL,
                             int logS, jac<F>* __restrict__ seg_acc, jac<F>* __restrict__ seg_run) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t S = (size_t)1 << logS;
  if (t >= (size_t)nw * S) return;
  const size_t w = t >> logS, j = t & (S - 1);
  const int L = 1 << logL;
  const jac<F>* base = buckets + w * M + (j << logL);
  jac<F> run = jac_identity<F>(), acc = jac_identity<F>();
  for (int k = L - 1; k >= 0; --k) {
    run = jac_add(run, base[k]);
    acc = jac_add(acc, run);
  }
  seg_acc[t] = acc;
  seg_run[t] = run;
}

// Per window w, 2 + logS sums of at most S/2 reduction segments each, one workgroup per sum and
// every thread busy in the strided phase: sum 0 = sum_{j < ceil(S/2)} acc_j, sum 1 = the other
// acc_j, sum 2 + b = D_b = sum_{j with bit b set} run_j; T_w = sum 0 + sum 1 + L sum_b 2^b D_b.
template <class F, int T>
__global__ void __launch_bounds__(T) k_msm_window_sums(const jac<F>* __restrict__ seg_acc,
                                                       const jac<F>* __restrict__ seg_run, int logS,
                                                       jac<F>* __restrict__ out) {
  __shared__ jac<F> sh[T];
  const int nsum = 2 + logS;
  const int w = blockIdx.x / nsum, sid = blockIdx.x % nsum;
  const size_t S = (size_t)1 << logS;
  const int tid = threadIdx.x;
  jac<F> acc = jac_identity<F>();
  if (sid < 2) {
    const size_t n0 = (S + 1) / 2;
    const size_t lo = sid ? n0 : 0, hi = sid ? S : n0;
    for (size_t j = lo + tid; j < hi; j += T) acc = jac_add(acc, seg_acc[w * S + j]);
  } else {
    const int b = sid - 2;
    const size_t low = ((size_t)1 << b) - 1;
    for (size_t i = tid; i < S / 2; i += T) {
      const size_t j = ((i >> b) << (b + 1)) | ((size_t)1 << b) | (i & low);
      acc = jac_add(acc, seg_run[w * S + j]);
    }
  }
  sh[tid] = acc;
  __syncthreads();
  for (int s = T / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = jac_add(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  if (tid == 0) out[blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------------------------------------
// Batched (segmented) MSM: many independent MSMs of uneven sizes through one pipeline.  Term t
// belongs to segment seg(t) (seg_offsets); its window digits go to bucket
// (seg * nw + w) * M + |d| - 1, so the sort / accumulate / fix-up kernels above run unchanged
// over all segments at once.  Each (segment, window) then reduces its M buckets with one
// running sum, and one thread per segment combines the windows (Horner, c doublings per
// window).  Nothing synchronises with the host, so the sequence can be graph-captured.
// ---------------------------------------------------------------------------------------------
static constexpr uint32_t NO_DIGIT = 0xffffffffu;

__global__ void k_msm_digits_seg(const uint32_t* __restrict__ scalars, const uint32_t* __restrict__ scalar_index,
                                 size_t nterms,
                                 const uint64_t* __restrict__ seg_offsets, uint32_t nseg, int c, int nw,
                                 uint32_t M, uint32_t* __restrict__ keys, uint32_t* __restrict__ counts) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= nterms) return;
  uint32_t lo = 0, hi = nseg;  // seg_offsets[lo] <= t < seg_offsets[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (seg_offsets[mid] <= t) lo = mid; else hi = mid;
  }
  const uint32_t seg = lo;
  const size_t si = scalar_index ? scalar_index[t] : t;
  uint32_t s[9];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = scalars[8 * si + k];
  s[7] &= 0x7fffffffu;
  s[8] = 0;
  const uint32_t mask = (1u << c) - 1u;
  int carry = 0;
  for (int w = 0; w < nw; ++w) {
    const int bit = w * c;
    const int word = bit >> 5, off = bit & 31;
    uint64_t two = (uint64_t)s[word] | ((uint64_t)s[word + 1] << 32);
    int d = (int)((two >> off) & mask) + carry;
    if (d > (int)M) {
      d -= (int)(1u << c);
      carry = 1;
    } else {
      carry = 0;
    }
    uint32_t e = NO_DIGIT;
    if (d) {
      const uint32_t key = (seg * (uint32_t)nw + (uint32_t)w) * M + (uint32_t)(d < 0 ? -d : d) - 1u;
      atomicAdd(&counts[key], 1u);
      e = (key << 1) | (d < 0 ? 1u : 0u);
    }
    keys[(size_t)w * nterms + t] = e;
  }
}

__global__ void k_msm_scatter_keys(const uint32_t* __restrict__ keys, size_t nterms, int nw,
                                   const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursor,
                                   uint32_t* __restrict__ sorted) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= nterms * (size_t)nw) return;
  const uint32_t e = keys[i];
  if (e == NO_DIGIT) return;
  const uint32_t key = e >> 1;
  const uint32_t term = (uint32_t)(i % nterms);
  const uint32_t pos = offsets[key] + atomicAdd(&cursor[key], 1u);
  sorted[pos] = term | ((e & 1u) << 31);
}

// Per (segment, window) t: T_t = sum_j acc_j + L * sum_j j * run_j over its S reduction
// segments (the bucket sum sum_k k B_k split at granularity L = 2^logL).
template <class F>
__global__ void __launch_bounds__(64) k_msm_seg_window(const jac<F>* __restrict__ seg_acc,
                                                       const jac<F>* __restrict__ seg_run, size_t nsw, int logL,
                                                       int logS, jac<F>* __restrict__ win) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= nsw) return;
  const size_t S = (size_t)1 << logS;
  const jac<F>* acc = seg_acc + t * S;
  const jac<F>* run = seg_run + t * S;
  jac<F> a = jac_identity<F>(), r = jac_identity<F>(), rs = jac_identity<F>();
  for (size_t j = S; j-- > 1;) {
    r = jac_add(r, run[j]);
    rs = jac_add(rs, r);  // sum_j j * run_j
  }
  for (int k = 0; k < logL; ++k) rs = jac_dbl(rs);
  for (size_t j = 0; j < S; ++j) a = jac_add(a, acc[j]);
  win[t] = jac_add(a, rs);
}

template <class F>
__global__ void __launch_bounds__(64) k_msm_combine(const jac<F>* __restrict__ win, uint32_t nseg, int nw, int c,
                                                    jac<F>* __restrict__ out) {
  const uint32_t sgm = blockIdx.x * blockDim.x + threadIdx.x;
  if (sgm >= nseg) return;
  jac<F> acc = jac_identity<F>();
  for (int w = nw - 1; w >= 0; --w) {
    if (!jac_is_identity(acc))
      for (int k = 0; k < c; ++k) acc = jac_dbl(acc);
    acc = jac_add(acc, win[(size_t)sgm * nw + w]);
  }
  out[sgm] = acc;
}

// ---------------------------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------------------------
// Accumulation threads: BTE_TUNE_MSM_ACC_ROUNDS (default 3) full-occupancy rounds of one-wavefront workgroups on this
// device (queried once per device), and no more than chunks of the minimum size need.
template <class F>
static uint32_t acc_threads(size_t E_max) {
  static std::mutex mu;
  static int slots[64];
  int dev = 0;
  (void)hipGetDevice(&dev);
  int sl;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (dev < 0 || dev >= 64) dev = 0;
    if (!slots[dev]) {
      int cus = 0, per_cu = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_msm_accumulate<F>),
                                                       ACC_BLOCK, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
      slots[dev] = cus * per_cu;
    }
    sl = slots[dev];
  }
  const size_t full = (size_t)sl * ACC_BLOCK * (size_t)tune_get(BTE_TUNE_MSM_ACC_ROUNDS);
  const size_t need = (E_max + MSM_CHUNK_MIN - 1) / MSM_CHUNK_MIN;
  return (uint32_t)std::max<size_t>(1, std::min(full, need));
}

// Slices per window for the LDS-histogram sort: about 512 workgroups in all, slices of at least
// 2048 points, S a multiple of 8 from 8 up, and at most 2^23 histogram counters.
static uint32_t sort_slices(size_t n, int nw, uint32_t M) {
  size_t S = std::max<size_t>(1, std::min<size_t>(512 / nw, n / 2048));
  if (S >= 8) S &= ~(size_t)7;
  while (S > 1 && (size_t)nw * S * M > ((size_t)1 << 23)) S = S > 8 ? S - 8 : S - 1;
  return (uint32_t)S;
}

// LDS tier by bucket count: counters per workgroup and threads per workgroup.
template <int W, int T>
static void lds_hist_launch(bool scatter, unsigned nwg, const uint32_t* scalars, size_t n, int c, uint32_t M,
                            uint32_t S, size_t P, uint32_t* hist, const uint32_t* offsets, uint32_t* sorted,
                            hipStream_t s) {
  if (!scatter)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_hist<W, T>), dim3(nwg), dim3(T), 0, s, scalars, n, c, M, S, P, hist);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_scatter_slices<W, T>), dim3(nwg), dim3(T), 0, s, scalars, n, c, M, S,
                       P, (const uint32_t*)hist, offsets, sorted);
}

static void lds_sort_pass(bool scatter, unsigned nwg, const uint32_t* scalars, size_t n, int c, uint32_t M,
                          uint32_t S, size_t P, uint32_t* hist, const uint32_t* offsets, uint32_t* sorted,
                          hipStream_t s) {
  if (M <= 2048)
    lds_hist_launch<2048, 256>(scatter, nwg, scalars, n, c, M, S, P, hist, offsets, sorted, s);
  else if (M <= 8192)
    lds_hist_launch<8192, 512>(scatter, nwg, scalars, n, c, M, S, P, hist, offsets, sorted, s);
  else
    lds_hist_launch<32768, 1024>(scatter, nwg, scalars, n, c, M, S, P, hist, offsets, sorted, s);
}

// Host tail of a single MSM: the window sums come back into pinned memory, a host callback on
// the stream combines them (Horner over the exponent: A_w sits at 2^(c w), D_{w,b} at
// 2^(c w + logL + b)) and the result goes back to the device, all in stream order, so the call
// neither waits nor allocates and can be captured into a graph.
template <class HF> struct HornerJob {
  int c, nw, logL, logS, err;
  bteh::Jac<HF> result;
  bteh::Jac<HF>* hw;  // nw x (2 + logS) sums
};

template <class HF> static void horner_host(void* data) {
  HornerJob<HF>* j = (HornerJob<HF>*)data;
  const int nsum = 2 + j->logS, top = j->c * j->nw;
  std::vector<std::vector<int>> at(top + 1);
  for (int w = 0; w < j->nw; ++w) {
    at[j->c * w].push_back(w * nsum);
    at[j->c * w].push_back(w * nsum + 1);
    for (int b = 0; b < j->logS; ++b) at[j->c * w + j->logL + b].push_back(w * nsum + 2 + b);
  }
  bteh::Jac<HF> acc = bteh::jac_identity<HF>();
  for (int e = top; e >= 0; --e) {
    acc = bteh::jac_dbl(acc);
    for (int k : at[e]) acc = bteh::jac_add(acc, j->hw[k]);
  }
  j->result = acc;
}

template <class F> __global__ void k_msm_identity(jac<F>* out) {
  if (threadIdx.x == 0) *out = jac_identity<F>();
}

// The same Horner pass on the device, one thread, in the host's order of operations (so the same
// Jacobian coordinates): used when the stream is being captured into a graph, where the host
// callback form is not relied on.  About 256 doublings and 240 additions in one lane: a few
// milliseconds, against tenths of a millisecond on a host core.
template <class F>
__global__ void __launch_bounds__(64) k_msm_horner(const jac<F>* __restrict__ wsum, int c, int nw, int logL, int logS,
                                                   jac<F>* __restrict__ out) {
  if (threadIdx.x != 0) return;
  const int nsum = 2 + logS, top = c * nw;
  jac<F> acc = jac_identity<F>();
  for (int e = top; e >= 0; --e) {
    acc = jac_dbl(acc);
    const int w = e / c, r = e - w * c, b = r - logL;
    if (w >= nw) continue;
    if (r == 0) {
      acc = jac_add(acc, wsum[w * nsum]);
      acc = jac_add(acc, wsum[w * nsum + 1]);
    }
    if (b >= 0 && b < logS) acc = jac_add(acc, wsum[w * nsum + 2 + b]);
  }
  *out = acc;
}

// The host callback's result, read from pinned host memory when the kernel runs.  (A host-to-
// device copy node of a captured graph may take its bytes when it is recorded, before any
// callback has run; a kernel argument holding the address reads them at replay time.)
template <class F> __global__ void k_msm_result_from_host(const jac<F>* __restrict__ host, jac<F>* __restrict__ out) {
  constexpr int W = sizeof(jac<F>) / 4;
  const volatile uint32_t* src = (const volatile uint32_t*)host;
  uint32_t* dst = (uint32_t*)out;
  for (int i = threadIdx.x; i < W; i += blockDim.x) dst[i] = src[i];
}

template <class F>
static int msm_run(int curve, const aff<F>* points, size_t n_points, const uint32_t* index,
                   const uint8_t* scalars, size_t n, jac<F>* out, hipStream_t s) {
  using HF = typename MsmTraits<F>::HF;
  if (n == 0) {
    if (!out) return set_error(BTE_EINVAL, "msm: null pointer");
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_identity<F>), dim3(1), dim3(64), 0, s, out);
    BTE_CHECK_LAUNCH();
    return BTE_OK;
  }
  if (!points || !scalars || !out) return set_error(BTE_EINVAL, "msm: null pointer");
  if (((uintptr_t)scalars & 3) != 0) return set_error(BTE_EINVAL, "msm: scalars must be 4-byte aligned");
  if (n > 0x7fffffffu / 2) return set_error(BTE_EINVAL, "msm: n too large");
  const MsmShape sh = msm_shape(curve, n);
  const uint32_t nb = (uint32_t)sh.nw * sh.M;
  const size_t E_max = n * (size_t)sh.nw;
  const uint32_t tstar = acc_threads<F>(E_max);
  const size_t S = (size_t)1 << sh.logS;
  const int nsum = 2 + sh.logS;
  const int ntiles = (int)((nb + 1023) / 1024);
  if (ntiles > (1 << 20)) return set_error(BTE_EINVAL, "msm: window too wide");

  // Windows of up to 16 bits sort through LDS histograms; wider ones through global counters.
  const bool lds_sort = sh.M <= 32768;
  const uint32_t S_sort = lds_sort ? sort_slices(n, sh.nw, sh.M) : 0;

  CallScope sc(s);
  int32_t* digits = nullptr;
  uint32_t *counts, *offsets, *cursor = nullptr, *sorted, *tile_sums, *hist = nullptr;
  jac<F> *buckets, *heads, *tails, *seg_acc, *seg_run, *wsum;
  int* err;
  uint32_t *big, *nbig, *mid, *nmid;
  const size_t max_big = tstar / FIX_SERIAL + 1;
  if (lds_sort) {
    BTE_TRY_RC(sc.dev(&hist, (size_t)sh.nw * S_sort * sh.M));
  } else {
    BTE_TRY_RC(sc.dev(&digits, E_max));
    BTE_TRY_RC(sc.dev(&cursor, nb));
  }
  BTE_TRY_RC(sc.dev(&counts, nb));
  BTE_TRY_RC(sc.dev(&offsets, (size_t)nb + 1));
  BTE_TRY_RC(sc.dev(&sorted, E_max));
  BTE_TRY_RC(sc.dev(&tile_sums, ntiles));
  BTE_TRY_RC(sc.dev(&buckets, nb));
  BTE_TRY_RC(sc.dev(&heads, tstar));
  BTE_TRY_RC(sc.dev(&tails, tstar));
  BTE_TRY_RC(sc.dev(&seg_acc, sh.nw * S));
  BTE_TRY_RC(sc.dev(&seg_run, sh.nw * S));
  BTE_TRY_RC(sc.dev(&wsum, (size_t)sh.nw * nsum));
  BTE_TRY_RC(sc.dev(&err, 1));
  BTE_TRY_RC(sc.dev(&big, max_big));
  BTE_TRY_RC(sc.dev(&nbig, 1));
  BTE_TRY_RC(sc.dev(&mid, nb));
  BTE_TRY_RC(sc.dev(&nmid, 1));
  const aff<F>* acc_points = points;
  uint32_t stride = (uint32_t)sizeof(aff<F>);
  if (!index) {  // contiguous bases: gather from line-aligned copies
    char* padded;
    BTE_TRY_RC(sc.dev(&padded, n * (size_t)PadStride<F>::value));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_pad_points<F>), dim3(blocks_for(n, 256)), dim3(256), 0, s, points, n,
                       padded);
    acc_points = (const aff<F>*)padded;
    stride = PadStride<F>::value;
  }
  HornerJob<HF>* job = nullptr;
  if (!sc.capturing()) {
    BTE_TRY_RC(sc.host(&job, 1));
    BTE_TRY_RC(sc.host(&job->hw, (size_t)sh.nw * nsum));
    job->c = sh.c;
    job->nw = sh.nw;
    job->logL = sh.logL;
    job->logS = sh.logS;
  }

  PhaseTimer pt(s);
  pt.begin();
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)err, (size_t)1, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nbig, (size_t)1, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nmid, (size_t)1, 0u);
  const size_t P_sort = lds_sort ? (n + S_sort - 1) / S_sort : 0;
  const unsigned nwg_sort = (unsigned)(sh.nw * S_sort);
  if (lds_sort) {
    lds_sort_pass(false, nwg_sort, (const uint32_t*)scalars, n, sh.c, sh.M, S_sort, P_sort, hist, nullptr, nullptr,
                  s);
    hipLaunchKernelGGL(k_msm_hist_colscan, dim3(blocks_for(nb, 256)), dim3(256), 0, s, hist, sh.nw, sh.M, S_sort,
                       counts);
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nb, 256)), dim3(256), 0, s, counts, (size_t)nb, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nb, 256)), dim3(256), 0, s, cursor, (size_t)nb, 0u);
    hipLaunchKernelGGL(k_msm_digits, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                       (const uint32_t*)scalars, n, sh.c, sh.nw, sh.M, digits, counts);
  }
  pt.mark("digits");
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(256), 0, s, counts, (size_t)nb, offsets, tile_sums);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(256), 0, s, tile_sums, ntiles, offsets + nb);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(nb, 256)), dim3(256), 0, s, offsets, (size_t)nb,
                     (const uint32_t*)tile_sums);
  if (lds_sort)
    lds_sort_pass(true, nwg_sort, (const uint32_t*)scalars, n, sh.c, sh.M, S_sort, P_sort, hist,
                  (const uint32_t*)offsets, sorted, s);
  else
    hipLaunchKernelGGL(k_msm_scatter, dim3(blocks_for(E_max, 256)), dim3(256), 0, s,
                       (const int32_t*)digits, n, sh.nw, sh.M, (const uint32_t*)offsets, cursor, sorted);
  pt.mark("sort");
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_accumulate<F>), dim3(blocks_for(tstar, ACC_BLOCK)), dim3(ACC_BLOCK), 0, s,
                     acc_points, stride, index, n_points, (const uint32_t*)sorted, (const uint32_t*)offsets, nb, tstar,
                     buckets, heads, tails, err);
  pt.mark("accumulate");
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_fixup_queue<F>), dim3(blocks_for(nb, QUEUE_BLOCK)), dim3(QUEUE_BLOCK), 0, s,
                     (const uint32_t*)offsets, nb, tstar, buckets, mid, nmid, big, nbig);
  {
    const unsigned gbig = (unsigned)std::min<size_t>(max_big, 2048);
    const unsigned gmid = (unsigned)std::min<size_t>(blocks_for(nb, 64), 4096);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_fixup<F>), dim3(gbig + gmid), dim3(64), 0, s, (const uint32_t*)offsets,
                       nb, tstar, buckets, (const jac<F>*)heads, (const jac<F>*)tails, (const uint32_t*)mid,
                       (const uint32_t*)nmid, (const uint32_t*)big, (const uint32_t*)nbig, gbig);
  }
  pt.mark("fixup");
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_segsum<F>), dim3(blocks_for(sh.nw * S, 64)), dim3(64), 0, s,
                     (const jac<F>*)buckets, sh.nw, sh.M, sh.logL, sh.logS, seg_acc, seg_run);
  pt.mark("segsum");
  constexpr int RT = MsmTraits<F>::REDUCE_THREADS;
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_window_sums<F, RT>), dim3(sh.nw * nsum), dim3(RT), 0, s,
                     (const jac<F>*)seg_acc, (const jac<F>*)seg_run, sh.logS, wsum);
  pt.mark("window_sums");
  BTE_CHECK_LAUNCH();
  if (sc.capturing()) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_horner<F>), dim3(1), dim3(64), 0, s, (const jac<F>*)wsum, sh.c, sh.nw,
                       sh.logL, sh.logS, out);
    BTE_CHECK_LAUNCH();
    return BTE_OK;
  }
  BTE_HIP(hipMemcpyAsync(job->hw, wsum, (size_t)sh.nw * nsum * sizeof(bteh::Jac<HF>), hipMemcpyDeviceToHost, s));
  BTE_HIP(hipMemcpyAsync(&job->err, err, sizeof(int), hipMemcpyDeviceToHost, s));
  BTE_HIP(hipLaunchHostFunc(s, horner_host<HF>, job));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_result_from_host<F>), dim3(1), dim3(64), 0, s,
                     (const jac<F>*)&job->result, out);
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  // Gathered points: an index out of range is reported by the call, which therefore waits for
  // its stream (outside graph capture).
  if (index && !sc.capturing()) {
    BTE_HIP(hipStreamSynchronize(s));
    if (job->err) return set_error(BTE_EINVAL, "msm: point index out of range");
  }
  return BTE_OK;
}


// Window width for a batch from its mean segment size (the bucket reduction costs ~2^c per
// window and segment, the accumulation one addition per term and window).
static int batched_window(size_t nterms, size_t nseg) {
  const size_t avg = nseg ? (nterms + nseg - 1) / nseg : nterms;
  int c = floor_log2(avg ? avg : 1) - 1;
  return std::max(3, std::min(12, c));
}

// status <- BTE_EINVAL when an index was out of range (the term was skipped), else BTE_OK.
__global__ void k_msm_status(const int* __restrict__ err, int* __restrict__ status) {
  if (threadIdx.x == 0) *status = *err ? BTE_EINVAL : BTE_OK;
}

// `status` (device int or null) receives BTE_EINVAL in stream order when some point index was
// out of range; the library's own callers build their index arrays on the host in range.
template <class F>
static int msm_batched_run(const aff<F>* points, size_t n_points, const uint32_t* index, const uint8_t* scalars,
                           const uint32_t* scalar_index, size_t nterms, const uint64_t* seg_offsets, size_t nseg, int c_force, jac<F>* out,
                           int* status, hipStream_t s) {
  if (nseg == 0) {
    if (status) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)status, (size_t)1, 0u);
    BTE_CHECK_LAUNCH();
    return BTE_OK;
  }
  if (!seg_offsets || !out || (nterms && (!points || !scalars)))
    return set_error(BTE_EINVAL, "msm_batched: null pointer");
  if (((uintptr_t)scalars & 3) != 0) return set_error(BTE_EINVAL, "msm_batched: scalars must be 4-byte aligned");
  if (nterms > 0x7fffffffu || nseg > 0x7fffffffu) return set_error(BTE_EINVAL, "msm_batched: too many terms");
  const int c = c_force > 0 ? c_force : batched_window(nterms, nseg);
  const int nw = (256 + c - 1) / c;
  const uint32_t M = 1u << (c - 1);
  const size_t nb64 = (size_t)nseg * nw * M;
  if (nb64 >= 0x7fffffffu) return set_error(BTE_EINVAL, "msm_batched: %zu buckets exceed the index range", nb64);
  const uint32_t nb = (uint32_t)nb64;
  const size_t E_max = nterms * (size_t)nw;
  const uint32_t tstar = acc_threads<F>(E_max);
  const int ntiles = (int)((nb + 1023) / 1024);
  const size_t max_big = tstar / FIX_SERIAL + 1;

  CallScope cv(s);
  uint32_t *keys, *counts, *offsets, *cursor, *sorted, *tile_sums, *big, *nbig, *mid, *nmid;
  jac<F> *buckets, *heads, *tails, *win, *seg_acc, *seg_run;
  int* err;
  // bucket reduction in two levels: runs of L = 2^logL buckets per thread, then per window
  const int logL = std::min(3, c - 1), logS = (c - 1) - logL;
  const size_t nsw = (size_t)nseg * nw, nss = nsw << logS;
  BTE_TRY_RC(cv.dev(&keys, E_max));
  BTE_TRY_RC(cv.dev(&counts, nb));
  BTE_TRY_RC(cv.dev(&offsets, (size_t)nb + 1));
  BTE_TRY_RC(cv.dev(&cursor, nb));
  BTE_TRY_RC(cv.dev(&sorted, E_max));
  BTE_TRY_RC(cv.dev(&tile_sums, ntiles));
  BTE_TRY_RC(cv.dev(&buckets, nb));
  BTE_TRY_RC(cv.dev(&heads, tstar));
  BTE_TRY_RC(cv.dev(&tails, tstar));
  BTE_TRY_RC(cv.dev(&win, nsw));
  BTE_TRY_RC(cv.dev(&seg_acc, nss));
  BTE_TRY_RC(cv.dev(&seg_run, nss));
  BTE_TRY_RC(cv.dev(&err, 1));
  BTE_TRY_RC(cv.dev(&big, max_big));
  BTE_TRY_RC(cv.dev(&nbig, 1));
  BTE_TRY_RC(cv.dev(&mid, nb));
  BTE_TRY_RC(cv.dev(&nmid, 1));

  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nb, 256)), dim3(256), 0, s, counts, (size_t)nb, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nb, 256)), dim3(256), 0, s, cursor, (size_t)nb, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)err, (size_t)1, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nbig, (size_t)1, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nmid, (size_t)1, 0u);
  if (nterms)
    hipLaunchKernelGGL(k_msm_digits_seg, dim3(blocks_for(nterms, 256)), dim3(256), 0, s, (const uint32_t*)scalars,
                       scalar_index, nterms, seg_offsets, (uint32_t)nseg, c, nw, M, keys, counts);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(256), 0, s, counts, (size_t)nb, offsets, tile_sums);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(256), 0, s, tile_sums, ntiles, offsets + nb);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(nb, 256)), dim3(256), 0, s, offsets, (size_t)nb,
                     (const uint32_t*)tile_sums);
  if (E_max) {
    hipLaunchKernelGGL(k_msm_scatter_keys, dim3(blocks_for(E_max, 256)), dim3(256), 0, s, (const uint32_t*)keys,
                       nterms, nw, (const uint32_t*)offsets, cursor, sorted);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_accumulate<F>), dim3(blocks_for(tstar, ACC_BLOCK)), dim3(ACC_BLOCK), 0,
                       s, points, (uint32_t)sizeof(aff<F>), index, index ? n_points : nterms, (const uint32_t*)sorted,
                       (const uint32_t*)offsets,
                       nb, tstar, buckets, heads, tails, err);
  }
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_fixup_queue<F>), dim3(blocks_for(nb, QUEUE_BLOCK)), dim3(QUEUE_BLOCK), 0, s,
                     (const uint32_t*)offsets, nb, tstar, buckets, mid, nmid, big, nbig);
  {
    const unsigned gbig = (unsigned)std::min<size_t>(max_big, 2048);
    const unsigned gmid = (unsigned)std::min<size_t>(blocks_for(nb, 64), 4096);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_fixup<F>), dim3(gbig + gmid), dim3(64), 0, s, (const uint32_t*)offsets,
                       nb, tstar, buckets, (const jac<F>*)heads, (const jac<F>*)tails, (const uint32_t*)mid,
                       (const uint32_t*)nmid, (const uint32_t*)big, (const uint32_t*)nbig, gbig);
  }
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_segsum<F>), dim3(blocks_for(nss, 64)), dim3(64), 0, s,
                     (const jac<F>*)buckets, (int)nsw, M, logL, logS, seg_acc, seg_run);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_seg_window<F>), dim3(blocks_for(nsw, 64)), dim3(64), 0, s,
                     (const jac<F>*)seg_acc, (const jac<F>*)seg_run, nsw, logL, logS, win);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_msm_combine<F>), dim3(blocks_for(nseg, 64)), dim3(64), 0, s,
                     (const jac<F>*)win, (uint32_t)nseg, nw, c, out);
  if (status) hipLaunchKernelGGL(k_msm_status, dim3(1), dim3(64), 0, s, (const int*)err, status);
  BTE_CHECK_LAUNCH();
  // The workspace is reused by the next call on this stream; stream order protects it.
  return BTE_OK;
}

static_assert(sizeof(bteh::Jac<bteh::Fp>) == sizeof(bte_p1), "layout");
static_assert(sizeof(bteh::Jac<bteh::Fp2>) == sizeof(bte_p2), "layout");
static_assert(sizeof(jac<fp>) == sizeof(bte_p1), "layout");
static_assert(sizeof(aff<fp2>) == sizeof(bte_p2_affine), "layout");

}  // namespace bte

using namespace bte;

extern "C" int bte_g1_msm(const bte_p1_affine* points, const uint8_t* scalars, size_t n, bte_p1* out,
                          void* stream) {
  return msm_run<fp>(0, (const aff<fp>*)points, n, nullptr, scalars, n, (jac<fp>*)out,
                     (hipStream_t)stream);
}

extern "C" int bte_g2_msm(const bte_p2_affine* points, const uint8_t* scalars, size_t n, bte_p2* out,
                          void* stream) {
  return msm_run<fp2>(1, (const aff<fp2>*)points, n, nullptr, scalars, n, (jac<fp2>*)out,
                      (hipStream_t)stream);
}

extern "C" int bte_g2_msm_indexed(const bte_p2_affine* points, size_t n_points, const uint32_t* indices,
                                  const uint8_t* scalars, size_t n, bte_p2* out, void* stream) {
  if (n && !indices) return set_error(BTE_EINVAL, "msm_indexed: null indices");
  return msm_run<fp2>(1, (const aff<fp2>*)points, n_points, indices, scalars, n, (jac<fp2>*)out,
                      (hipStream_t)stream);
}

extern "C" int bte_msm_set_window(int g1_or_g2, size_t n, int window_bits) {
  if ((g1_or_g2 != 1 && g1_or_g2 != 2) || window_bits < 0 || window_bits > 20 ||
      (window_bits > 0 && window_bits < 2))
    return set_error(BTE_EINVAL, "msm_set_window: bad arguments");
  std::lock_guard<std::mutex> lk(g_window_mu);
  g_window_override[g1_or_g2 - 1][floor_log2(n)] = window_bits;
  return BTE_OK;
}

extern "C" int bte_g1_msm_batched(const bte_p1_affine* points, size_t n_points, const uint32_t* indices,
                                  const uint8_t* scalars, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                                  bte_p1* out, int* status, void* stream) {
  return msm_batched_run<fp>((const aff<fp>*)points, n_points, indices, scalars, nullptr, nterms, seg_offsets, nseg, 0,
                             (jac<fp>*)out, status, (hipStream_t)stream);
}

extern "C" int bte_g2_msm_batched(const bte_p2_affine* points, size_t n_points, const uint32_t* indices,
                                  const uint8_t* scalars, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                                  bte_p2* out, int* status, void* stream) {
  return msm_batched_run<fp2>((const aff<fp2>*)points, n_points, indices, scalars, nullptr, nterms, seg_offsets, nseg, 0,
                              (jac<fp2>*)out, status, (hipStream_t)stream);
}

// Batched MSMs whose terms gather both the point (points[point_index[t]]) and the scalar
// (scalars[scalar_index[t]]); used by the scheme drivers to reuse one coefficient vector across
// many bases without replicating it.
namespace bte {
int g1_msm_batched_gather(const void* points, size_t n_points, const uint32_t* point_index, const uint8_t* scalars,
                          const uint32_t* scalar_index, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                          int c, void* out, hipStream_t s) {
  return msm_batched_run<fp>((const aff<fp>*)points, n_points, point_index, scalars, scalar_index, nterms,
                             seg_offsets, nseg, c, (jac<fp>*)out, nullptr, s);
}
int g2_msm_batched_gather(const void* points, size_t n_points, const uint32_t* point_index, const uint8_t* scalars,
                          const uint32_t* scalar_index, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                          int c, void* out, hipStream_t s) {
  return msm_batched_run<fp2>((const aff<fp2>*)points, n_points, point_index, scalars, scalar_index, nterms,
                              seg_offsets, nseg, c, (jac<fp2>*)out, nullptr, s);
}
}  // namespace bte
