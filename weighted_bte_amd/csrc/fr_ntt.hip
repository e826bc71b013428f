// Radix-2 transforms over the scalar field Fr, and the Lagrange-at-zero coefficients of
// weighted_btx/src/interpolation.rs computed on the device with a product tree of such
// transforms.
//
// Transform (weighted_btx/src/fft.rs:197-234 Radix2Domain::{fft, ifft} over Scalar): the
// natural-order DFT X_k = sum_j x_j w^(jk), w = ROOT_OF_UNITY^(2^(32 - log n)), w^-1 for the
// inverse, n^-1 for the scaled inverse; bit reversal then decimation-in-time stages.  Values are
// held in Montgomery form (R = 2^256) inside; the C ABI takes canonical 32-byte scalars.  Stages
// below 2^11 run in LDS, one workgroup per block of 2^min(log n, 11) consecutive elements (one
// launch for all of them); the remaining stages are one launch each, one thread per butterfly.
//
// Lagrange at zero (interpolation.rs:31-140): for the selected points x_i = w_M^(v_i) of the
// radix-2 domain of size M = nextpow2(W),
//     L_i(0) = -P(0) / (x_i P'(x_i)),   P(X) = prod_i (X - x_i),
// with P from a product tree (products of low degree as plain convolutions, the rest through
// batched transforms, every pair of a tree level in one launch sequence) and P' evaluated on the
// whole domain by one transform of size M.  Small sets use the direct formula on the host and a
// complete power-of-two domain gives 1/M, as in the reference.  The coefficients are unique field
// values, so any correct method returns the same bytes.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "fr.hpp"
#include "host_field.hpp"

namespace bte {

BTE_DEV fr fr_neg(const fr& a) {
  fr z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z.l[i] = 0;
  return fr_sub(z, a);
}

BTE_DEV fr fr_mont_small(uint32_t v) {
  fr a;
#pragma unroll
  for (int i = 0; i < 8; ++i) a.l[i] = 0;
  a.l[0] = v;
  return fr_to_mont(a);
}

// a^(r-2) (Montgomery in and out); zero for zero.
BTE_DEV fr fr_mont_inv(const fr& a) {
  constexpr uint32_t E[8] = {0xffffffffu, 0xfffffffeu, 0xfffe5bfeu, 0x53bda402u,
                             0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
  fr r = fr_mont_small(1);
  for (int i = 7; i >= 0; --i)
    for (int b = 31; b >= 0; --b) {
      r = fr_mont_mul(r, r);
      if ((E[i] >> b) & 1u) r = fr_mont_mul(r, a);
    }
  return r;
}

// ---------------------------------------------------------------------------------------------
// Domain tables: the n/2 twiddles w^j and n^-1, Montgomery form, per (device, log n, direction).
// ---------------------------------------------------------------------------------------------
struct FrDomain {
  fr* tw = nullptr;
  fr* ninv = nullptr;
};
static std::mutex g_frdom_mu;
static std::map<std::tuple<int, int, int>, FrDomain> g_frdom;

static bteh::Fr host_to_mont(const bteh::Fr& x) {
  static const bteh::Fr R256 = bteh::fr_pow_u64(bteh::fr_from_u64(2), 256);  // 2^256 mod r
  return bteh::fr_mul(x, R256);
}

static int fr_domain(int logn, bool inverse, const FrDomain** out) {
  int dev = 0;
  BTE_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_frdom_mu);
  FrDomain& t = g_frdom[std::make_tuple(dev, logn, inverse ? 1 : 0)];
  if (!t.tw) {
    const size_t half = logn ? ((size_t)1 << (logn - 1)) : 1;
    bteh::Fr w;
    memcpy(w.l, bteh::FR_ROOT_OF_UNITY, 32);
    for (int i = logn; i < bteh::FR_S; ++i) w = bteh::fr_mul(w, w);
    if (inverse) w = bteh::fr_inv(w);
    std::vector<bteh::Fr> tw(half);
    bteh::Fr cur = bteh::fr_from_u64(1);
    for (size_t j = 0; j < half; ++j) {
      tw[j] = host_to_mont(cur);
      cur = bteh::fr_mul(cur, w);
    }
    const bteh::Fr ninv = host_to_mont(bteh::fr_inv(bteh::fr_from_u64((uint64_t)1 << logn)));
    fr *d_tw = nullptr, *d_ninv = nullptr;
    hipError_t e = test_alloc_should_fail() ? hipErrorOutOfMemory : hipMalloc((void**)&d_tw, half * sizeof(fr));
    if (e == hipSuccess) e = hipMemcpy(d_tw, tw.data(), half * sizeof(fr), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&d_ninv, sizeof(fr));
    if (e == hipSuccess) e = hipMemcpy(d_ninv, &ninv, sizeof(fr), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (d_tw) (void)hipFree(d_tw);
      if (d_ninv) (void)hipFree(d_ninv);
      return set_error(BTE_ENOMEM, "fr domain tables (2^%d): %s", logn, hipGetErrorString(e));
    }
    t.ninv = d_ninv;
    t.tw = d_tw;  // published last
  }
  *out = &t;
  return BTE_OK;
}

// ---------------------------------------------------------------------------------------------
// Transform kernels (count transforms of size n = 2^logn stored back to back)
// ---------------------------------------------------------------------------------------------
__global__ void k_fr_convert(fr* v, size_t n, int to_mont) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = to_mont ? fr_to_mont(v[i]) : fr_from_mont(v[i]);
}

__global__ void k_fr_bitrev(fr* v, size_t n, int logn, size_t count) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= n * count) return;
  fr* b = v + (g >> logn) * n;
  const size_t i = g & (n - 1);
  const size_t r = logn ? (size_t)(__builtin_bitreverse64((uint64_t)i) >> (64 - logn)) : 0;
  if (i < r) {
    const fr a = b[i];
    b[i] = b[r];
    b[r] = a;
  }
}

BTE_DEV void fr_butterfly(fr& u, fr& x, const fr& w) {
  const fr t = fr_mont_mul(x, w);
  x = fr_sub(u, t);
  u = fr_add(u, t);
}

// Stages 0 .. s1-1 of every block of 2^s1 consecutive elements, in LDS (s1 <= FR_LDS_LOG).
static constexpr int FR_LDS_LOG = 11;
static constexpr int FR_LDS_THREADS = 256;
__global__ void __launch_bounds__(FR_LDS_THREADS) k_fr_ntt_lds(fr* __restrict__ v, int logn, int s1,
                                                               const fr* __restrict__ tw) {
  __shared__ fr sh[1 << FR_LDS_LOG];
  const int bs = 1 << s1;
  fr* base = v + (size_t)blockIdx.x * bs;
  for (int i = threadIdx.x; i < bs; i += FR_LDS_THREADS) sh[i] = base[i];
  __syncthreads();
  for (int s = 0; s < s1; ++s) {
    for (int b = threadIdx.x; b < bs / 2; b += FR_LDS_THREADS) {
      const int k = b & ((1 << s) - 1);
      const int i0 = ((b >> s) << (s + 1)) + k, i1 = i0 + (1 << s);
      fr_butterfly(sh[i0], sh[i1], tw[(size_t)k << (logn - 1 - s)]);
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < bs; i += FR_LDS_THREADS) base[i] = sh[i];
}

// Stage s (butterfly half-size 2^s) over all transforms, one thread per butterfly.
__global__ void __launch_bounds__(256) k_fr_stage(fr* __restrict__ v, int logn, int s, size_t total,
                                                  const fr* __restrict__ tw) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const size_t half_n = (size_t)1 << (logn - 1);
  const size_t t = g >> (logn - 1), b = g & (half_n - 1);
  const size_t k = b & (((size_t)1 << s) - 1);
  const size_t i0 = (t << logn) + ((b >> s) << (s + 1)) + k, i1 = i0 + ((size_t)1 << s);
  fr u = v[i0], x = v[i1];
  fr_butterfly(u, x, tw[k << (logn - 1 - s)]);
  v[i0] = u;
  v[i1] = x;
}

__global__ void k_fr_scale(fr* v, size_t n, const fr* __restrict__ f) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = fr_mont_mul(v[i], *f);
}

// count transforms of size n (Montgomery values), in place, on stream s.
static int fr_ntt(fr* v, size_t n, size_t count, bool inverse, bool scaled, hipStream_t s) {
  if (n == 0 || (n & (n - 1))) return set_error(BTE_EINVAL, "fft_fr: InvalidBatchSize(%zu)", n);
  int logn = 0;
  while (((size_t)1 << logn) < n) ++logn;
  if (logn > bteh::FR_S) return set_error(BTE_EINVAL, "fft_fr: InvalidBatchSize(%zu)", n);
  if (count == 0 || n == 1) return BTE_OK;
  const FrDomain* d;
  BTE_TRY_RC(fr_domain(logn, inverse, &d));
  const size_t total = n * count;
  hipLaunchKernelGGL(k_fr_bitrev, dim3(blocks_for(total, 256)), dim3(256), 0, s, v, n, logn, count);
  const int s1 = std::min(logn, FR_LDS_LOG);
  hipLaunchKernelGGL(k_fr_ntt_lds, dim3((unsigned)(total >> s1)), dim3(FR_LDS_THREADS), 0, s, v, logn, s1,
                     (const fr*)d->tw);
  for (int st = s1; st < logn; ++st)
    hipLaunchKernelGGL(k_fr_stage, dim3(blocks_for(total / 2, 256)), dim3(256), 0, s, v, logn, st, total / 2,
                       (const fr*)d->tw);
  if (inverse && scaled)
    hipLaunchKernelGGL(k_fr_scale, dim3(blocks_for(total, 256)), dim3(256), 0, s, v, total, (const fr*)d->ninv);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// ---------------------------------------------------------------------------------------------
// Product tree and Lagrange coefficients
// ---------------------------------------------------------------------------------------------
// Level 0: node i = X - x_i with x_i = w^(sel_i) read from the forward table of the size-M
// domain (w^j = -w^(j - M/2) for j >= M/2).  Coefficients low degree first.
BTE_DEV fr domain_point(const fr* __restrict__ tw, size_t half_m, uint32_t e) {
  return e < half_m ? tw[e] : fr_neg(tw[e - half_m]);
}

__global__ void k_tree_leaves(const uint32_t* __restrict__ sel, size_t n, const fr* __restrict__ tw, size_t half_m,
                              fr* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = fr_neg(domain_point(tw, half_m, sel[i]));
  out[2 * i + 1] = fr_mont_small(1);
}

// One tree level by plain convolution: out node p = in node 2p * in node 2p+1 (or a copy of the
// last node when the level has an odd count).  Node k of a level holds off[k+1] - off[k]
// coefficients at off[k].  One thread per output coefficient.
__global__ void __launch_bounds__(256) k_tree_level_naive(const fr* __restrict__ in, const uint64_t* __restrict__ in_off,
                                                          uint32_t nin, fr* __restrict__ out,
                                                          const uint64_t* __restrict__ out_off, uint32_t nout) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= out_off[nout]) return;
  const uint32_t p = seg_of(out_off, nout, g);
  const uint64_t j = g - out_off[p];
  const fr* a = in + in_off[2 * p];
  if (2 * p + 1 >= nin) {
    out[g] = a[j];
    return;
  }
  const int64_t da = (int64_t)(in_off[2 * p + 1] - in_off[2 * p]) - 1;
  const fr* b = in + in_off[2 * p + 1];
  const int64_t db = (int64_t)(in_off[2 * p + 2] - in_off[2 * p + 1]) - 1;
  const int64_t lo = (int64_t)j - db > 0 ? (int64_t)j - db : 0, hi = (int64_t)j < da ? (int64_t)j : da;
  fr acc;
#pragma unroll
  for (int k = 0; k < 8; ++k) acc.l[k] = 0;
  for (int64_t i = lo; i <= hi; ++i) acc = fr_add(acc, fr_mont_mul(a[i], b[j - i]));
  out[g] = acc;
}

// Transform path of a level: buf[(2p + side) T + j] = coefficient j of in node 2p + side, zero
// beyond its degree (npairs = nin / 2 pairs).
__global__ void k_tree_pack(const fr* __restrict__ in, const uint64_t* __restrict__ in_off, uint32_t npairs, size_t T,
                            fr* __restrict__ buf) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= 2 * (size_t)npairs * T) return;
  const size_t node = g / T, j = g - node * T;
  const size_t len = in_off[node + 1] - in_off[node];
  if (j < len) {
    buf[g] = in[in_off[node] + j];
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) buf[g].l[k] = 0;
  }
}

__global__ void k_tree_pointwise(const fr* __restrict__ buf, uint32_t npairs, size_t T, fr* __restrict__ prod) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= (size_t)npairs * T) return;
  const size_t p = g / T, j = g - p * T;
  prod[g] = fr_mont_mul(buf[2 * p * T + j], buf[(2 * p + 1) * T + j]);
}

__global__ void k_tree_unpack(const fr* __restrict__ prod, size_t T, uint32_t npairs, const fr* __restrict__ in,
                              const uint64_t* __restrict__ in_off, fr* __restrict__ out,
                              const uint64_t* __restrict__ out_off, uint32_t nout) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= out_off[nout]) return;
  const uint32_t p = seg_of(out_off, nout, g);
  const uint64_t j = g - out_off[p];
  out[g] = p < npairs ? prod[(size_t)p * T + j] : in[in_off[2 * p] + j];
}

// der[j] = (j + 1) P[j + 1] for j < deg P, zero up to M.
__global__ void k_derivative(const fr* __restrict__ P, size_t deg, size_t M, fr* __restrict__ der) {
  const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (j >= M) return;
  if (j < deg) {
    der[j] = fr_mont_mul(fr_mont_small((uint32_t)(j + 1)), P[j + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) der[j].l[k] = 0;
  }
}

// out_i = -P(0) / (x_i P'(x_i)) (canonical), scaled by *scale (Montgomery).  The selection is
// validated on the host (distinct points of the domain), so x_i P'(x_i) = x_i prod_{j != i}
// (x_i - x_j) is never zero.
__global__ void __launch_bounds__(64) k_lagrange_out(const uint32_t* __restrict__ sel, size_t n, const fr* __restrict__ tw,
                                                     size_t half_m, const fr* __restrict__ der_eval,
                                                     const fr* __restrict__ P, const fr* __restrict__ scale,
                                                     fr* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t e = sel[i];
  const fr den = fr_mont_mul(domain_point(tw, half_m, e), der_eval[e]);
  const fr v = fr_mont_mul(fr_mont_mul(fr_neg(P[0]), fr_mont_inv(den)), *scale);
  out[i] = fr_from_mont(v);
}

__global__ void k_fr_fill_canonical(fr* out, size_t n, fr value) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = value;
}

static constexpr size_t LAGRANGE_DIRECT_MAX = 64;  // interpolation.rs:20 DIRECT_INTERPOLATION_CUTOFF
static constexpr size_t NAIVE_PRODUCT_MAX = 32;    // interpolation.rs:24 NAIVE_MULTIPLICATION_CUTOFF

// Validation of interpolation.rs:55-67 (validate_selection): nonempty, in range, no repeats.
static int validate_selection(size_t W, const uint64_t* sel, size_t n) {
  if (W == 0 || n == 0 || !sel) return set_error(BTE_EINVAL, "InvalidInterpolationSet");
  std::vector<char> seen(W, 0);
  for (size_t i = 0; i < n; ++i) {
    if (sel[i] >= W || seen[sel[i]]) return set_error(BTE_EINVAL, "InvalidInterpolationSet");
    seen[sel[i]] = 1;
  }
  return BTE_OK;
}

// interpolation.rs:69-99 direct formula on the host (few points): L_i = prod_{j != i} x_j / (x_j - x_i).
static void lagrange_direct_host(int logm, const uint64_t* sel, size_t n, std::vector<bteh::Fr>& out) {
  bteh::Fr w;
  memcpy(w.l, bteh::FR_ROOT_OF_UNITY, 32);
  for (int i = logm; i < bteh::FR_S; ++i) w = bteh::fr_mul(w, w);
  std::vector<bteh::Fr> x(n);
  for (size_t i = 0; i < n; ++i) x[i] = bteh::fr_pow_u64(w, sel[i]);
  out.resize(n);
  for (size_t i = 0; i < n; ++i) {
    bteh::Fr num = bteh::fr_from_u64(1), den = bteh::fr_from_u64(1);
    for (size_t j = 0; j < n; ++j)
      if (j != i) {
        num = bteh::fr_mul(num, x[j]);
        den = bteh::fr_mul(den, bteh::fr_sub(x[j], x[i]));
      }
    out[i] = bteh::fr_mul(num, bteh::fr_inv(den));
  }
}

int lagrange_at_zero_device(size_t W, const uint64_t* sel, size_t n, const uint8_t* scale_canonical,
                            uint8_t* d_out, hipStream_t s) {
  BTE_TRY_RC(validate_selection(W, sel, n));
  int logm = 0;
  while (((size_t)1 << logm) < W) ++logm;
  const size_t M = (size_t)1 << logm;
  bteh::Fr scale = bteh::fr_from_u64(1);
  if (scale_canonical) memcpy(scale.l, scale_canonical, 32);
  if (n == W && M == W) {  // interpolation.rs:36-44: every point of a complete domain, 1/M each
    const bteh::Fr v = bteh::fr_mul(bteh::fr_inv(bteh::fr_from_u64(M)), scale);
    fr dv;
    memcpy(dv.l, v.l, 32);
    hipLaunchKernelGGL(k_fr_fill_canonical, dim3(blocks_for(n, 256)), dim3(256), 0, s, (fr*)d_out, n, dv);
    BTE_CHECK_LAUNCH();
    return BTE_OK;
  }
  if (n <= LAGRANGE_DIRECT_MAX) {
    std::vector<bteh::Fr> c;
    lagrange_direct_host(logm, sel, n, c);
    for (auto& v : c) v = bteh::fr_mul(v, scale);
    CallScope sc(s);
    void* stage;
    BTE_TRY_RC(sc.host_bytes(&stage, 32 * n));
    memcpy(stage, c.data(), 32 * n);
    BTE_HIP(hipMemcpyAsync(d_out, stage, 32 * n, hipMemcpyHostToDevice, s));
    return BTE_OK;
  }
  // Tree levels: offsets of every node of every level (host), uploaded once.
  std::vector<std::vector<uint64_t>> offs;
  {
    std::vector<uint64_t> o(n + 1);
    for (size_t i = 0; i <= n; ++i) o[i] = 2 * i;
    offs.push_back(o);
    while (offs.back().size() > 2) {
      const std::vector<uint64_t>& in = offs.back();
      const size_t nin = in.size() - 1, nout = (nin + 1) / 2;
      std::vector<uint64_t> o2(nout + 1, 0);
      for (size_t p = 0; p < nout; ++p) {
        const uint64_t la = in[2 * p + 1] - in[2 * p];
        const uint64_t len = 2 * p + 1 < nin ? la + (in[2 * p + 2] - in[2 * p + 1]) - 1 : la;
        o2[p + 1] = o2[p] + len;
      }
      offs.push_back(o2);
    }
  }
  std::vector<uint64_t> flat;
  std::vector<size_t> at;
  for (auto& o : offs) {
    at.push_back(flat.size());
    flat.insert(flat.end(), o.begin(), o.end());
  }
  std::vector<uint32_t> sel32(n);
  for (size_t i = 0; i < n; ++i) sel32[i] = (uint32_t)sel[i];
  size_t maxT = 1, maxcoef = 2 * n;
  for (size_t l = 0; l + 1 < offs.size(); ++l) {
    const std::vector<uint64_t>& in = offs[l];
    const size_t nin = in.size() - 1;
    size_t T = 1;
    for (size_t p = 0; 2 * p + 1 < nin; ++p) {
      const size_t need = (in[2 * p + 1] - in[2 * p]) + (in[2 * p + 2] - in[2 * p + 1]) - 1;
      while (T < need) T <<= 1;
    }
    if ((in[1] - in[0]) - 1 > NAIVE_PRODUCT_MAX) maxT = std::max(maxT, T * nin);
    maxcoef = std::max<size_t>(maxcoef, offs[l + 1].back());
  }
  const FrDomain* dom;
  BTE_TRY_RC(fr_domain(logm, false, &dom));
  CallScope sc(s);
  uint64_t* d_offs;
  uint32_t* d_sel;
  fr *lv[2], *buf, *prod, *der, *d_scale;
  const bteh::Fr scale_m = host_to_mont(scale);
  BTE_TRY_RC(sc.upload(&d_offs, flat.data(), flat.size()));
  BTE_TRY_RC(sc.upload(&d_sel, sel32.data(), n));
  BTE_TRY_RC(sc.upload(&d_scale, (const fr*)&scale_m, 1));
  BTE_TRY_RC(sc.dev(&lv[0], maxcoef));
  BTE_TRY_RC(sc.dev(&lv[1], maxcoef));
  BTE_TRY_RC(sc.dev(&buf, maxT));
  BTE_TRY_RC(sc.dev(&prod, maxT / 2 + 1));
  BTE_TRY_RC(sc.dev(&der, M));
  const size_t half_m = M / 2;
  hipLaunchKernelGGL(k_tree_leaves, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const uint32_t*)d_sel, n,
                     (const fr*)dom->tw, half_m, lv[0]);
  int cur = 0;
  for (size_t l = 0; l + 1 < offs.size(); ++l) {
    const std::vector<uint64_t>& in = offs[l];
    const uint32_t nin = (uint32_t)(in.size() - 1), nout = (uint32_t)(offs[l + 1].size() - 1);
    const uint64_t* din = d_offs + at[l];
    const uint64_t* dout = d_offs + at[l + 1];
    const size_t total_out = offs[l + 1].back();
    fr *src = lv[cur], *dst = lv[cur ^ 1];
    if ((in[1] - in[0]) - 1 <= NAIVE_PRODUCT_MAX) {
      hipLaunchKernelGGL(k_tree_level_naive, dim3(blocks_for(total_out, 256)), dim3(256), 0, s, (const fr*)src, din,
                         nin, dst, dout, nout);
    } else {
      const uint32_t npairs = nin / 2;
      size_t T = 1;
      for (size_t p = 0; p < npairs; ++p) {
        const size_t need = (in[2 * p + 1] - in[2 * p]) + (in[2 * p + 2] - in[2 * p + 1]) - 1;
        while (T < need) T <<= 1;
      }
      hipLaunchKernelGGL(k_tree_pack, dim3(blocks_for(2 * npairs * T, 256)), dim3(256), 0, s, (const fr*)src, din,
                         npairs, T, buf);
      BTE_TRY_RC(fr_ntt(buf, T, 2 * (size_t)npairs, false, false, s));
      hipLaunchKernelGGL(k_tree_pointwise, dim3(blocks_for(npairs * T, 256)), dim3(256), 0, s, (const fr*)buf, npairs,
                         T, prod);
      BTE_TRY_RC(fr_ntt(prod, T, npairs, true, true, s));
      hipLaunchKernelGGL(k_tree_unpack, dim3(blocks_for(total_out, 256)), dim3(256), 0, s, (const fr*)prod, T, npairs,
                         (const fr*)src, din, dst, dout, nout);
    }
    cur ^= 1;
  }
  const fr* P = lv[cur];  // degree n, n + 1 coefficients
  hipLaunchKernelGGL(k_derivative, dim3(blocks_for(M, 256)), dim3(256), 0, s, P, n, M, der);
  BTE_TRY_RC(fr_ntt(der, M, 1, false, false, s));
  hipLaunchKernelGGL(k_lagrange_out, dim3(blocks_for(n, 64)), dim3(64), 0, s, (const uint32_t*)d_sel, n,
                     (const fr*)dom->tw, half_m, (const fr*)der, P, (const fr*)d_scale, (fr*)d_out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

}  // namespace bte

using namespace bte;

static int fft_fr_entry(uint8_t* values, size_t n, size_t count, int inverse, int scaled, hipStream_t s) {
  if (count == 0) return BTE_OK;
  if (!values) return set_error(BTE_EINVAL, "fft_fr: null pointer");
  if (n == 0 || (n & (n - 1))) return set_error(BTE_EINVAL, "fft_fr: InvalidBatchSize(%zu)", n);
  const size_t total = n * count;
  hipLaunchKernelGGL(k_fr_convert, dim3(blocks_for(total, 256)), dim3(256), 0, s, (fr*)values, total, 1);
  BTE_TRY_RC(fr_ntt((fr*)values, n, count, inverse != 0, scaled != 0, s));
  hipLaunchKernelGGL(k_fr_convert, dim3(blocks_for(total, 256)), dim3(256), 0, s, (fr*)values, total, 0);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_fft_fr(uint8_t* values, size_t n, int inverse, int scaled, void* stream) {
  return fft_fr_entry(values, n, 1, inverse, scaled, (hipStream_t)stream);
}

extern "C" int bte_fft_fr_batched(uint8_t* values, size_t n, size_t count, int inverse, int scaled, void* stream) {
  return fft_fr_entry(values, n, count, inverse, scaled, (hipStream_t)stream);
}

// interpolation.rs:31-52 lagrange_at_zero with host output.  Complete power-of-two domains and
// sets of up to 64 points are computed on the host; larger sets take the device product tree on
// the null stream (the host interface has no stream), which this call waits for.
extern "C" int bte_lagrange_at_zero(size_t total_weight, const uint64_t* selected, size_t n, uint8_t* out) {
  if (!selected || !out) return set_error(BTE_EINVAL, "lagrange_at_zero: null pointer");
  BTE_TRY_RC(validate_selection(total_weight, selected, n));
  int logm = 0;
  while (((size_t)1 << logm) < total_weight) ++logm;
  if (n == total_weight && ((size_t)1 << logm) == total_weight) {
    const bteh::Fr v = bteh::fr_inv(bteh::fr_from_u64(total_weight));
    for (size_t i = 0; i < n; ++i) memcpy(out + 32 * i, v.l, 32);
    return BTE_OK;
  }
  if (n <= LAGRANGE_DIRECT_MAX) {
    std::vector<bteh::Fr> c;
    lagrange_direct_host(logm, selected, n, c);
    memcpy(out, c.data(), 32 * n);
    return BTE_OK;
  }
  uint8_t* d = nullptr;
  BTE_HIP(hipMalloc((void**)&d, 32 * n));
  int rc = lagrange_at_zero_device(total_weight, selected, n, nullptr, d, nullptr);
  if (rc == BTE_OK) {
    const hipError_t e = hipMemcpy(out, d, 32 * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_error(BTE_EDEVICE, "lagrange_at_zero: %s", hipGetErrorString(e));
  }
  (void)hipFree(d);
  return rc;
}

extern "C" int bte_lagrange_at_zero_device(size_t total_weight, const uint64_t* selected, size_t n, uint8_t* out,
                                           void* stream) {
  if (!out) return set_error(BTE_EINVAL, "lagrange_at_zero: null pointer");
  return lagrange_at_zero_device(total_weight, selected, n, nullptr, out, (hipStream_t)stream);
}
