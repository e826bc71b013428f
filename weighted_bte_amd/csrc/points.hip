// Point utilities: batch Jacobian -> affine normalisation (weighted_btx/src/blst_utils.rs:11-36,
// blst_p{1,2}s_to_affine) and fixed-base multiples of the standard generators.
#include <mutex>

#include "common.hpp"
#include "ctx.hpp"
#include "curve.hpp"

namespace bte {

static constexpr int NORM_CHUNK = 16;

// Montgomery's simultaneous inversion over chunks of NORM_CHUNK points, one chunk per thread.
// out[i].x holds the running prefix product until point i is finished.  Identities (z = 0)
// become the all-zero affine point and are skipped in the product.
template <class F>
__global__ void __launch_bounds__(64) k_batch_normalize(const jac<F>* __restrict__ in, size_t n, aff<F>* __restrict__ out) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t start = t * NORM_CHUNK;
  if (start >= n) return;
  const size_t stop = start + NORM_CHUNK < n ? start + NORM_CHUNK : n;
  F acc = f_one<F>();
  for (size_t i = start; i < stop; ++i) {
    F z = in[i].z;
    if (!f_is_zero(z)) {
      out[i].x = acc;
      acc = f_mul(acc, z);
    }
  }
  F inv = f_inv(acc);
  for (size_t i = stop; i-- > start;) {
    jac<F> p = in[i];
    if (f_is_zero(p.z)) {
      out[i] = {f_zero<F>(), f_zero<F>()};
      continue;
    }
    F zi = f_mul(inv, out[i].x);
    inv = f_mul(inv, p.z);
    F zi2 = f_sqr(zi);
    out[i] = {f_mul(p.x, zi2), f_mul(p.y, f_mul(zi2, zi))};
  }
}

template <class F> static int batch_normalize(const jac<F>* in, size_t n, aff<F>* out, hipStream_t s) {
  if (n == 0) return BTE_OK;
  if (!in || !out) return set_error(BTE_EINVAL, "batch_normalize: null pointer");
  size_t chunks = (n + NORM_CHUNK - 1) / NORM_CHUNK;
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_batch_normalize<F>), dim3(blocks_for(chunks, 64)), dim3(64), 0, s,
                     in, n, out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

BTE_DEV g1a g1_generator() {
  g1a g;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    g.x.l[j] = G1_GEN[0][j];
    g.y.l[j] = G1_GEN[1][j];
  }
  return g;
}

BTE_DEV g2a g2_generator() {
  g2a g;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    g.x.c0.l[j] = G2_GEN[0][j];
    g.x.c1.l[j] = G2_GEN[1][j];
    g.y.c0.l[j] = G2_GEN[2][j];
    g.y.c1.l[j] = G2_GEN[3][j];
  }
  return g;
}

template <class F> BTE_DEV aff<F> generator();
template <> BTE_DEV g1a generator<fp>() { return g1_generator(); }
template <> BTE_DEV g2a generator<fp2>() { return g2_generator(); }

// Byte-window table: table[256 w + j] = j 2^(8w) G for w < 32, j < 256 (entry 0 of each window is
// the identity).  One thread per window builds its 256 multiples; run once per device.
static constexpr int GEN_WINDOWS = 32, GEN_ENTRIES = 256;

template <class F> __global__ void k_generator_table(jac<F>* table) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= GEN_WINDOWS) return;
  jac<F> base = jac_from_aff(generator<F>());
  for (int i = 0; i < 8 * w; ++i) base = jac_dbl(base);
  jac<F> acc = jac_identity<F>();
  for (int j = 0; j < GEN_ENTRIES; ++j) {
    table[w * GEN_ENTRIES + j] = acc;
    acc = jac_add(acc, base);
  }
}

// out[i] = k_i * G as the sum of one table entry per nonzero byte of k_i (full 256 bits).
template <class F>
__global__ void __launch_bounds__(64) k_fixed_base_mul(const aff<F>* __restrict__ table, const uint32_t* __restrict__ scalars,
                                 size_t n, jac<F>* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  jac<F> acc = jac_identity<F>();
  for (int word = 0; word < 8; ++word) {
    const uint32_t k = scalars[8 * i + word];
#pragma unroll 1
    for (int b = 0; b < 4; ++b) {
      const uint32_t byte = (k >> (8 * b)) & 0xffu;
      if (byte) acc = jac_add_aff(acc, table[(4 * word + b) * GEN_ENTRIES + byte]);
    }
  }
  out[i] = acc;
}

template <class F> struct GenTable {
  std::mutex mu;
  aff<F>* table[64] = {};
};

template <class F> static int generator_table(const aff<F>** out, hipStream_t s) {
  static GenTable<F> g;
  int dev = 0;
  BTE_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_error(BTE_EDEVICE, "generator_table: device id %d", dev);
  std::lock_guard<std::mutex> lk(g.mu);
  if (!g.table[dev]) {
    jac<F>* tj = nullptr;
    aff<F>* ta = nullptr;
    const size_t nt = (size_t)GEN_WINDOWS * GEN_ENTRIES;
    BTE_HIP(hipMalloc((void**)&tj, nt * sizeof(jac<F>)));
    BTE_HIP(hipMalloc((void**)&ta, nt * sizeof(aff<F>)));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_generator_table<F>), dim3(1), dim3(64), 0, s, tj);
    int rc = batch_normalize<F>(tj, nt, ta, s);
    if (rc) return rc;
    BTE_HIP(hipStreamSynchronize(s));
    BTE_HIP(hipFree(tj));
    g.table[dev] = ta;
  }
  *out = g.table[dev];
  return BTE_OK;
}

template <class F> static int generator_mul(const uint8_t* scalars, size_t n, aff<F>* out, hipStream_t s) {
  if (n == 0) return BTE_OK;
  if (!scalars || !out) return set_error(BTE_EINVAL, "generator_mul: null pointer");
  if (((uintptr_t)scalars & 3) != 0) return set_error(BTE_EINVAL, "generator_mul: scalars must be 4-byte aligned");
  const aff<F>* table;
  int rc = generator_table<F>(&table, s);
  if (rc) return rc;
  CallScope sc(s);
  jac<F>* tmp;
  BTE_TRY_RC(sc.dev(&tmp, n));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fixed_base_mul<F>), dim3(blocks_for(n, 64)), dim3(64), 0, s, table,
                     (const uint32_t*)scalars, n, tmp);
  rc = batch_normalize<F>(tmp, n, out, s);
  if (rc) return rc;
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

}  // namespace bte

using namespace bte;

extern "C" int bte_g1_batch_normalize(const bte_p1* in, size_t n, bte_p1_affine* out, void* stream) {
  return batch_normalize<fp>((const jac<fp>*)in, n, (aff<fp>*)out, (hipStream_t)stream);
}

extern "C" int bte_g2_batch_normalize(const bte_p2* in, size_t n, bte_p2_affine* out, void* stream) {
  return batch_normalize<fp2>((const jac<fp2>*)in, n, (aff<fp2>*)out, (hipStream_t)stream);
}

extern "C" int bte_g1_generator_mul(const uint8_t* scalars, size_t n, bte_p1_affine* out, void* stream) {
  return generator_mul<fp>(scalars, n, (aff<fp>*)out, (hipStream_t)stream);
}

extern "C" int bte_g2_generator_mul(const uint8_t* scalars, size_t n, bte_p2_affine* out, void* stream) {
  return generator_mul<fp2>(scalars, n, (aff<fp2>*)out, (hipStream_t)stream);
}
