// Canonical compressed encodings of G1 and G2 points on the GPU (the ZCash / blst format that
// blstrs' to_compressed / from_compressed use, e.g. weighted_pfe/src/proof.rs:62-65 and the size
// KATs of weighted_pfe/src/encoding.rs:31-45): x big-endian (G2: x.c1 then x.c0) with flag bits
// 0x80 compressed, 0x40 infinity, 0x20 "y is the larger of y, -y" (G2: compared on y.c1, or on
// y.c0 when y.c1 = 0).  Decompression checks the flags, x < p, that x is on the curve and that
// the point lies in the order-r subgroup (r P = O), and reports each element's validity.
// Square roots: y = a^((p+1)/4) in Fp (p = 3 mod 4); in Fp2 the method for q = 3 mod 4 of Adj
// and Rodriguez-Henriquez (a^((q-3)/4), then (1 + a^((q-1)/2))^((q-1)/2) or the factor i).
#include "common.hpp"
#include "curve.hpp"

namespace bte {

BTE_DEV fp fp_to_mont(const fp& c) {
  constexpr uint32_t R2[12] = {BTE_R2_LIMBS};
  fp r2;
#pragma unroll
  for (int i = 0; i < 12; ++i) r2.l[i] = R2[i];
  return fp_mul(c, r2);
}

// Big-endian 48 bytes <-> canonical limbs.
BTE_DEV fp fp_from_be(const uint8_t* b, bool& ok) {
  fp c;
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    const uint8_t* q = b + 44 - 4 * i;
    c.l[i] = ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3];
  }
  uint32_t borrow = 0;  // canonical: c < p
#pragma unroll
  for (int i = 0; i < 12; ++i) (void)__builtin_subc(c.l[i], p_limb(i), borrow, &borrow);
  ok = ok && borrow;
  return c;
}

BTE_DEV void fp_to_be(const fp& c, uint8_t* b) {
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    uint8_t* q = b + 44 - 4 * i;
    q[0] = (uint8_t)(c.l[i] >> 24);
    q[1] = (uint8_t)(c.l[i] >> 16);
    q[2] = (uint8_t)(c.l[i] >> 8);
    q[3] = (uint8_t)c.l[i];
  }
}

BTE_DEV fp2 fp2_pow_limbs(const fp2& a, const uint32_t* e) {
  fp2 r = fp2_one();
  bool started = false;
#pragma unroll 1
  for (int i = 11; i >= 0; --i)
#pragma unroll 1
    for (int b = 31; b >= 0; --b) {
      if (started) r = fp2_sqr(r);
      if ((e[i] >> b) & 1u) {
        r = started ? fp2_mul(r, a) : a;
        started = true;
      }
    }
  return r;
}

// (p+1)/4, (p-3)/4 and (p-1)/2 as 12 x u32 little-endian.
__device__ __constant__ const uint32_t EXP_P1_4[12] = {0xffffeaabu, 0xee7fbfffu, 0xac54ffffu, 0x07aaffffu,
                                                      0x3dac3d89u, 0xd9cc34a8u, 0x3ce144afu, 0xd91dd2e1u,
                                                      0x90d2eb35u, 0x92c6e9edu, 0x8e5ff9a6u, 0x0680447au};
__device__ __constant__ const uint32_t EXP_P3_4[12] = {0xffffeaaau, 0xee7fbfffu, 0xac54ffffu, 0x07aaffffu,
                                                      0x3dac3d89u, 0xd9cc34a8u, 0x3ce144afu, 0xd91dd2e1u,
                                                      0x90d2eb35u, 0x92c6e9edu, 0x8e5ff9a6u, 0x0680447au};
__device__ __constant__ const uint32_t EXP_P1_2[12] = {0xffffd555u, 0xdcff7fffu, 0x58a9ffffu, 0x0f55ffffu,
                                                      0x7b587b12u, 0xb3986950u, 0x79c2895fu, 0xb23ba5c2u,
                                                      0x21a5d66bu, 0x258dd3dbu, 0x1cbff34du, 0x0d0088f5u};
// r, the order of G1 and G2, as 8 x u32.
__device__ __constant__ const uint32_t ORDER_R[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                                                    0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};

BTE_DEV bool fp_sqrt(const fp& a, fp& y) {
  uint32_t e[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) e[i] = EXP_P1_4[i];
  y = fp_pow_limbs(a, e, 12);
  return fp_eq(fp_sqr(y), a);
}

BTE_DEV bool fp2_sqrt(const fp2& a, fp2& x) {
  uint32_t e[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) e[i] = EXP_P3_4[i];
  const fp2 a1 = fp2_pow_limbs(a, e);
  const fp2 alpha = fp2_mul(fp2_sqr(a1), a);
  const fp2 x0 = fp2_mul(a1, a);
  const fp2 minus_one = fp2_neg(fp2_one());
  if (fp2_eq(alpha, minus_one)) {
    x = {fp_neg(x0.c1), x0.c0};  // i x0
  } else {
#pragma unroll
    for (int i = 0; i < 12; ++i) e[i] = EXP_P1_2[i];
    x = fp2_mul(fp2_pow_limbs(fp2_add(alpha, fp2_one()), e), x0);
  }
  return fp2_eq(fp2_sqr(x), a);
}

BTE_DEV bool fp2_is_lex_largest(const fp2& y) {
  return fp_is_zero(y.c1) ? fp_is_lex_largest(y.c0) : fp_is_lex_largest(y.c1);
}

template <class F> BTE_DEV bool in_subgroup(const aff<F>& p) {
  uint32_t r[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) r[i] = ORDER_R[i];
  return jac_is_identity(jac_mul_scalar(jac_from_aff(p), r, 8));
}

__global__ void __launch_bounds__(64) k_g1_compress(const g1a* __restrict__ in, size_t n, uint8_t* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t b[48];
  const g1a p = in[i];
  if (aff_is_identity(p)) {
    for (int k = 0; k < 48; ++k) b[k] = 0;
    b[0] = 0xc0;
  } else {
    fp_to_be(fp_from_mont(p.x), b);
    b[0] |= 0x80 | (fp_is_lex_largest(p.y) ? 0x20 : 0);
  }
  for (int k = 0; k < 48; ++k) out[48 * i + k] = b[k];
}

__global__ void __launch_bounds__(64) k_g2_compress(const g2a* __restrict__ in, size_t n, uint8_t* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t b[96];
  const g2a p = in[i];
  if (aff_is_identity(p)) {
    for (int k = 0; k < 96; ++k) b[k] = 0;
    b[0] = 0xc0;
  } else {
    fp_to_be(fp_from_mont(p.x.c1), b);
    fp_to_be(fp_from_mont(p.x.c0), b + 48);
    b[0] |= 0x80 | (fp2_is_lex_largest(p.y) ? 0x20 : 0);
  }
  for (int k = 0; k < 96; ++k) out[96 * i + k] = b[k];
}

// Flags of an encoding: compressed bit set; the infinity encoding is exactly 0xc0 then zeros.
BTE_DEV bool read_flags(const uint8_t* b, int len, bool& inf, bool& big) {
  const uint8_t f = b[0];
  if (!(f & 0x80)) return false;
  inf = (f & 0x40) != 0;
  big = (f & 0x20) != 0;
  if (inf) {
    if (big || (f & 0x1f)) return false;
    for (int k = 1; k < len; ++k)
      if (b[k]) return false;
  }
  return true;
}

__global__ void __launch_bounds__(64) k_g1_decompress(const uint8_t* __restrict__ in, size_t n, g1a* __restrict__ out,
                                                      uint8_t* __restrict__ ok) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t b[48];
  for (int k = 0; k < 48; ++k) b[k] = in[48 * i + k];
  bool inf = false, big = false;
  bool good = read_flags(b, 48, inf, big);
  g1a p = {fp_zero(), fp_zero()};
  if (good && !inf) {
    b[0] &= 0x1f;
    const fp x = fp_to_mont(fp_from_be(b, good));
    const fp rhs = fp_add(fp_mul(fp_sqr(x), x), fp_mul4(fp_one()));
    fp y;
    good = good && fp_sqrt(rhs, y);
    if (good) {
      if (fp_is_lex_largest(y) != big) y = fp_neg(y);
      p = {x, y};
      good = in_subgroup(p);
    }
  }
  out[i] = good ? p : g1a{fp_zero(), fp_zero()};
  ok[i] = good ? 1 : 0;
}

__global__ void __launch_bounds__(64) k_g2_decompress(const uint8_t* __restrict__ in, size_t n, g2a* __restrict__ out,
                                                      uint8_t* __restrict__ ok) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t b[96];
  for (int k = 0; k < 96; ++k) b[k] = in[96 * i + k];
  bool inf = false, big = false;
  bool good = read_flags(b, 96, inf, big);
  g2a p = {fp2_zero(), fp2_zero()};
  if (good && !inf) {
    b[0] &= 0x1f;
    fp2 x;
    x.c1 = fp_to_mont(fp_from_be(b, good));
    x.c0 = fp_to_mont(fp_from_be(b + 48, good));
    const fp2 b4 = {fp_mul4(fp_one()), fp_mul4(fp_one())};  // 4 (u + 1)
    const fp2 rhs = fp2_add(fp2_mul(fp2_sqr(x), x), b4);
    fp2 y;
    good = good && fp2_sqrt(rhs, y);
    if (good) {
      if (fp2_is_lex_largest(y) != big) y = fp2_neg(y);
      p = {x, y};
      good = in_subgroup(p);
    }
  }
  out[i] = good ? p : g2a{fp2_zero(), fp2_zero()};
  ok[i] = good ? 1 : 0;
}

}  // namespace bte

using namespace bte;

extern "C" int bte_g1_compress(const bte_p1_affine* points, size_t n, uint8_t* out, void* stream) {
  if (n == 0) return BTE_OK;
  if (!points || !out) return set_error(BTE_EINVAL, "g1_compress: null pointer");
  hipLaunchKernelGGL(k_g1_compress, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream, (const g1a*)points, n,
                     out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_g2_compress(const bte_p2_affine* points, size_t n, uint8_t* out, void* stream) {
  if (n == 0) return BTE_OK;
  if (!points || !out) return set_error(BTE_EINVAL, "g2_compress: null pointer");
  hipLaunchKernelGGL(k_g2_compress, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream, (const g2a*)points, n,
                     out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_g1_decompress(const uint8_t* in, size_t n, bte_p1_affine* points, uint8_t* ok, void* stream) {
  if (n == 0) return BTE_OK;
  if (!in || !points || !ok) return set_error(BTE_EINVAL, "g1_decompress: null pointer");
  hipLaunchKernelGGL(k_g1_decompress, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream, in, n, (g1a*)points,
                     ok);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_g2_decompress(const uint8_t* in, size_t n, bte_p2_affine* points, uint8_t* ok, void* stream) {
  if (n == 0) return BTE_OK;
  if (!in || !points || !ok) return set_error(BTE_EINVAL, "g2_decompress: null pointer");
  hipLaunchKernelGGL(k_g2_decompress, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream, in, n, (g2a*)points,
                     ok);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}
