// BLS12-381 base field Fp on gfx950: 12 x 32-bit little-endian limbs in Montgomery form with
// R = 2^384.  That is bit-for-bit the blst_fp layout (6 x u64 LE, R = 2^384), so values cross the
// C ABI without conversion.  Every function on fp keeps values fully reduced to [0, p), except
// fp_add_raw, whose sum below 2p feeds only the double-width products of lazy reduction.
//
// Multiplication is Montgomery with product scanning on v_mad_u64_u32 carry chains.  Because
// p < 2^381 the unreduced product is below 2p < 2^384, so one conditional subtraction finishes it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bte {

#define BTE_DEV __device__ __forceinline__

struct fp {
  uint32_t l[12];
};

// p, -p^-1 mod 2^32, R mod p, R^2 mod p (12 x u32 LE).
#define BTE_P_LIMBS                                                                              \
  0xffffaaabu, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu, 0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu,    \
      0x64774b84u, 0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau
#define BTE_ONE_LIMBS                                                                            \
  0x0002fffdu, 0x76090000u, 0xc40c0002u, 0xebf4000bu, 0x53c758bau, 0x5f489857u, 0x70525745u,    \
      0x77ce5853u, 0xa256ec6du, 0x5c071a97u, 0xfa80e493u, 0x15f65ec3u
#define BTE_R2_LIMBS                                                                             \
  0x1c341746u, 0xf4df1f34u, 0x09d104f1u, 0x0a76e6a6u, 0x4c95b6d5u, 0x8de5476cu, 0x939d83c0u,    \
      0x67eb88a9u, 0xb519952du, 0x9a793e85u, 0x92cae3aau, 0x11988fe5u
static constexpr uint32_t BTE_PINV = 0xfffcfffdu;

BTE_DEV uint32_t p_limb(int i) {
  constexpr uint32_t P[12] = {BTE_P_LIMBS};
  return P[i];
}

BTE_DEV fp fp_zero() {
  fp r;
#pragma unroll
  for (int i = 0; i < 12; ++i) r.l[i] = 0;
  return r;
}

BTE_DEV fp fp_one() {
  constexpr uint32_t ONE[12] = {BTE_ONE_LIMBS};
  fp r;
#pragma unroll
  for (int i = 0; i < 12; ++i) r.l[i] = ONE[i];
  return r;
}

BTE_DEV bool fp_is_zero(const fp& a) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) acc |= a.l[i];
  return acc == 0;
}

BTE_DEV bool fp_eq(const fp& a, const fp& b) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) acc |= a.l[i] ^ b.l[i];
  return acc == 0;
}

// r = a - p if a >= p else a, for a < 2p.
BTE_DEV fp fp_reduce_once(const fp& a) {
  fp t;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) t.l[i] = __builtin_subc(a.l[i], p_limb(i), borrow, &borrow);
  // borrow == 1 means a < p: keep a.
  fp r;
#pragma unroll
  for (int i = 0; i < 12; ++i) r.l[i] = borrow ? a.l[i] : t.l[i];
  return r;
}

BTE_DEV fp fp_add(const fp& a, const fp& b) {
  fp s;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) s.l[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
  return fp_reduce_once(s);
}

BTE_DEV fp fp_sub(const fp& a, const fp& b) {
  fp d;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) d.l[i] = __builtin_subc(a.l[i], b.l[i], borrow, &borrow);
  // if borrow, add p back
  uint32_t mask = 0u - borrow;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) d.l[i] = __builtin_addc(d.l[i], p_limb(i) & mask, c, &c);
  return d;
}

BTE_DEV fp fp_dbl(const fp& a) { return fp_add(a, a); }

BTE_DEV fp fp_neg(const fp& a) {
  uint32_t mask = fp_is_zero(a) ? 0u : 0xffffffffu;
  fp r;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) r.l[i] = __builtin_subc(p_limb(i) & mask, a.l[i], borrow, &borrow);
  return r;
}

// a / 2 mod p.
BTE_DEV fp fp_half(const fp& a) {
  uint32_t mask = 0u - (a.l[0] & 1u);
  fp t;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) t.l[i] = __builtin_addc(a.l[i], p_limb(i) & mask, c, &c);
  fp r;
#pragma unroll
  for (int i = 0; i < 11; ++i) r.l[i] = (t.l[i] >> 1) | (t.l[i + 1] << 31);
  r.l[11] = (t.l[11] >> 1) | (c << 31);
  return r;
}

// One multiply-accumulate step of the product-scanning loop: the 96-bit column accumulator
// {hi:acc} += a * b.  v_mad_u64_u32 returns the 64-bit sum and its carry-out mask; the carry goes
// into the third word with v_addc_co_u32, two VALU instructions per 32x32 product.
BTE_DEV void mac(uint64_t& acc, uint32_t& hi, uint32_t a, uint32_t b) {
  uint64_t c;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(acc), "=s"(c) : "v"(a), "v"(b), "v"(acc));
  asm("v_addc_co_u32 %0, %1, %2, 0, %3" : "=v"(hi), "=s"(c) : "v"(hi), "s"(c));
}

// Same with b held in a scalar register (the limbs of p are wave-uniform).
BTE_DEV void mac_s(uint64_t& acc, uint32_t& hi, uint32_t a, uint32_t b) {
  uint64_t c;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(acc), "=s"(c) : "v"(a), "s"(b), "v"(acc));
  asm("v_addc_co_u32 %0, %1, %2, 0, %3" : "=v"(hi), "=s"(c) : "v"(hi), "s"(c));
}

// Montgomery product a * b * 2^-384 mod p, product scanning with the reduction interleaved per
// column (column k collects a_i b_{k-i} and m_i p_{k-i}; m_k zeroes the low word of column k).
// With a, b < p the result before the final subtraction is < 2p.
BTE_DEV fp fp_mul(const fp& a, const fp& b) {
  uint32_t m[12];
  fp r;
  uint64_t acc = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int k = 0; k < 23; ++k) {
    const int ilo = k < 12 ? 0 : k - 11;
    const int ihi = k < 12 ? k : 11;
#pragma unroll
    for (int i = ilo; i <= ihi; ++i) mac(acc, hi, a.l[i], b.l[k - i]);
#pragma unroll
    for (int i = ilo; i < (k < 12 ? k : 12); ++i) mac_s(acc, hi, m[i], p_limb(k - i));
    if (k < 12) {
      m[k] = (uint32_t)acc * BTE_PINV;
      mac_s(acc, hi, m[k], p_limb(0));
    } else {
      r.l[k - 12] = (uint32_t)acc;
    }
    acc = (acc >> 32) | ((uint64_t)hi << 32);
    hi = 0;
  }
  r.l[11] = (uint32_t)acc;
  return fp_reduce_once(r);
}

BTE_DEV fp fp_sqr(const fp& a) { return fp_mul(a, a); }

// ---- double-width products for lazy reduction ---------------------------------------------------
// A 768-bit value as 24 little-endian limbs: a product of two elements below 2p, or a sum or
// difference of such products, reduced once by fp_redc_wide.
struct fpw {
  uint32_t l[24];
};

// p^2 (24 x u32 LE).
#define BTE_P2_LIMBS                                                                             \
  0x1c718e39u, 0x26aa0000u, 0x76382eabu, 0x7ced6b1du, 0x62113cfdu, 0x162c3383u, 0x3e71b743u,    \
      0x66bf91edu, 0x7091a049u, 0x292e85a8u, 0x86185c7bu, 0x1d68619cu, 0x0978ef01u, 0xf5314933u, \
      0x16ddca6eu, 0x50a62cfdu, 0x349e8bd0u, 0x66e59e49u, 0x0e7046b4u, 0xe2dc90e5u, 0xa22f25e9u, \
      0x4bd278eau, 0xb8c35fc7u, 0x02a437a4u

// a + b without reduction (a, b < p: the sum is below 2p < 2^382).
BTE_DEV fp fp_add_raw(const fp& a, const fp& b) {
  fp s;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) s.l[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
  return s;
}

// The full 768-bit product a * b (product scanning, no reduction).
BTE_DEV fpw fp_mul_wide(const fp& a, const fp& b) {
  fpw r;
  uint64_t acc = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int k = 0; k < 23; ++k) {
    const int ilo = k < 12 ? 0 : k - 11;
    const int ihi = k < 12 ? k : 11;
#pragma unroll
    for (int i = ilo; i <= ihi; ++i) mac(acc, hi, a.l[i], b.l[k - i]);
    r.l[k] = (uint32_t)acc;
    acc = (acc >> 32) | ((uint64_t)hi << 32);
    hi = 0;
  }
  r.l[23] = (uint32_t)acc;
  return r;
}

BTE_DEV fpw fpw_add(const fpw& a, const fpw& b) {
  fpw r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 24; ++i) r.l[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
  return r;
}

// a - b for a >= b.
BTE_DEV fpw fpw_sub(const fpw& a, const fpw& b) {
  fpw r;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 24; ++i) r.l[i] = __builtin_subc(a.l[i], b.l[i], borrow, &borrow);
  return r;
}

// k p^2 - b for k in {1, 2, 4} and b < k p^2.
BTE_DEV fpw fpw_kp2_sub(int k, const fpw& b) {
  constexpr uint32_t P2[24] = {BTE_P2_LIMBS};
  const int sh = k == 4 ? 2 : k == 2 ? 1 : 0;
  fpw kp;
  uint32_t prev = 0;
#pragma unroll
  for (int i = 0; i < 24; ++i) {  // k p^2 < 2^765: the shift loses no bits
    kp.l[i] = sh ? (P2[i] << sh) | (prev >> (32 - sh)) : P2[i];
    prev = P2[i];
  }
  fpw r;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 24; ++i) r.l[i] = __builtin_subc(kp.l[i], b.l[i], borrow, &borrow);
  return r;
}

// p^2 - b for b < p^2.
BTE_DEV fpw fpw_p2_sub(const fpw& b) {
  constexpr uint32_t P2[24] = {BTE_P2_LIMBS};
  fpw r;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 24; ++i) r.l[i] = __builtin_subc(P2[i], b.l[i], borrow, &borrow);
  return r;
}

// Montgomery reduction T 2^-384 mod p for T < p 2^384 (product scanning; the value before the
// final subtraction is below 2p).
BTE_DEV fp fp_redc_wide(const fpw& T) {
  uint32_t m[12];
  fp r;
  uint64_t acc = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int k = 0; k < 24; ++k) {
    {  // acc += T_k
      const uint64_t s = acc + T.l[k];
      hi += s < acc;
      acc = s;
    }
    const int ilo = k < 12 ? 0 : k - 11;
#pragma unroll
    for (int i = ilo; i < (k < 12 ? k : 12); ++i) mac_s(acc, hi, m[i], p_limb(k - i));
    if (k < 12) {
      m[k] = (uint32_t)acc * BTE_PINV;
      mac_s(acc, hi, m[k], p_limb(0));
    } else {
      r.l[k - 12] = (uint32_t)acc;
    }
    acc = (acc >> 32) | ((uint64_t)hi << 32);
    hi = 0;
  }
  return fp_reduce_once(r);
}

// Small-constant multiples used by the curve formulas.
BTE_DEV fp fp_mul3(const fp& a) { return fp_add(fp_dbl(a), a); }
BTE_DEV fp fp_mul4(const fp& a) { return fp_dbl(fp_dbl(a)); }
BTE_DEV fp fp_mul8(const fp& a) { return fp_dbl(fp_mul4(a)); }

// a^e for a public exponent given as 12 x u32 limbs, most significant bit first.
BTE_DEV fp fp_pow_limbs(const fp& a, const uint32_t* e, int nlimbs) {
  fp r = fp_one();
  bool started = false;
  for (int i = nlimbs - 1; i >= 0; --i) {
    for (int b = 31; b >= 0; --b) {
      if (started) r = fp_sqr(r);
      if ((e[i] >> b) & 1u) {
        r = started ? fp_mul(r, a) : a;
        started = true;
      }
    }
  }
  return r;
}

// a^-1 = a^(p-2) (Fermat); returns 0 for 0.
BTE_DEV fp fp_inv(const fp& a) {
  constexpr uint32_t E[12] = {0xffffaaa9u, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu,
                              0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu, 0x64774b84u,
                              0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
  return fp_pow_limbs(a, E, 12);
}

// Canonical (non-Montgomery) value of a Montgomery-form element: a * 1 * 2^-384.
BTE_DEV fp fp_from_mont(const fp& a) {
  fp one;
#pragma unroll
  for (int i = 0; i < 12; ++i) one.l[i] = 0;
  one.l[0] = 1;
  return fp_mul(a, one);
}

// Lexicographic "y > (p-1)/2" on the canonical value (the ZCash sign bit).
BTE_DEV bool fp_is_lex_largest(const fp& a_mont) {
  fp a = fp_from_mont(a_mont);
  // (p-1)/2
  constexpr uint32_t H[12] = {0xffffd555u, 0xdcff7fffu, 0x58a9ffffu, 0x0f55ffffu,
                              0x7b587b12u, 0xb3986950u, 0x79c2895fu, 0xb23ba5c2u,
                              0x21a5d66bu, 0x258dd3dbu, 0x1cbff34du, 0x0d0088f5u};
  // a > H  <=>  H - a borrows
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i) (void)__builtin_subc(H[i], a.l[i], borrow, &borrow);
  return borrow != 0;
}

}  // namespace bte
