// The BLS12-381 scalar field Fr on gfx950: 8 x 32-bit little-endian limbs, Montgomery form with
// R = 2^256 inside arithmetic, canonical values at every interface.  Used by key generation,
// where it is a small share of the work next to the group operations, so the product is plain
// CIOS over 64-bit products.
#pragma once
#include "fp.hpp"

namespace bte {

struct fr {
  uint32_t l[8];
};

BTE_DEV uint32_t fr_limb(int i) {
  constexpr uint32_t R[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                             0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
  return R[i];
}
static constexpr uint32_t FR_NINV = 0xffffffffu;  // -r^-1 mod 2^32

BTE_DEV fr fr_reduce_once(const fr& a) {
  fr t;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) t.l[i] = __builtin_subc(a.l[i], fr_limb(i), borrow, &borrow);
  return borrow ? a : t;
}

BTE_DEV fr fr_mont_mul(const fr& a, const fr& b) {
  uint32_t t[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) t[j] = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint64_t u = (uint64_t)a.l[j] * b.l[i] + t[j] + c;
      t[j] = (uint32_t)u;
      c = u >> 32;
    }
    uint64_t u = (uint64_t)t[8] + c;
    t[8] = (uint32_t)u;
    t[9] = (uint32_t)(u >> 32);
    const uint32_t m = t[0] * FR_NINV;
    u = (uint64_t)m * fr_limb(0) + t[0];
    c = u >> 32;
#pragma unroll
    for (int j = 1; j < 8; ++j) {
      u = (uint64_t)m * fr_limb(j) + t[j] + c;
      t[j - 1] = (uint32_t)u;
      c = u >> 32;
    }
    u = (uint64_t)t[8] + c;
    t[7] = (uint32_t)u;
    t[8] = t[9] + (uint32_t)(u >> 32);
  }
  fr r;
#pragma unroll
  for (int j = 0; j < 8; ++j) r.l[j] = t[j];
  return fr_reduce_once(r);  // inputs below r give a result below 2r
}

BTE_DEV fr fr_to_mont(const fr& a) {
  constexpr uint32_t R2[8] = {0xf3f29c6du, 0xc999e990u, 0x87925c23u, 0x2b6cedcbu,
                              0x7254398fu, 0x05d31496u, 0x9f59ff11u, 0x0748d9d9u};
  fr r2;
#pragma unroll
  for (int i = 0; i < 8; ++i) r2.l[i] = R2[i];
  return fr_mont_mul(a, r2);
}

BTE_DEV fr fr_from_mont(const fr& a) {
  fr one;
#pragma unroll
  for (int i = 0; i < 8; ++i) one.l[i] = 0;
  one.l[0] = 1;
  return fr_mont_mul(a, one);
}

BTE_DEV fr fr_add(const fr& a, const fr& b) {
  fr r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.l[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
  return fr_reduce_once(r);  // a, b < r < 2^255: no carry out of the top limb
}

BTE_DEV fr fr_sub(const fr& a, const fr& b) {
  fr r;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.l[i] = __builtin_subc(a.l[i], b.l[i], borrow, &borrow);
  if (borrow) {
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.l[i] = __builtin_addc(r.l[i], fr_limb(i), c, &c);
  }
  return r;
}

}  // namespace bte
