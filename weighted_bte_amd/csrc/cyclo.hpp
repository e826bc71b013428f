// Cyclotomic-subgroup and GT exponentiation on gfx950.
//
// cyclo_pow_plan: the four-way BLS-x / Frobenius decomposed action of
// weighted_btx/src/fft.rs:74-154.  A canonical scalar s = sum_i d_i |x|^i (four 64-bit d_i,
// fft.rs:543-557) acts as prod_i phi_i(v)^(d_i) with phi_i = Frobenius^i, conjugated for odd i
// (x < 0).  The d_i are given as interleaved width-4 NAF digits (fft.rs:569-594) that share one
// chain of at most 64 cyclotomic squarings; the odd multiples v, v^3, v^5, v^7 and their three
// Frobenius images are computed once per exponentiation.  The result is the same Fp12 value
// the reference computes (the product prod_i phi_i(v)^(d_i) does not depend on how the chain is
// organised), also for cyclotomic inputs outside GT.
#pragma once
#include "tower.hpp"

namespace bte {

static constexpr int CYC_DIGITS = 65;

// Host-built plan for one public scalar (fft.rs:33-50 CyclotomicScalarPlan).
struct cyc_plan {
  int8_t naf[4][CYC_DIGITS];
  uint8_t count;
  uint8_t pad[3];
};
static_assert(sizeof(cyc_plan) == 264, "cyc_plan layout");

BTE_DEV fp12 cyclo_pow_plan(const fp12& v, const cyc_plan* __restrict__ plan) {
  const int count = plan->count;
  if (count == 0) return fp12_one();
  // tab[comp][i] = phi_comp(v^(2i+1)), conjugated for odd comp: the 12 Frobenius images are
  // taken once here instead of once per nonzero digit of components 1..3 (about 39 per scalar).
  fp12 tab[4][4];
  tab[0][0] = v;
  const fp12 two = fp12_cyclotomic_sqr(v);
#pragma unroll 1
  for (int i = 1; i < 4; ++i) tab[0][i] = fp12_mul(tab[0][i - 1], two);
#pragma unroll 1
  for (int comp = 1; comp < 4; ++comp)
#pragma unroll 1
    for (int i = 0; i < 4; ++i) {
      const fp12 m = fp12_frobenius(tab[0][i], comp);
      tab[comp][i] = (comp & 1) ? fp12_conj(m) : m;
    }
  fp12 acc = fp12_one();
  bool started = false;
#pragma unroll 1
  for (int pos = count - 1; pos >= 0; --pos) {
    if (started) acc = fp12_cyclotomic_sqr(acc);
#pragma unroll 1
    for (int comp = 0; comp < 4; ++comp) {
      const int d = plan->naf[comp][pos];
      if (d == 0) continue;
      fp12 m = tab[comp][((d < 0 ? -d : d) - 1) >> 1];
      if (d < 0) m = fp12_conj(m);  // a negative digit inverts (conjugates)
      acc = started ? fp12_mul(acc, m) : m;
      started = true;
    }
  }
  return acc;
}

// v^c for a small signed integer by square-and-multiply with cyclotomic squaring
// (weighted_pfe/src/fft.rs:671-697 mul_cyclotomic_signed_small).
BTE_DEV fp12 cyclo_pow_small(const fp12& v, int64_t c) {
  if (c == 0) return fp12_one();
  uint64_t mag = c < 0 ? (uint64_t)(-c) : (uint64_t)c;
  fp12 base = v, result = fp12_one();
  bool started = false;
#pragma unroll 1
  while (mag) {
    if (mag & 1) {
      result = started ? fp12_mul(result, base) : base;
      started = true;
    }
    mag >>= 1;
    if (mag) base = fp12_cyclotomic_sqr(base);
  }
  return c < 0 ? fp12_conj(result) : result;
}

// ---- final exponentiation (weighted_btx/src/final_exponentiation.rs) -------------------------

// f^((p^6 - 1)(p^2 + 1)) given f^-1 (final_exponentiation.rs:356-371).
BTE_DEV fp12 easy_part_with_inverse(const fp12& f, const fp12& finv) {
  fp12 r = fp12_mul(fp12_conj(f), finv);
  return fp12_mul(r, fp12_frobenius(r, 2));
}

BTE_DEV fp12 mul_then_sqr(const fp12& acc, const fp12& m, int n) {
  fp12 r = fp12_mul(acc, m);
#pragma unroll 1
  for (int i = 0; i < n; ++i) r = fp12_cyclotomic_sqr(r);
  return r;
}

// final_exponentiation.rs:452-463: a^(|x|/2), conjugated.
BTE_DEV fp12 raise_to_z_div_by_2(const fp12& a) {
  fp12 r = fp12_cyclotomic_sqr(a);
  r = mul_then_sqr(r, a, 2);
  r = mul_then_sqr(r, a, 3);
  r = mul_then_sqr(r, a, 9);
  r = mul_then_sqr(r, a, 32);
  r = mul_then_sqr(r, a, 15);
  return fp12_conj(r);
}

BTE_DEV fp12 raise_to_z(const fp12& a) { return fp12_cyclotomic_sqr(raise_to_z_div_by_2(a)); }

// Hard part, exponent 3 * (p^4 - p^2 + 1) / r, as the addition chain of
// final_exponentiation.rs:308-336.
BTE_DEV fp12 hard_part(const fp12& f) {
  fp12 y0 = fp12_cyclotomic_sqr(f);
  fp12 y1 = raise_to_z(y0);
  fp12 y2 = raise_to_z_div_by_2(y1);
  fp12 y3 = fp12_conj(f);
  y1 = fp12_mul(y1, y3);
  y1 = fp12_conj(y1);
  y1 = fp12_mul(y1, y2);
  y2 = raise_to_z(y1);
  y3 = raise_to_z(y2);
  y1 = fp12_conj(y1);
  y3 = fp12_mul(y3, y1);
  y1 = fp12_conj(y1);
  y1 = fp12_frobenius(y1, 3);
  y2 = fp12_frobenius(y2, 2);
  y1 = fp12_mul(y1, y2);
  y2 = raise_to_z(y3);
  y2 = fp12_mul(y2, y0);
  y2 = fp12_mul(y2, f);
  y1 = fp12_mul(y1, y2);
  y2 = fp12_frobenius(y3, 1);
  return fp12_mul(y1, y2);
}

}  // namespace bte
