// Extension tower of BLS12-381 on gfx950:
//   Fp2  = Fp[u]  / (u^2 + 1)
//   Fp6  = Fp2[v] / (v^3 - xi),  xi = u + 1
//   Fp12 = Fp6[w] / (w^2 - v)
// Struct layouts equal blst_fp2 {fp[2]}, blst_fp6 {fp2[3]}, blst_fp12 {fp6[2]}.
#pragma once
#include "constants.hpp"
#include "fp.hpp"

namespace bte {

// Out-of-line device functions: every Fp12-level operation (and the inversions) is a real call
// whose body has all Fp / Fp2 arithmetic inlined, so a call costs a few hundred bytes of
// argument traffic against tens of thousands of instructions of work, and kernels built from
// them (Miller loops, final exponentiation, GT transforms) stay small enough for the
// instruction cache and for the compiler.
#define BTE_NOINL static __device__ __attribute__((noinline))

struct fp2 {
  fp c0, c1;
};
struct fp6 {
  fp2 c0, c1, c2;
};
struct fp12 {
  fp6 c0, c1;
};

// ---------------------------------------------------------------------------------------------
// Fp2
// ---------------------------------------------------------------------------------------------
BTE_DEV fp2 fp2_zero() { return {fp_zero(), fp_zero()}; }
BTE_DEV fp2 fp2_one() { return {fp_one(), fp_zero()}; }
BTE_DEV bool fp2_is_zero(const fp2& a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
BTE_DEV bool fp2_eq(const fp2& a, const fp2& b) { return fp_eq(a.c0, b.c0) && fp_eq(a.c1, b.c1); }
BTE_DEV fp2 fp2_add(const fp2& a, const fp2& b) { return {fp_add(a.c0, b.c0), fp_add(a.c1, b.c1)}; }
BTE_DEV fp2 fp2_sub(const fp2& a, const fp2& b) { return {fp_sub(a.c0, b.c0), fp_sub(a.c1, b.c1)}; }
BTE_DEV fp2 fp2_dbl(const fp2& a) { return {fp_dbl(a.c0), fp_dbl(a.c1)}; }
BTE_DEV fp2 fp2_neg(const fp2& a) { return {fp_neg(a.c0), fp_neg(a.c1)}; }
BTE_DEV fp2 fp2_conj(const fp2& a) { return {a.c0, fp_neg(a.c1)}; }
BTE_DEV fp2 fp2_half(const fp2& a) { return {fp_half(a.c0), fp_half(a.c1)}; }

// Karatsuba: 3 Fp multiplications.
BTE_DEV fp2 fp2_mul(const fp2& a, const fp2& b) {
  fp t0 = fp_mul(a.c0, b.c0);
  fp t1 = fp_mul(a.c1, b.c1);
  fp t2 = fp_mul(fp_add(a.c0, a.c1), fp_add(b.c0, b.c1));
  return {fp_sub(t0, t1), fp_sub(t2, fp_add(t0, t1))};
}

// The same product with lazy reduction: three 768-bit products combined at double width
// (c0 = a0 b0 + (p^2 - a1 b1), c1 = (a0 + a1)(b0 + b1) - a0 b0 - a1 b1, both below 2 p^2) and two
// Montgomery reductions instead of three.  Bit-identical to fp2_mul; used inside Fp6/Fp12
// products, where the extra double-width registers are affordable.
BTE_DEV fp2 fp2_mul_lazy(const fp2& a, const fp2& b) {
  const fpw t0 = fp_mul_wide(a.c0, b.c0);
  const fpw t1 = fp_mul_wide(a.c1, b.c1);
  const fpw t2 = fp_mul_wide(fp_add_raw(a.c0, a.c1), fp_add_raw(b.c0, b.c1));
  return {fp_redc_wide(fpw_add(t0, fpw_p2_sub(t1))), fp_redc_wide(fpw_sub(fpw_sub(t2, t0), t1))};
}

// (a0 + a1 u)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 u: 2 Fp multiplications.
BTE_DEV fp2 fp2_sqr(const fp2& a) {
  fp t = fp_mul(a.c0, a.c1);
  return {fp_mul(fp_add(a.c0, a.c1), fp_sub(a.c0, a.c1)), fp_dbl(t)};
}

BTE_DEV fp2 fp2_mul_fp(const fp2& a, const fp& b) { return {fp_mul(a.c0, b), fp_mul(a.c1, b)}; }

// multiplication by xi = 1 + u
BTE_DEV fp2 fp2_mul_xi(const fp2& a) { return {fp_sub(a.c0, a.c1), fp_add(a.c0, a.c1)}; }

BTE_DEV fp2 fp2_mul3(const fp2& a) { return fp2_add(fp2_dbl(a), a); }

BTE_NOINL fp2 fp2_inv(const fp2& a) {
  fp n = fp_add(fp_sqr(a.c0), fp_sqr(a.c1));
  fp ni = fp_inv(n);
  return {fp_mul(a.c0, ni), fp_neg(fp_mul(a.c1, ni))};
}

// ---------------------------------------------------------------------------------------------
// Fp6
// ---------------------------------------------------------------------------------------------
BTE_DEV fp6 fp6_zero() { return {fp2_zero(), fp2_zero(), fp2_zero()}; }
BTE_DEV fp6 fp6_one() { return {fp2_one(), fp2_zero(), fp2_zero()}; }
BTE_DEV fp6 fp6_add(const fp6& a, const fp6& b) {
  return {fp2_add(a.c0, b.c0), fp2_add(a.c1, b.c1), fp2_add(a.c2, b.c2)};
}
BTE_DEV fp6 fp6_sub(const fp6& a, const fp6& b) {
  return {fp2_sub(a.c0, b.c0), fp2_sub(a.c1, b.c1), fp2_sub(a.c2, b.c2)};
}
BTE_DEV fp6 fp6_neg(const fp6& a) { return {fp2_neg(a.c0), fp2_neg(a.c1), fp2_neg(a.c2)}; }
BTE_DEV bool fp6_is_zero(const fp6& a) {
  return fp2_is_zero(a.c0) && fp2_is_zero(a.c1) && fp2_is_zero(a.c2);
}
BTE_DEV bool fp6_eq(const fp6& a, const fp6& b) {
  return fp2_eq(a.c0, b.c0) && fp2_eq(a.c1, b.c1) && fp2_eq(a.c2, b.c2);
}

// Karatsuba over Fp2: 6 Fp2 multiplications.
BTE_DEV fp6 fp6_mul(const fp6& a, const fp6& b) {
  fp2 t0 = fp2_mul_lazy(a.c0, b.c0);
  fp2 t1 = fp2_mul_lazy(a.c1, b.c1);
  fp2 t2 = fp2_mul_lazy(a.c2, b.c2);
  fp2 c0 = fp2_add(t0, fp2_mul_xi(fp2_sub(fp2_mul_lazy(fp2_add(a.c1, a.c2), fp2_add(b.c1, b.c2)),
                                          fp2_add(t1, t2))));
  fp2 c1 = fp2_add(fp2_sub(fp2_mul_lazy(fp2_add(a.c0, a.c1), fp2_add(b.c0, b.c1)), fp2_add(t0, t1)),
                   fp2_mul_xi(t2));
  fp2 c2 = fp2_add(fp2_sub(fp2_mul_lazy(fp2_add(a.c0, a.c2), fp2_add(b.c0, b.c2)), fp2_add(t0, t2)), t1);
  return {c0, c1, c2};
}

// multiplication by v
BTE_DEV fp6 fp6_mul_v(const fp6& a) { return {fp2_mul_xi(a.c2), a.c0, a.c1}; }

// a * (b0 + b1 v): 5 Fp2 multiplications.
BTE_DEV fp6 fp6_mul_01(const fp6& a, const fp2& b0, const fp2& b1) {
  fp2 t0 = fp2_mul_lazy(a.c0, b0);
  fp2 t1 = fp2_mul_lazy(a.c1, b1);
  fp2 c0 = fp2_add(t0, fp2_mul_xi(fp2_mul_lazy(a.c2, b1)));
  fp2 c1 = fp2_sub(fp2_mul_lazy(fp2_add(a.c0, a.c1), fp2_add(b0, b1)), fp2_add(t0, t1));
  fp2 c2 = fp2_add(fp2_mul_lazy(a.c2, b0), t1);
  return {c0, c1, c2};
}

// a * (b1 v): 3 Fp2 multiplications.
BTE_DEV fp6 fp6_mul_1(const fp6& a, const fp2& b1) {
  return {fp2_mul_xi(fp2_mul_lazy(a.c2, b1)), fp2_mul_lazy(a.c0, b1), fp2_mul_lazy(a.c1, b1)};
}

BTE_NOINL fp6 fp6_inv(const fp6& a) {
  fp2 t0 = fp2_sub(fp2_sqr(a.c0), fp2_mul_xi(fp2_mul(a.c1, a.c2)));
  fp2 t1 = fp2_sub(fp2_mul_xi(fp2_sqr(a.c2)), fp2_mul(a.c0, a.c1));
  fp2 t2 = fp2_sub(fp2_sqr(a.c1), fp2_mul(a.c0, a.c2));
  fp2 den = fp2_add(fp2_mul(a.c0, t0), fp2_mul_xi(fp2_add(fp2_mul(a.c2, t1), fp2_mul(a.c1, t2))));
  fp2 di = fp2_inv(den);
  return {fp2_mul(t0, di), fp2_mul(t1, di), fp2_mul(t2, di)};
}

// ---------------------------------------------------------------------------------------------
// Fp12
// ---------------------------------------------------------------------------------------------
BTE_DEV fp12 fp12_one() { return {fp6_one(), fp6_zero()}; }
BTE_DEV bool fp12_is_one(const fp12& a) { return fp6_eq(a.c0, fp6_one()) && fp6_is_zero(a.c1); }
BTE_DEV bool fp12_eq(const fp12& a, const fp12& b) { return fp6_eq(a.c0, b.c0) && fp6_eq(a.c1, b.c1); }
BTE_DEV fp12 fp12_conj(const fp12& a) { return {a.c0, fp6_neg(a.c1)}; }

// Karatsuba over Fp6: 18 Fp2 multiplications.
BTE_NOINL fp12 fp12_mul(const fp12& a, const fp12& b) {
  fp6 t0 = fp6_mul(a.c0, b.c0);
  fp6 t1 = fp6_mul(a.c1, b.c1);
  fp6 c1 = fp6_sub(fp6_mul(fp6_add(a.c0, a.c1), fp6_add(b.c0, b.c1)), fp6_add(t0, t1));
  return {fp6_add(t0, fp6_mul_v(t1)), c1};
}

// Complex squaring: 2 Fp6 multiplications.
BTE_NOINL fp12 fp12_sqr(const fp12& a) {
  fp6 t = fp6_mul(a.c0, a.c1);
  fp6 c0 = fp6_sub(fp6_mul(fp6_add(a.c0, a.c1), fp6_add(a.c0, fp6_mul_v(a.c1))),
                   fp6_add(t, fp6_mul_v(t)));
  return {c0, fp6_add(t, t)};
}

// f * l for a Miller-loop line l = (l0 + l1 v) + (l4 v) w.
BTE_NOINL fp12 fp12_mul_line(const fp12& f, const fp2& l0, const fp2& l1, const fp2& l4) {
  fp6 a = fp6_mul_01(f.c0, l0, l1);
  fp6 b = fp6_mul_1(f.c1, l4);
  fp6 c = fp6_mul_01(fp6_add(f.c0, f.c1), l0, fp2_add(l1, l4));
  return {fp6_add(a, fp6_mul_v(b)), fp6_sub(c, fp6_add(a, b))};
}

BTE_NOINL fp12 fp12_inv(const fp12& a) {
  fp6 den = fp6_sub(fp6_mul(a.c0, a.c0), fp6_mul_v(fp6_mul(a.c1, a.c1)));
  fp6 di = fp6_inv(den);
  return {fp6_mul(a.c0, di), fp6_neg(fp6_mul(a.c1, di))};
}

BTE_DEV fp2 frob_gamma(int k, int i) {
  fp2 g;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    g.c0.l[j] = FROB_GAMMA[k - 1][i][0][j];
    g.c1.l[j] = FROB_GAMMA[k - 1][i][1][j];
  }
  return g;
}

// x -> x^(p^k), k in 1..3.  In the w basis x = sum g_i w^i with (g0, g2, g4) = c0 and
// (g1, g3, g5) = c1; each coefficient is conjugated for odd k and scaled by gamma_{k,i}.
BTE_NOINL fp12 fp12_frobenius(const fp12& a, int k) {
  const bool odd = k & 1;
  auto f = [&](const fp2& g, int i) {
    fp2 x = odd ? fp2_conj(g) : g;
    return i == 0 ? x : fp2_mul(x, frob_gamma(k, i));
  };
  fp12 r;
  r.c0.c0 = f(a.c0.c0, 0);
  r.c1.c0 = f(a.c1.c0, 1);
  r.c0.c1 = f(a.c0.c1, 2);
  r.c1.c1 = f(a.c1.c1, 3);
  r.c0.c2 = f(a.c0.c2, 4);
  r.c1.c2 = f(a.c1.c2, 5);
  return r;
}

// (a + b s)^2 in Fp4 = Fp2[s]/(s^2 - xi).
// With lazy reduction: each Fp2 square x^2 = (x0 + x1)(x0 - x1) + 2 x0 x1 u is kept at double
// width (operands reduced, so the parts are below p^2 and 2 p^2), c0 = xi t1 + t0 and
// c1 = (a + b)^2 - t0 - t1 are formed there with multiples of p^2 keeping every value
// non-negative and below 6 p^2 < p 2^384, and four Montgomery reductions replace six.
BTE_DEV void fp4_sqr(fp2& c0, fp2& c1, const fp2& a, const fp2& b) {
  const fpw A0 = fp_mul_wide(fp_add(a.c0, a.c1), fp_sub(a.c0, a.c1));
  const fpw a01 = fp_mul_wide(a.c0, a.c1);
  const fpw A1 = fpw_add(a01, a01);
  const fpw B0 = fp_mul_wide(fp_add(b.c0, b.c1), fp_sub(b.c0, b.c1));
  const fpw b01 = fp_mul_wide(b.c0, b.c1);
  const fpw B1 = fpw_add(b01, b01);
  c0.c0 = fp_redc_wide(fpw_add(fpw_add(A0, B0), fpw_kp2_sub(2, B1)));  // B0 - B1 + A0
  c0.c1 = fp_redc_wide(fpw_add(fpw_add(A1, B0), B1));                 // B0 + B1 + A1
  const fpw P0 = fpw_add(A0, B0), P1 = fpw_add(A1, B1);               // < 2 p^2, < 4 p^2
  const fp2 s = fp2_add(a, b);
  const fpw S0 = fp_mul_wide(fp_add(s.c0, s.c1), fp_sub(s.c0, s.c1));
  const fpw s01 = fp_mul_wide(s.c0, s.c1);
  c1.c0 = fp_redc_wide(fpw_add(S0, fpw_kp2_sub(2, P0)));
  c1.c1 = fp_redc_wide(fpw_add(fpw_add(s01, s01), fpw_kp2_sub(4, P1)));
}

// Granger-Scott squaring, valid on the cyclotomic subgroup.  With f = A + B w + C w^2 over
// Fp4 = Fp2[w^3]: f^2 = (3A^2 - 2conj A) + (3 s C^2 + 2 conj B) w + (3B^2 - 2 conj C) w^2.
BTE_NOINL fp12 fp12_cyclotomic_sqr(const fp12& f) {
  fp2 z0 = f.c0.c0, z4 = f.c0.c1, z3 = f.c0.c2, z2 = f.c1.c0, z1 = f.c1.c1, z5 = f.c1.c2;
  fp2 t0, t1, t2, t3;
  fp4_sqr(t0, t1, z0, z1);
  z0 = fp2_add(fp2_dbl(fp2_sub(t0, z0)), t0);
  z1 = fp2_add(fp2_dbl(fp2_add(t1, z1)), t1);
  fp4_sqr(t0, t1, z2, z3);
  fp4_sqr(t2, t3, z4, z5);
  z4 = fp2_add(fp2_dbl(fp2_sub(t0, z4)), t0);
  z5 = fp2_add(fp2_dbl(fp2_add(t1, z5)), t1);
  t0 = fp2_mul_xi(t3);
  z2 = fp2_add(fp2_dbl(fp2_add(t0, z2)), t0);
  z3 = fp2_add(fp2_dbl(fp2_sub(t2, z3)), t2);
  fp12 r;
  r.c0 = {z0, z4, z3};
  r.c1 = {z2, z1, z5};
  return r;
}

}  // namespace bte
