// Fp12 spread over three lanes of a wavefront (gfx950).
//
// Writing f = sum_k c_k w^k (c_k in Fp2, w^6 = xi = u + 1) and t = w^3 (so t^2 = xi), f is
// A_0 + A_1 w + A_2 w^2 over Fp4 = Fp2[t] with A_r = c_r + c_{r+3} t, and Fp12 = Fp4[w]/(w^3 - t).
// Lane r of a group of three consecutive lanes holds A_r (48 words instead of 144), so a
// wavefront carries 21 elements (lane 63 idles).  The operations split evenly:
//   * product: Karatsuba over the cubic extension, V_r = A_r B_r and M_r = (A_r + A_{r+1})
//     (B_r + B_{r+1}) on lane r (two Fp4 products = 18 Fp products per lane), then
//     H_0 = V_0 + t (M_1 - V_1 - V_2), H_1 = M_0 - V_0 - V_1 + t V_2, H_2 = M_2 - V_0 - V_2 + V_1;
//   * cyclotomic square (Granger-Scott): lane r squares its own A_r in Fp4; lane 0 combines its
//     own square, lanes 1 and 2 swap theirs (6 Fp products per lane);
//   * conjugation and Frobenius act coefficient-wise, on each lane's two coefficients.
// Values move between the lanes of a group through a per-wavefront LDS buffer laid out
// [word][lane]; the three lanes of a group always run the same instructions (they share the
// element and its exponent), so an exchange is a store followed by a load in program order, which
// the LDS executes in order for one wavefront.
#pragma once
#include "cyclo.hpp"

namespace bte {

struct fp4 {
  fp2 a, b;  // a + b t
};

BTE_DEV fp4 fp4_add(const fp4& p, const fp4& q) { return {fp2_add(p.a, q.a), fp2_add(p.b, q.b)}; }
BTE_DEV fp4 fp4_sub(const fp4& p, const fp4& q) { return {fp2_sub(p.a, q.a), fp2_sub(p.b, q.b)}; }
BTE_DEV fp4 fp4_mul_t(const fp4& p) { return {fp2_mul_xi(p.b), p.a}; }

BTE_DEV fp4 fp4_mul(const fp4& p, const fp4& q) {
  const fp2 v0 = fp2_mul(p.a, q.a);
  const fp2 v1 = fp2_mul(p.b, q.b);
  const fp2 m = fp2_mul(fp2_add(p.a, p.b), fp2_add(q.a, q.b));
  return {fp2_add(v0, fp2_mul_xi(v1)), fp2_sub(m, fp2_add(v0, v1))};
}

BTE_DEV fp4 fp4_sqr(const fp4& p) {
  const fp2 a2 = fp2_sqr(p.a);
  const fp2 b2 = fp2_sqr(p.b);
  const fp2 s = fp2_sqr(fp2_add(p.a, p.b));
  return {fp2_add(a2, fp2_mul_xi(b2)), fp2_sub(s, fp2_add(a2, b2))};
}

// Lane-group context.
struct c3ctx {
  volatile uint32_t* buf;  // 48 x 64 words of LDS for this wavefront
  int lane, r, base;       // lane in the wavefront, role 0..2, lane of role 0
};

BTE_DEV c3ctx c3_make(volatile uint32_t* buf) {
  c3ctx c;
  c.buf = buf;
  c.lane = threadIdx.x & 63;
  c.r = c.lane % 3;
  c.base = c.lane - c.r;
  return c;
}

BTE_DEV void c3_put_fp(const c3ctx& c, int w0, const fp& x) {
#pragma unroll
  for (int i = 0; i < 12; ++i) c.buf[(w0 + i) * 64 + c.lane] = x.l[i];
}
BTE_DEV fp c3_get_fp(const c3ctx& c, int w0, int src) {
  fp x;
#pragma unroll
  for (int i = 0; i < 12; ++i) x.l[i] = c.buf[(w0 + i) * 64 + src];
  return x;
}

BTE_DEV void c3_put(const c3ctx& c, const fp4& v) {
  c3_put_fp(c, 0, v.a.c0);
  c3_put_fp(c, 12, v.a.c1);
  c3_put_fp(c, 24, v.b.c0);
  c3_put_fp(c, 36, v.b.c1);
}
// The value role `role` of this lane's group last stored.
BTE_DEV fp4 c3_get(const c3ctx& c, int role) {
  const int src = c.base + role;
  return {{c3_get_fp(c, 0, src), c3_get_fp(c, 12, src)}, {c3_get_fp(c, 24, src), c3_get_fp(c, 36, src)}};
}

BTE_DEV fp4 c3_one(const c3ctx& c) {
  return c.r == 0 ? fp4{fp2_one(), fp2_zero()} : fp4{fp2_zero(), fp2_zero()};
}

// The exchanges are ordered so that at most three Fp4 values are live besides the operands of
// the Fp4 product in flight (the kernels target two wavefronts per SIMD, 256 registers).
BTE_DEV fp4 c3_mul(const c3ctx& c, const fp4& A, const fp4& B) {
  const int r1 = c.r == 2 ? 0 : c.r + 1, r2 = c.r == 0 ? 2 : c.r - 1;
  const fp4 V = fp4_mul(A, B);
  c3_put(c, A);
  const fp4 As = fp4_add(A, c3_get(c, r1));
  c3_put(c, B);
  const fp4 Bs = fp4_add(B, c3_get(c, r1));
  const fp4 M = fp4_mul(As, Bs);
  // Q folds the neighbours' V: role 0 -(V1 + V2), role 1 t V2 - V0, role 2 V1 - V0
  c3_put(c, V);
  const fp4 Vn = c3_get(c, r1);
  fp4 Q = c.r == 1 ? fp4_mul_t(Vn) : fp4_sub(fp4{fp2_zero(), fp2_zero()}, Vn);
  const fp4 Vnn = c3_get(c, r2);
  Q = c.r == 0 ? fp4_sub(Q, Vnn) : (c.r == 1 ? fp4_sub(Q, Vnn) : fp4_add(Q, Vnn));
  c3_put(c, M);
  const fp4 Mx = c3_get(c, c.r == 0 ? 1 : (c.r == 1 ? 0 : 2));
  // H_0 = V_0 + t (M_1 - V_1 - V_2); H_1 = M_0 - V_1 + t V_2 - V_0; H_2 = M_2 - V_2 + V_1 - V_0
  if (c.r == 0) return fp4_add(V, fp4_mul_t(fp4_add(Mx, Q)));
  return fp4_add(fp4_sub(Mx, V), Q);
}

BTE_DEV fp2 fp2_triple(const fp2& a) { return fp2_add(fp2_dbl(a), a); }

// Granger-Scott cyclotomic square (cyclotomic inputs only).
BTE_DEV fp4 c3_cyclotomic_sqr(const c3ctx& c, const fp4& A) {
  c3_put(c, fp4_sqr(A));
  const fp4 P = c3_get(c, c.r == 1 ? 2 : (c.r == 2 ? 1 : 0));
  if (c.r == 1) return {fp2_add(fp2_triple(fp2_mul_xi(P.b)), fp2_dbl(A.a)), fp2_sub(fp2_triple(P.a), fp2_dbl(A.b))};
  return {fp2_sub(fp2_triple(P.a), fp2_dbl(A.a)), fp2_add(fp2_triple(P.b), fp2_dbl(A.b))};
}

// c_k -> (-1)^k c_k: the inverse of a cyclotomic element.
BTE_DEV fp4 c3_conj(const c3ctx& c, const fp4& A) {
  return c.r == 1 ? fp4{fp2_neg(A.a), A.b} : fp4{A.a, fp2_neg(A.b)};
}

// Frobenius^k (k = 1, 2, 3): c_j -> conj^k(c_j) gamma_{k,j}.
BTE_DEV fp4 c3_frobenius(const c3ctx& c, const fp4& A, int k) {
  const fp2 a = (k & 1) ? fp2_conj(A.a) : A.a;
  const fp2 b = (k & 1) ? fp2_conj(A.b) : A.b;
  return {fp2_mul(a, frob_gamma(k, c.r)), fp2_mul(b, frob_gamma(k, c.r + 3))};
}

// Position of w-basis coefficient c_k in the blst fp12 layout {c0.c0, c0.c1, c0.c2, c1.c0, ...}.
BTE_DEV int c3_slot(int k) { return (k & 1) ? 3 + (k >> 1) : (k >> 1); }

BTE_DEV fp4 c3_load(const c3ctx& c, const fp12* __restrict__ p) {
  const fp2* q = (const fp2*)p;
  return {q[c3_slot(c.r)], q[c3_slot(c.r + 3)]};
}
BTE_DEV void c3_store(const c3ctx& c, fp12* __restrict__ p, const fp4& A) {
  fp2* q = (fp2*)p;
  q[c3_slot(c.r)] = A.a;
  q[c3_slot(c.r + 3)] = A.b;
}

// Per-lane spill area for the odd powers of the plan exponentiation: entry e, word i of global
// lane gl at tab[(e * 48 + i) * stride + gl] (coalesced across a wavefront).
struct c3tab {
  uint32_t* base;
  size_t stride;
  size_t gl;
};

BTE_DEV void c3_tab_put(const c3tab& t, int e, const fp4& v) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const fp x = j == 0 ? v.a.c0 : j == 1 ? v.a.c1 : j == 2 ? v.b.c0 : v.b.c1;
#pragma unroll
    for (int i = 0; i < 12; ++i) t.base[(size_t)(e * 48 + 12 * j + i) * t.stride + t.gl] = x.l[i];
  }
}
BTE_DEV fp4 c3_tab_get(const c3tab& t, int e) {
  fp x[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 12; ++i) x[j].l[i] = t.base[(size_t)(e * 48 + 12 * j + i) * t.stride + t.gl];
  return {{x[0], x[1]}, {x[2], x[3]}};
}

// The decomposed action of a plan (cyclo_pow_plan) on a three-lane element; the table
// phi_comp(v^(2i+1)) (conjugated for odd comp) of C3_TAB_ENTRIES entries lives in the spill area,
// so each Frobenius image is taken once per scalar.  Same result as cyclo_pow_plan.
static constexpr int C3_TAB_ENTRIES = 16;
static constexpr size_t C3_TAB_BYTES = (size_t)C3_TAB_ENTRIES * 48 * sizeof(uint32_t);  // per lane

BTE_DEV fp4 c3_cyclo_pow_plan(const c3ctx& c, const c3tab& t, const fp4& v, const cyc_plan* __restrict__ plan) {
  const int count = plan->count;
  if (count == 0) return c3_one(c);
  const fp4 two = c3_cyclotomic_sqr(c, v);
  fp4 cur = v;
  c3_tab_put(t, 0, cur);
#pragma unroll 1
  for (int i = 1; i < 4; ++i) {
    cur = c3_mul(c, cur, two);
    c3_tab_put(t, i, cur);
  }
#pragma unroll 1
  for (int comp = 1; comp < 4; ++comp)
#pragma unroll 1
    for (int i = 0; i < 4; ++i) {
      fp4 m = c3_frobenius(c, c3_tab_get(t, i), comp);
      if (comp & 1) m = c3_conj(c, m);
      c3_tab_put(t, 4 * comp + i, m);
    }
  fp4 acc = c3_one(c);
  bool started = false;
#pragma unroll 1
  for (int pos = count - 1; pos >= 0; --pos) {
    if (started) acc = c3_cyclotomic_sqr(c, acc);
#pragma unroll 1
    for (int comp = 0; comp < 4; ++comp) {
      const int d = plan->naf[comp][pos];
      if (d == 0) continue;
      fp4 m = c3_tab_get(t, 4 * comp + (((d < 0 ? -d : d) - 1) >> 1));
      if (d < 0) m = c3_conj(c, m);  // a negative digit inverts (conjugates)
      acc = started ? c3_mul(c, acc, m) : m;
      started = true;
    }
  }
  return acc;
}

}  // namespace bte
