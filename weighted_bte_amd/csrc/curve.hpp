// Short-Weierstrass group law (a = 0) for G1 over Fp (y^2 = x^3 + 4) and G2 over Fp2 on the
// M-type twist (y^2 = x^3 + 4(u+1)), Jacobian coordinates x = X/Z^2, y = Y/Z^3, Z = 0 the
// identity.  Affine points with x = y = 0 are the identity, as in blst_p{1,2}_affine.
// Struct layouts equal blst_p1 / blst_p1_affine / blst_p2 / blst_p2_affine.
#pragma once
#include "tower.hpp"

namespace bte {

// Field dispatch so the same formulas serve Fp and Fp2.
BTE_DEV fp f_add(const fp& a, const fp& b) { return fp_add(a, b); }
BTE_DEV fp f_sub(const fp& a, const fp& b) { return fp_sub(a, b); }
BTE_DEV fp f_mul(const fp& a, const fp& b) { return fp_mul(a, b); }
BTE_DEV fp f_sqr(const fp& a) { return fp_sqr(a); }
BTE_DEV fp f_dbl(const fp& a) { return fp_dbl(a); }
BTE_DEV fp f_neg(const fp& a) { return fp_neg(a); }
BTE_DEV fp f_inv(const fp& a) { return fp_inv(a); }
BTE_DEV bool f_is_zero(const fp& a) { return fp_is_zero(a); }
BTE_DEV bool f_eq(const fp& a, const fp& b) { return fp_eq(a, b); }
BTE_DEV fp2 f_add(const fp2& a, const fp2& b) { return fp2_add(a, b); }
BTE_DEV fp2 f_sub(const fp2& a, const fp2& b) { return fp2_sub(a, b); }
BTE_DEV fp2 f_mul(const fp2& a, const fp2& b) { return fp2_mul(a, b); }
BTE_DEV fp2 f_sqr(const fp2& a) { return fp2_sqr(a); }
BTE_DEV fp2 f_dbl(const fp2& a) { return fp2_dbl(a); }
BTE_DEV fp2 f_neg(const fp2& a) { return fp2_neg(a); }
BTE_DEV fp2 f_inv(const fp2& a) { return fp2_inv(a); }
BTE_DEV bool f_is_zero(const fp2& a) { return fp2_is_zero(a); }
BTE_DEV bool f_eq(const fp2& a, const fp2& b) { return fp2_eq(a, b); }

template <class F> BTE_DEV F f_zero();
template <> BTE_DEV fp f_zero<fp>() { return fp_zero(); }
template <> BTE_DEV fp2 f_zero<fp2>() { return fp2_zero(); }
template <class F> BTE_DEV F f_one();
template <> BTE_DEV fp f_one<fp>() { return fp_one(); }
template <> BTE_DEV fp2 f_one<fp2>() { return fp2_one(); }

template <class F> struct jac {
  F x, y, z;
};
template <class F> struct aff {
  F x, y;
};

using g1j = jac<fp>;
using g1a = aff<fp>;
using g2j = jac<fp2>;
using g2a = aff<fp2>;

template <class F> BTE_DEV jac<F> jac_identity() { return {f_one<F>(), f_one<F>(), f_zero<F>()}; }
template <class F> BTE_DEV bool jac_is_identity(const jac<F>& p) { return f_is_zero(p.z); }
template <class F> BTE_DEV bool aff_is_identity(const aff<F>& p) {
  return f_is_zero(p.x) && f_is_zero(p.y);
}
template <class F> BTE_DEV jac<F> jac_from_aff(const aff<F>& a) {
  if (aff_is_identity(a)) return jac_identity<F>();
  return {a.x, a.y, f_one<F>()};
}
template <class F> BTE_DEV jac<F> jac_neg(const jac<F>& p) { return {p.x, f_neg(p.y), p.z}; }
template <class F> BTE_DEV aff<F> aff_neg(const aff<F>& p) {
  return aff_is_identity(p) ? p : aff<F>{p.x, f_neg(p.y)};
}

// dbl-2009-l: 2M + 5S.
template <class F> BTE_DEV jac<F> jac_dbl(const jac<F>& p) {
  F A = f_sqr(p.x);
  F B = f_sqr(p.y);
  F C = f_sqr(B);
  F D = f_dbl(f_sub(f_sqr(f_add(p.x, B)), f_add(A, C)));
  F E = f_add(f_dbl(A), A);
  F Fv = f_sqr(E);
  jac<F> r;
  r.x = f_sub(Fv, f_dbl(D));
  F C8 = f_dbl(f_dbl(f_dbl(C)));
  r.y = f_sub(f_mul(E, f_sub(D, r.x)), C8);
  r.z = f_dbl(f_mul(p.y, p.z));
  return r;  // Z = 0 in, Z = 0 out; Y = 0 cannot occur for points of order r.
}

// add-2007-bl: 11M + 5S, with the exceptional cases handled.
template <class F> BTE_DEV jac<F> jac_add(const jac<F>& p, const jac<F>& q) {
  if (jac_is_identity(p)) return q;
  if (jac_is_identity(q)) return p;
  F Z1Z1 = f_sqr(p.z);
  F Z2Z2 = f_sqr(q.z);
  F U1 = f_mul(p.x, Z2Z2);
  F U2 = f_mul(q.x, Z1Z1);
  F S1 = f_mul(f_mul(p.y, q.z), Z2Z2);
  F S2 = f_mul(f_mul(q.y, p.z), Z1Z1);
  F H = f_sub(U2, U1);
  F R = f_sub(S2, S1);
  if (f_is_zero(H)) {
    if (f_is_zero(R)) return jac_dbl(p);
    return jac_identity<F>();
  }
  F I = f_sqr(f_dbl(H));
  F J = f_mul(H, I);
  F r2 = f_dbl(R);
  F V = f_mul(U1, I);
  jac<F> o;
  o.x = f_sub(f_sub(f_sqr(r2), J), f_dbl(V));
  o.y = f_sub(f_mul(r2, f_sub(V, o.x)), f_dbl(f_mul(S1, J)));
  o.z = f_mul(f_sub(f_sqr(f_add(p.z, q.z)), f_add(Z1Z1, Z2Z2)), H);
  return o;
}

// madd-2007-bl: 7M + 4S, q affine (identity allowed).
template <class F> BTE_DEV jac<F> jac_add_aff(const jac<F>& p, const aff<F>& q) {
  if (aff_is_identity(q)) return p;
  if (jac_is_identity(p)) return {q.x, q.y, f_one<F>()};
  F Z1Z1 = f_sqr(p.z);
  F U2 = f_mul(q.x, Z1Z1);
  F S2 = f_mul(f_mul(q.y, p.z), Z1Z1);
  F H = f_sub(U2, p.x);
  F R = f_sub(S2, p.y);
  if (f_is_zero(H)) {
    if (f_is_zero(R)) return jac_dbl(p);
    return jac_identity<F>();
  }
  F HH = f_sqr(H);
  F I = f_dbl(f_dbl(HH));
  F J = f_mul(H, I);
  F r2 = f_dbl(R);
  F V = f_mul(p.x, I);
  jac<F> o;
  o.x = f_sub(f_sub(f_sqr(r2), J), f_dbl(V));
  o.y = f_sub(f_mul(r2, f_sub(V, o.x)), f_dbl(f_mul(p.y, J)));
  o.z = f_sub(f_sub(f_sqr(f_add(p.z, H)), Z1Z1), HH);
  return o;
}

template <class F> BTE_DEV jac<F> jac_sub(const jac<F>& p, const jac<F>& q) {
  return jac_add(p, jac_neg(q));
}

// Affine form using one inversion (the identity maps to x = y = 0).
template <class F> BTE_DEV aff<F> jac_to_aff(const jac<F>& p) {
  if (jac_is_identity(p)) return {f_zero<F>(), f_zero<F>()};
  F zi = f_inv(p.z);
  F zi2 = f_sqr(zi);
  return {f_mul(p.x, zi2), f_mul(p.y, f_mul(zi2, zi))};
}

// k * p for a scalar given as nwords little-endian u32 words (left-to-right double-and-add).
template <class F> BTE_DEV jac<F> jac_mul_scalar(const jac<F>& p, const uint32_t* k, int nwords) {
  jac<F> acc = jac_identity<F>();
#pragma unroll 1
  for (int i = nwords * 32 - 1; i >= 0; --i) {
    acc = jac_dbl(acc);
    if ((k[i >> 5] >> (i & 31)) & 1u) acc = jac_add(acc, p);
  }
  return acc;
}

// Joint double-and-add for k1 P + k2 Q with k1, k2 below 2^128 (four u32 words each).
template <class F>
BTE_DEV jac<F> jac_mul2(const jac<F>& p, const jac<F>& q, const uint32_t* __restrict__ k1,
                        const uint32_t* __restrict__ k2) {
  const jac<F> pq = jac_add(p, q);
  int top = 127;
  while (top >= 0 && !(((k1[top >> 5] | k2[top >> 5]) >> (top & 31)) & 1u)) --top;
  jac<F> acc = jac_identity<F>();
#pragma unroll 1
  for (int i = top; i >= 0; --i) {
    acc = jac_dbl(acc);
    const uint32_t b1 = (k1[i >> 5] >> (i & 31)) & 1u, b2 = (k2[i >> 5] >> (i & 31)) & 1u;
    if (b1 | b2) acc = jac_add(acc, b1 ? (b2 ? pq : p) : q);
  }
  return acc;
}

// G1: phi(X, Y, Z) = (beta X, Y, Z) is multiplication by lambda = |z|^2 - 1, so for
// k = k1 + k2 lambda (k[0..3] = k1, k[4..7] = k2, both below 2^128) k P = k1 P + k2 phi(P).
BTE_DEV jac<fp> g1_glv_mul(const jac<fp>& p, const uint32_t* __restrict__ k) {
  fp beta;
#pragma unroll
  for (int i = 0; i < 12; ++i) beta.l[i] = G1_BETA[i];
  const jac<fp> q = {fp_mul(p.x, beta), p.y, p.z};
  return jac_mul2(p, q, k, k + 4);
}

// G2: psi(X, Y, Z) = (conj(X) c_x, conj(Y) c_y, conj(Z)) is multiplication by z, so psi^2 is
// multiplication by |z|^2; for k = k1 + k2 |z|^2 (k[0..3], k[4..7]) k Q = k1 Q + k2 psi^2(Q).
BTE_DEV jac<fp2> g2_psi(const jac<fp2>& p) {
  fp2 cx, cy;
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    cx.c0.l[i] = G2_PSI[0][0][i];
    cx.c1.l[i] = G2_PSI[0][1][i];
    cy.c0.l[i] = G2_PSI[1][0][i];
    cy.c1.l[i] = G2_PSI[1][1][i];
  }
  return {fp2_mul(fp2_conj(p.x), cx), fp2_mul(fp2_conj(p.y), cy), fp2_conj(p.z)};
}

BTE_DEV jac<fp2> g2_gls_mul(const jac<fp2>& p, const uint32_t* __restrict__ k) {
  return jac_mul2(p, g2_psi(g2_psi(p)), k, k + 4);
}

// Multiplication by an FFT twiddle from its two-part decomposition (8 x u32 for G1, then 8 x u32
// for G2: 16 words per twiddle) by the joint binary double-and-add: the shorter per-butterfly
// chain for transforms too small to fill the device.
BTE_DEV jac<fp> twiddle_mul(const jac<fp>& p, const uint32_t* __restrict__ dec) { return g1_glv_mul(p, dec); }
BTE_DEV jac<fp2> twiddle_mul(const jac<fp2>& p, const uint32_t* __restrict__ dec) {
  return g2_gls_mul(p, dec + 8);
}
BTE_DEV const uint32_t* twiddle_at(const uint32_t* dec, size_t k) { return dec + 16 * k; }

// Multiplication by an FFT twiddle k = k1 + k2 lambda (G1, lambda = |z|^2 - 1, endomorphism phi)
// or k = k1 + k2 |z|^2 (G2, psi^2), from host-built width-3 NAF digits of k1 and k2 (0, +-1, +-3,
// least significant first): one shared chain of at most 129 doublings and about a quarter as many
// additions per component, with P and 3P as the only precomputed points (the endomorphism is
// applied to them when a k2 digit uses them).  A third fewer additions than the binary form, at
// a longer chain per addition: the form for transforms that fill the device.
static constexpr int TW_NAF_LEN = 130;
static constexpr int TW_NAF_BLOCKS = 5;  // 16-byte blocks of 32 four-bit digits
struct tw_naf {
  // [group: 0 G1, 1 G2][component: k1, k2]: digit i (two's complement, 4 bits) in bits
  // 4 (i mod 8) .. of word i / 8
  uint32_t d[2][2][TW_NAF_BLOCKS * 4];
  int16_t top[2];  // highest nonzero position per group, -1 when k = 0
  int16_t pad[6];
};
static_assert(sizeof(tw_naf) % 16 == 0, "tw_naf arrays keep 16-byte digit blocks aligned");
BTE_DEV const tw_naf* twiddle_at(const tw_naf* t, size_t k) { return t + k; }

BTE_DEV uint32_t tw_word(const uint4& b, int j) { return j == 0 ? b.x : j == 1 ? b.y : j == 2 ? b.z : b.w; }
BTE_DEV int tw_digit(const uint4& b, int i) {
  const uint32_t nib = (tw_word(b, (i >> 3) & 3) >> (4 * (i & 7))) & 15u;
  return ((int)(nib << 28)) >> 28;
}

// The digit blocks are read 32 digits at a time, the next block fetched one block ahead, so a
// lone wavefront does not wait on a load per digit.
template <class F, class Endo>
BTE_DEV jac<F> jac_mul2_naf(const jac<F>& p, const uint32_t* __restrict__ d1, const uint32_t* __restrict__ d2,
                            int top, Endo endo) {
  jac<F> acc = jac_identity<F>();
  if (top < 0) return acc;
  const uint4* b1 = reinterpret_cast<const uint4*>(d1);
  const uint4* b2 = reinterpret_cast<const uint4*>(d2);
  int blk = top >> 5;
  uint4 c1 = b1[blk], c2 = b2[blk];
  uint4 n1 = c1, n2 = c2;
  if (blk) {
    n1 = b1[blk - 1];
    n2 = b2[blk - 1];
  }
  const jac<F> p3 = jac_add(jac_dbl(p), p);
#pragma unroll 1
  for (int i = top; i >= 0; --i) {
    if ((i & 31) == 31 && i != top) {  // entering block i / 32
      c1 = n1;
      c2 = n2;
      blk = i >> 5;
      if (blk) {
        n1 = b1[blk - 1];
        n2 = b2[blk - 1];
      }
    }
    acc = jac_dbl(acc);
    const int a = tw_digit(c1, i), b = tw_digit(c2, i);
    if (a) {
      const jac<F> t = (a == 1 || a == -1) ? p : p3;
      acc = jac_add(acc, a < 0 ? jac_neg(t) : t);
    }
    if (b) {
      const jac<F> t = endo((b == 1 || b == -1) ? p : p3);
      acc = jac_add(acc, b < 0 ? jac_neg(t) : t);
    }
  }
  return acc;
}

BTE_DEV jac<fp> twiddle_mul(const jac<fp>& p, const tw_naf* __restrict__ t) {
  fp beta;
#pragma unroll
  for (int i = 0; i < 12; ++i) beta.l[i] = G1_BETA[i];
  return jac_mul2_naf(p, t->d[0][0], t->d[0][1], t->top[0],
                      [&](const jac<fp>& q) { return jac<fp>{fp_mul(q.x, beta), q.y, q.z}; });
}

// psi^2(X, Y, Z) = (N(c_x) X, N(c_y) Y, Z) with the norms N(c) = conj(c) c in Fp of psi's constants.
BTE_DEV jac<fp2> twiddle_mul(const jac<fp2>& p, const tw_naf* __restrict__ t) {
  fp2 cx, cy;
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    cx.c0.l[i] = G2_PSI[0][0][i];
    cx.c1.l[i] = G2_PSI[0][1][i];
    cy.c0.l[i] = G2_PSI[1][0][i];
    cy.c1.l[i] = G2_PSI[1][1][i];
  }
  const fp nx = fp2_mul(fp2_conj(cx), cx).c0, ny = fp2_mul(fp2_conj(cy), cy).c0;
  return jac_mul2_naf(p, t->d[1][0], t->d[1][1], t->top[1],
                      [&](const jac<fp2>& q) { return jac<fp2>{fp2_mul_fp(q.x, nx), fp2_mul_fp(q.y, ny), q.z}; });
}

// Left-to-right double-and-add by a canonical scalar (8 x u32, below 2^255).
template <class F> BTE_DEV jac<F> group_scalar_mul(const jac<F>& p, const uint32_t* __restrict__ k) {
  jac<F> acc = jac_identity<F>();
  int top = 255;
  while (top >= 0 && !((k[top >> 5] >> (top & 31)) & 1u)) --top;
#pragma unroll 1
  for (int i = top; i >= 0; --i) {
    acc = jac_dbl(acc);
    if ((k[i >> 5] >> (i & 31)) & 1u) acc = jac_add(acc, p);
  }
  return acc;
}

}  // namespace bte
