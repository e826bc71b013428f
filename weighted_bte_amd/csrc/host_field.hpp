// Host-side BLS12-381 arithmetic used by the library's own host code: the last (serial) step of
// a Pippenger MSM, FFT twiddle tables and GT scalar plans.  Fp uses 6 x 64-bit Montgomery limbs
// with R = 2^384 (the same bits as the device's 12 x 32-bit form); Fr uses 4 x 64-bit limbs,
// canonical values, Montgomery R = 2^256 internally.
#pragma once
#include <stdint.h>
#include <string.h>

namespace bteh {

typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------------------------
// Generic Montgomery arithmetic over N 64-bit limbs.
// ---------------------------------------------------------------------------------------------
template <int N> inline bool geq(const uint64_t* a, const uint64_t* b) {
  for (int i = N - 1; i >= 0; --i) {
    if (a[i] != b[i]) return a[i] > b[i];
  }
  return true;
}

template <int N> inline void sub_n(uint64_t* r, const uint64_t* a, const uint64_t* b) {
  uint64_t borrow = 0;
  for (int i = 0; i < N; ++i) {
    u128 d = (u128)a[i] - b[i] - borrow;
    r[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) & 1;
  }
}

template <int N> inline uint64_t add_n(uint64_t* r, const uint64_t* a, const uint64_t* b) {
  uint64_t c = 0;
  for (int i = 0; i < N; ++i) {
    u128 s = (u128)a[i] + b[i] + c;
    r[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  return c;
}

// CIOS Montgomery product for moduli below 2^(64N - 2): the running value stays below 2p, so
// the accumulator needs one extra word and a single final conditional subtraction.
template <int N> inline void mont_mul(uint64_t* r, const uint64_t* a, const uint64_t* b,
                                      const uint64_t* p, uint64_t pinv) {
  uint64_t t[N + 1] = {0};
  for (int i = 0; i < N; ++i) {
    const uint64_t bi = b[i];
    uint64_t c = 0;
    for (int j = 0; j < N; ++j) {
      u128 x = (u128)a[j] * bi + t[j] + c;
      t[j] = (uint64_t)x;
      c = (uint64_t)(x >> 64);
    }
    const uint64_t tn = t[N] + c;
    const uint64_t m = t[0] * pinv;
    u128 x = (u128)m * p[0] + t[0];
    c = (uint64_t)(x >> 64);
    for (int j = 1; j < N; ++j) {
      x = (u128)m * p[j] + t[j] + c;
      t[j - 1] = (uint64_t)x;
      c = (uint64_t)(x >> 64);
    }
    t[N - 1] = tn + c;
    t[N] = 0;
  }
  uint64_t d[N];
  uint64_t borrow = 0;
  for (int i = 0; i < N; ++i) {
    u128 v = (u128)t[i] - p[i] - borrow;
    d[i] = (uint64_t)v;
    borrow = (uint64_t)(v >> 64) & 1;
  }
  const uint64_t keep = 0 - borrow;  // all ones when t < p
  for (int i = 0; i < N; ++i) r[i] = (t[i] & keep) | (d[i] & ~keep);
}

// ---------------------------------------------------------------------------------------------
// Fp
// ---------------------------------------------------------------------------------------------
static const uint64_t FP_P[6] = {0xb9feffffffffaaabull, 0x1eabfffeb153ffffull, 0x6730d2a0f6b0f624ull,
                                 0x64774b84f38512bfull, 0x4b1ba7b6434bacd7ull, 0x1a0111ea397fe69aull};
static const uint64_t FP_PINV = 0x89f3fffcfffcfffdull;
static const uint64_t FP_ONE[6] = {0x760900000002fffdull, 0xebf4000bc40c0002ull, 0x5f48985753c758baull,
                                   0x77ce585370525745ull, 0x5c071a97a256ec6dull, 0x15f65ec3fa80e493ull};

struct Fp {
  uint64_t l[6];
};

inline Fp fp_zero() { Fp r; memset(r.l, 0, 48); return r; }
inline Fp fp_one() { Fp r; memcpy(r.l, FP_ONE, 48); return r; }
inline bool fp_is_zero(const Fp& a) {
  uint64_t acc = 0;
  for (int i = 0; i < 6; ++i) acc |= a.l[i];
  return acc == 0;
}
inline Fp fp_add(const Fp& a, const Fp& b) {
  Fp r;
  add_n<6>(r.l, a.l, b.l);
  if (geq<6>(r.l, FP_P)) sub_n<6>(r.l, r.l, FP_P);
  return r;
}
inline Fp fp_sub(const Fp& a, const Fp& b) {
  Fp r;
  if (geq<6>(a.l, b.l)) {
    sub_n<6>(r.l, a.l, b.l);
  } else {
    uint64_t t[6];
    add_n<6>(t, a.l, FP_P);
    sub_n<6>(r.l, t, b.l);
  }
  return r;
}
inline Fp fp_mul(const Fp& a, const Fp& b) {
  Fp r;
  mont_mul<6>(r.l, a.l, b.l, FP_P, FP_PINV);
  return r;
}
inline Fp fp_neg(const Fp& a) { return fp_is_zero(a) ? a : fp_sub(fp_zero(), a); }

struct Fp2 {
  Fp c0, c1;
};
inline Fp2 fp2_zero() { return {fp_zero(), fp_zero()}; }
inline Fp2 fp2_one() { return {fp_one(), fp_zero()}; }

// Field dispatch for the curve templates.
inline Fp f_add(const Fp& a, const Fp& b) { return fp_add(a, b); }
inline Fp f_sub(const Fp& a, const Fp& b) { return fp_sub(a, b); }
inline Fp f_mul(const Fp& a, const Fp& b) { return fp_mul(a, b); }
inline bool f_is_zero(const Fp& a) { return fp_is_zero(a); }
inline Fp2 f_add(const Fp2& a, const Fp2& b) { return {fp_add(a.c0, b.c0), fp_add(a.c1, b.c1)}; }
inline Fp2 f_sub(const Fp2& a, const Fp2& b) { return {fp_sub(a.c0, b.c0), fp_sub(a.c1, b.c1)}; }
inline Fp2 f_mul(const Fp2& a, const Fp2& b) {
  Fp t0 = fp_mul(a.c0, b.c0), t1 = fp_mul(a.c1, b.c1);
  Fp t2 = fp_mul(fp_add(a.c0, a.c1), fp_add(b.c0, b.c1));
  return {fp_sub(t0, t1), fp_sub(t2, fp_add(t0, t1))};
}
inline bool f_is_zero(const Fp2& a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
template <class F> inline F f_one();
template <> inline Fp f_one<Fp>() { return fp_one(); }
template <> inline Fp2 f_one<Fp2>() { return fp2_one(); }
template <class F> inline F f_zero();
template <> inline Fp f_zero<Fp>() { return fp_zero(); }
template <> inline Fp2 f_zero<Fp2>() { return fp2_zero(); }

// Jacobian points (same layout as the device / blst structs).
template <class F> struct Jac {
  F x, y, z;
};

template <class F> inline Jac<F> jac_identity() { return {f_one<F>(), f_one<F>(), f_zero<F>()}; }
template <class F> inline bool jac_is_identity(const Jac<F>& p) { return f_is_zero(p.z); }

template <class F> inline Jac<F> jac_dbl(const Jac<F>& p) {
  if (jac_is_identity(p)) return p;
  F A = f_mul(p.x, p.x), B = f_mul(p.y, p.y), C = f_mul(B, B);
  F xb = f_add(p.x, B);
  F D = f_sub(f_mul(xb, xb), f_add(A, C));
  D = f_add(D, D);
  F E = f_add(f_add(A, A), A);
  F Fv = f_mul(E, E);
  Jac<F> r;
  r.x = f_sub(Fv, f_add(D, D));
  F C2 = f_add(C, C), C4 = f_add(C2, C2), C8 = f_add(C4, C4);
  r.y = f_sub(f_mul(E, f_sub(D, r.x)), C8);
  F yz = f_mul(p.y, p.z);
  r.z = f_add(yz, yz);
  return r;
}

template <class F> inline Jac<F> jac_add(const Jac<F>& p, const Jac<F>& q) {
  if (jac_is_identity(p)) return q;
  if (jac_is_identity(q)) return p;
  F Z1Z1 = f_mul(p.z, p.z), Z2Z2 = f_mul(q.z, q.z);
  F U1 = f_mul(p.x, Z2Z2), U2 = f_mul(q.x, Z1Z1);
  F S1 = f_mul(f_mul(p.y, q.z), Z2Z2), S2 = f_mul(f_mul(q.y, p.z), Z1Z1);
  F H = f_sub(U2, U1), R = f_sub(S2, S1);
  if (f_is_zero(H)) {
    if (f_is_zero(R)) return jac_dbl(p);
    return jac_identity<F>();
  }
  F H2 = f_add(H, H);
  F I = f_mul(H2, H2);
  F J = f_mul(H, I);
  F r2 = f_add(R, R);
  F V = f_mul(U1, I);
  Jac<F> o;
  o.x = f_sub(f_sub(f_mul(r2, r2), J), f_add(V, V));
  F s1j = f_mul(S1, J);
  o.y = f_sub(f_mul(r2, f_sub(V, o.x)), f_add(s1j, s1j));
  F zz = f_add(p.z, q.z);
  o.z = f_mul(f_sub(f_mul(zz, zz), f_add(Z1Z1, Z2Z2)), H);
  return o;
}

// ---------------------------------------------------------------------------------------------
// Fr (canonical 4 x u64 values at the interface; Montgomery inside mul).
// ---------------------------------------------------------------------------------------------
static const uint64_t FR_R[4] = {0xffffffff00000001ull, 0x53bda402fffe5bfeull, 0x3339d80809a1d805ull,
                                 0x73eda753299d7d48ull};
static const uint64_t FR_RINV = 0xfffffffeffffffffull;
static const uint64_t FR_R2[4] = {0xc999e990f3f29c6dull, 0x2b6cedcb87925c23ull, 0x05d314967254398full,
                                  0x0748d9d99f59ff11ull};
// 7^((r-1)/2^32), a primitive 2^32-th root of unity (the ff PrimeField::ROOT_OF_UNITY of Fr).
static const uint64_t FR_ROOT_OF_UNITY[4] = {0x3829971f439f0d2bull, 0xb63683508c2280b9ull, 0xd09b681922c813b4ull, 0x16a2a19edfe81f20ull};
static const int FR_S = 32;

struct Fr {
  uint64_t l[4];  // canonical, little-endian
};

inline Fr fr_from_u64(uint64_t v) { Fr r = {{v, 0, 0, 0}}; return r; }
inline bool fr_is_zero(const Fr& a) { return (a.l[0] | a.l[1] | a.l[2] | a.l[3]) == 0; }
inline Fr fr_mul(const Fr& a, const Fr& b) {
  // canonical * canonical: (a*R2*R^-1) = aR, then (aR * b * R^-1) = ab.
  uint64_t am[4];
  mont_mul<4>(am, a.l, FR_R2, FR_R, FR_RINV);
  Fr r;
  mont_mul<4>(r.l, am, b.l, FR_R, FR_RINV);
  return r;
}
inline Fr fr_add(const Fr& a, const Fr& b) {
  Fr r;
  add_n<4>(r.l, a.l, b.l);
  if (geq<4>(r.l, FR_R)) sub_n<4>(r.l, r.l, FR_R);
  return r;
}
inline Fr fr_sub(const Fr& a, const Fr& b) {
  Fr r;
  if (geq<4>(a.l, b.l)) {
    sub_n<4>(r.l, a.l, b.l);
  } else {
    uint64_t t[4];
    add_n<4>(t, a.l, FR_R);
    sub_n<4>(r.l, t, b.l);
  }
  return r;
}
inline Fr fr_pow(Fr base, const uint64_t* e, int nlimbs) {
  Fr r = fr_from_u64(1);
  for (int i = nlimbs * 64 - 1; i >= 0; --i) {
    r = fr_mul(r, r);
    if ((e[i >> 6] >> (i & 63)) & 1) r = fr_mul(r, base);
  }
  return r;
}
inline Fr fr_pow_u64(Fr base, uint64_t e) { return fr_pow(base, &e, 1); }
inline Fr fr_inv(const Fr& a) {
  uint64_t e[4];
  memcpy(e, FR_R, 32);
  e[0] -= 2;
  return fr_pow(a, e, 4);
}

}  // namespace bteh
