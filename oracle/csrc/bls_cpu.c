/*
 * CPU oracle, C restatement (TEST INFRASTRUCTURE ONLY: used by tests/ as a checker and by
 * bench.py's cpu_baseline leg; the product never links it).
 *
 * BLS12-381 Fp / Fp2 Montgomery arithmetic on 6 x 64-bit limbs (R = 2^384, the blst layout),
 * Jacobian G1/G2 group law and a Pippenger MSM parallelised with OpenMP over windows, following
 * the semantics of blst_p{1,2}s_mult_pippenger as the reference calls it
 * (weighted_btx/src/blst_utils.rs:43-152): sum_i s_i P_i over the low 255 bits of 32-byte
 * little-endian scalars; affine inputs with x = y = 0 are the identity.
 *
 * Pinned in tests/test_oracle_cpu.py against the Python oracle (oracle/bls12_381.py) on random
 * inputs and against the closed form sum(k_i s_i) G.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

typedef struct { uint64_t l[6]; } fp;
typedef struct { fp c0, c1; } fp2;

static const fp P = {{0xb9feffffffffaaabull, 0x1eabfffeb153ffffull, 0x6730d2a0f6b0f624ull,
                      0x64774b84f38512bfull, 0x4b1ba7b6434bacd7ull, 0x1a0111ea397fe69aull}};
static const uint64_t PINV = 0x89f3fffcfffcfffdull;
static const fp ONE = {{0x760900000002fffdull, 0xebf4000bc40c0002ull, 0x5f48985753c758baull,
                        0x77ce585370525745ull, 0x5c071a97a256ec6dull, 0x15f65ec3fa80e493ull}};

static int fp_geq(const fp* a, const fp* b) {
  for (int i = 5; i >= 0; --i)
    if (a->l[i] != b->l[i]) return a->l[i] > b->l[i];
  return 1;
}
static void fp_sub_raw(fp* r, const fp* a, const fp* b) {
  uint64_t br = 0;
  for (int i = 0; i < 6; ++i) {
    u128 d = (u128)a->l[i] - b->l[i] - br;
    r->l[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
}
static fp fp_add(fp a, fp b) {
  fp r;
  uint64_t c = 0;
  for (int i = 0; i < 6; ++i) {
    u128 s = (u128)a.l[i] + b.l[i] + c;
    r.l[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  if (fp_geq(&r, &P)) fp_sub_raw(&r, &r, &P);
  return r;
}
static fp fp_sub(fp a, fp b) {
  fp r;
  if (fp_geq(&a, &b)) {
    fp_sub_raw(&r, &a, &b);
  } else {
    fp t;
    uint64_t c = 0;
    for (int i = 0; i < 6; ++i) {
      u128 s = (u128)a.l[i] + P.l[i] + c;
      t.l[i] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
    fp_sub_raw(&r, &t, &b);
  }
  return r;
}
static int fp_is_zero(fp a) { return (a.l[0] | a.l[1] | a.l[2] | a.l[3] | a.l[4] | a.l[5]) == 0; }
static fp fp_zero(void) { fp r; memset(&r, 0, sizeof r); return r; }
static fp fp_neg(fp a) { return fp_is_zero(a) ? a : fp_sub(fp_zero(), a); }
static fp fp_dbl(fp a) { return fp_add(a, a); }

/* CIOS Montgomery multiplication. */
static fp fp_mul(fp a, fp b) {
  uint64_t t[8] = {0};
  for (int i = 0; i < 6; ++i) {
    uint64_t c = 0;
    for (int j = 0; j < 6; ++j) {
      u128 x = (u128)a.l[j] * b.l[i] + t[j] + c;
      t[j] = (uint64_t)x;
      c = (uint64_t)(x >> 64);
    }
    u128 s = (u128)t[6] + c;
    t[6] = (uint64_t)s;
    t[7] = (uint64_t)(s >> 64);
    uint64_t m = t[0] * PINV;
    u128 x = (u128)m * P.l[0] + t[0];
    c = (uint64_t)(x >> 64);
    for (int j = 1; j < 6; ++j) {
      x = (u128)m * P.l[j] + t[j] + c;
      t[j - 1] = (uint64_t)x;
      c = (uint64_t)(x >> 64);
    }
    s = (u128)t[6] + c;
    t[5] = (uint64_t)s;
    t[6] = t[7] + (uint64_t)(s >> 64);
  }
  fp r;
  memcpy(r.l, t, 48);
  if (t[6] || fp_geq(&r, &P)) fp_sub_raw(&r, &r, &P);
  return r;
}

static fp2 fp2_add(fp2 a, fp2 b) { fp2 r = {fp_add(a.c0, b.c0), fp_add(a.c1, b.c1)}; return r; }
static fp2 fp2_sub(fp2 a, fp2 b) { fp2 r = {fp_sub(a.c0, b.c0), fp_sub(a.c1, b.c1)}; return r; }
static fp2 fp2_neg(fp2 a) { fp2 r = {fp_neg(a.c0), fp_neg(a.c1)}; return r; }
static fp2 fp2_dbl(fp2 a) { return fp2_add(a, a); }
static int fp2_is_zero(fp2 a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
static fp2 fp2_mul(fp2 a, fp2 b) {
  fp t0 = fp_mul(a.c0, b.c0), t1 = fp_mul(a.c1, b.c1);
  fp t2 = fp_mul(fp_add(a.c0, a.c1), fp_add(b.c0, b.c1));
  fp2 r = {fp_sub(t0, t1), fp_sub(t2, fp_add(t0, t1))};
  return r;
}

/* ---- Jacobian group law, generated for Fp (G1) and Fp2 (G2) ---------------------------- */
#define DEFINE_CURVE(PFX, F, F_ADD, F_SUB, F_MUL, F_DBL, F_NEG, F_ISZ, F_ONE, F_ZERO)          \
  typedef struct { F x, y, z; } PFX##_jac;                                                   \
  typedef struct { F x, y; } PFX##_aff;                                                      \
  static PFX##_jac PFX##_identity(void) { PFX##_jac r = {F_ONE, F_ONE, F_ZERO}; return r; }  \
  static int PFX##_is_identity(const PFX##_jac* p) { return F_ISZ(p->z); }                   \
  static int PFX##_aff_is_identity(const PFX##_aff* p) { return F_ISZ(p->x) && F_ISZ(p->y); } \
  static PFX##_jac PFX##_dbl(PFX##_jac p) {                                                  \
    if (F_ISZ(p.z)) return p;                                                                \
    F A = F_MUL(p.x, p.x), B = F_MUL(p.y, p.y), C = F_MUL(B, B);                             \
    F xb = F_ADD(p.x, B);                                                                    \
    F D = F_DBL(F_SUB(F_MUL(xb, xb), F_ADD(A, C)));                                          \
    F E = F_ADD(F_DBL(A), A);                                                                \
    PFX##_jac r;                                                                             \
    r.x = F_SUB(F_MUL(E, E), F_DBL(D));                                                      \
    r.y = F_SUB(F_MUL(E, F_SUB(D, r.x)), F_DBL(F_DBL(F_DBL(C))));                            \
    r.z = F_DBL(F_MUL(p.y, p.z));                                                            \
    return r;                                                                                \
  }                                                                                          \
  static PFX##_jac PFX##_add(PFX##_jac p, PFX##_jac q) {                                     \
    if (F_ISZ(p.z)) return q;                                                                \
    if (F_ISZ(q.z)) return p;                                                                \
    F Z1Z1 = F_MUL(p.z, p.z), Z2Z2 = F_MUL(q.z, q.z);                                        \
    F U1 = F_MUL(p.x, Z2Z2), U2 = F_MUL(q.x, Z1Z1);                                          \
    F S1 = F_MUL(F_MUL(p.y, q.z), Z2Z2), S2 = F_MUL(F_MUL(q.y, p.z), Z1Z1);                  \
    F H = F_SUB(U2, U1), R = F_SUB(S2, S1);                                                  \
    if (F_ISZ(H)) {                                                                          \
      if (F_ISZ(R)) return PFX##_dbl(p);                                                     \
      return PFX##_identity();                                                               \
    }                                                                                        \
    F H2 = F_DBL(H);                                                                         \
    F I = F_MUL(H2, H2), J = F_MUL(H, I), r2 = F_DBL(R), V = F_MUL(U1, I);                   \
    PFX##_jac o;                                                                             \
    o.x = F_SUB(F_SUB(F_MUL(r2, r2), J), F_DBL(V));                                          \
    o.y = F_SUB(F_MUL(r2, F_SUB(V, o.x)), F_DBL(F_MUL(S1, J)));                              \
    F zz = F_ADD(p.z, q.z);                                                                  \
    o.z = F_MUL(F_SUB(F_MUL(zz, zz), F_ADD(Z1Z1, Z2Z2)), H);                                 \
    return o;                                                                                \
  }                                                                                          \
  static PFX##_jac PFX##_add_aff(PFX##_jac p, PFX##_aff q) {                                 \
    if (PFX##_aff_is_identity(&q)) return p;                                                 \
    if (F_ISZ(p.z)) { PFX##_jac r = {q.x, q.y, F_ONE}; return r; }                           \
    F Z1Z1 = F_MUL(p.z, p.z);                                                                \
    F U2 = F_MUL(q.x, Z1Z1), S2 = F_MUL(F_MUL(q.y, p.z), Z1Z1);                              \
    F H = F_SUB(U2, p.x), R = F_SUB(S2, p.y);                                                \
    if (F_ISZ(H)) {                                                                          \
      if (F_ISZ(R)) return PFX##_dbl(p);                                                     \
      return PFX##_identity();                                                               \
    }                                                                                        \
    F HH = F_MUL(H, H), I = F_DBL(F_DBL(HH)), J = F_MUL(H, I), r2 = F_DBL(R);                \
    F V = F_MUL(p.x, I);                                                                     \
    PFX##_jac o;                                                                             \
    o.x = F_SUB(F_SUB(F_MUL(r2, r2), J), F_DBL(V));                                          \
    o.y = F_SUB(F_MUL(r2, F_SUB(V, o.x)), F_DBL(F_MUL(p.y, J)));                             \
    F zh = F_ADD(p.z, H);                                                                    \
    o.z = F_SUB(F_SUB(F_MUL(zh, zh), Z1Z1), HH);                                             \
    return o;                                                                                \
  }                                                                                          \
  static PFX##_aff PFX##_aff_neg(PFX##_aff q) {                                              \
    if (!PFX##_aff_is_identity(&q)) q.y = F_NEG(q.y);                                        \
    return q;                                                                                \
  }                                                                                          \
  /* Pippenger with signed c-bit digits; windows in parallel.  Same digit rule as the GPU. */ \
  static void PFX##_msm(const PFX##_aff* pts, const uint8_t* sc, size_t n, int c,            \
                        PFX##_jac* out) {                                                    \
    int nw = (256 + c - 1) / c;                                                              \
    uint32_t M = 1u << (c - 1);                                                              \
    int32_t* dig = (int32_t*)malloc(sizeof(int32_t) * n * (size_t)nw);                       \
    for (size_t i = 0; i < n; ++i) {                                                         \
      uint32_t s[9];                                                                         \
      memcpy(s, sc + 32 * i, 32);                                                            \
      s[7] &= 0x7fffffffu;                                                                   \
      s[8] = 0;                                                                              \
      int carry = 0;                                                                         \
      for (int w = 0; w < nw; ++w) {                                                         \
        int bit = w * c, word = bit >> 5, off = bit & 31;                                    \
        uint64_t two = (uint64_t)s[word] | ((uint64_t)s[word + 1] << 32);                    \
        int d = (int)((two >> off) & ((1u << c) - 1u)) + carry;                              \
        if (d > (int)M) { d -= (int)(1u << c); carry = 1; } else { carry = 0; }              \
        dig[(size_t)w * n + i] = d;                                                          \
      }                                                                                      \
    }                                                                                        \
    PFX##_jac* wsum = (PFX##_jac*)malloc(sizeof(PFX##_jac) * nw);                            \
    _Pragma("omp parallel for schedule(dynamic, 1)")                                         \
    for (int w = 0; w < nw; ++w) {                                                           \
      PFX##_jac* bk = (PFX##_jac*)malloc(sizeof(PFX##_jac) * M);                             \
      for (uint32_t b = 0; b < M; ++b) bk[b] = PFX##_identity();                             \
      for (size_t i = 0; i < n; ++i) {                                                       \
        int d = dig[(size_t)w * n + i];                                                      \
        if (!d) continue;                                                                    \
        PFX##_aff q = pts[i];                                                                \
        if (d < 0) { q = PFX##_aff_neg(q); d = -d; }                                         \
        bk[d - 1] = PFX##_add_aff(bk[d - 1], q);                                             \
      }                                                                                      \
      PFX##_jac run = PFX##_identity(), acc = PFX##_identity();                              \
      for (int b = (int)M - 1; b >= 0; --b) {                                                \
        run = PFX##_add(run, bk[b]);                                                         \
        acc = PFX##_add(acc, run);                                                           \
      }                                                                                      \
      wsum[w] = acc;                                                                         \
      free(bk);                                                                              \
    }                                                                                        \
    PFX##_jac acc = PFX##_identity();                                                        \
    for (int w = nw - 1; w >= 0; --w) {                                                      \
      for (int k = 0; k < c; ++k) acc = PFX##_dbl(acc);                                      \
      acc = PFX##_add(acc, wsum[w]);                                                         \
    }                                                                                        \
    *out = acc;                                                                              \
    free(wsum);                                                                              \
    free(dig);                                                                               \
  }

static fp2 FP2_ONE_V(void) { fp2 r = {ONE, {{0}}}; return r; }
static fp2 FP2_ZERO_V(void) { fp2 r; memset(&r, 0, sizeof r); return r; }

DEFINE_CURVE(g1, fp, fp_add, fp_sub, fp_mul, fp_dbl, fp_neg, fp_is_zero, ONE, fp_zero())
DEFINE_CURVE(g2, fp2, fp2_add, fp2_sub, fp2_mul, fp2_dbl, fp2_neg, fp2_is_zero, FP2_ONE_V(),
             FP2_ZERO_V())

static int default_window(size_t n) {
  int lg = 0;
  while (((size_t)1 << (lg + 1)) <= n) ++lg;
  int c = lg - 4;
  if (c < 4) c = 4;
  if (c > 16) c = 16;
  return c;
}

static fp fp_inv(fp a) {
  /* a^(p-2) */
  fp e = P;
  e.l[0] -= 2;
  fp r = ONE;
  for (int i = 383; i >= 0; --i) {
    r = fp_mul(r, r);
    if ((e.l[i >> 6] >> (i & 63)) & 1) r = fp_mul(r, a);
  }
  return r;
}


/* ======================================================================================== *
 * Pairing path (final_exponentiation.rs, fft.rs): Fp2 inverse, Fp6, Fp12, Miller loop with
 * projective lines, split final exponentiation with the blst hard chain, decomposed cyclotomic
 * power; G1 scalar multiplication and radix-2 G1 / cyclotomic transforms.  Values after the
 * final exponentiation are compared with the Python oracle in tests/test_oracle_cpu.py.
 * ======================================================================================== */
typedef struct { fp2 c0, c1, c2; } fp6;
typedef struct { fp6 c0, c1; } fp12;

static fp2 fp2_mul_xi(fp2 a) { fp2 r = {fp_sub(a.c0, a.c1), fp_add(a.c0, a.c1)}; return r; }
static fp2 fp2_conj(fp2 a) { fp2 r = {a.c0, fp_neg(a.c1)}; return r; }
static fp2 fp2_sqr(fp2 a) { return fp2_mul(a, a); }
static fp2 fp2_inv(fp2 a) {
  fp n = fp_add(fp_mul(a.c0, a.c0), fp_mul(a.c1, a.c1));
  fp ni = fp_inv(n);
  fp2 r = {fp_mul(a.c0, ni), fp_neg(fp_mul(a.c1, ni))};
  return r;
}
static fp2 fp2_mul_fp(fp2 a, fp b) { fp2 r = {fp_mul(a.c0, b), fp_mul(a.c1, b)}; return r; }

static fp6 fp6_add(fp6 a, fp6 b) { fp6 r = {fp2_add(a.c0, b.c0), fp2_add(a.c1, b.c1), fp2_add(a.c2, b.c2)}; return r; }
static fp6 fp6_sub(fp6 a, fp6 b) { fp6 r = {fp2_sub(a.c0, b.c0), fp2_sub(a.c1, b.c1), fp2_sub(a.c2, b.c2)}; return r; }
static fp6 fp6_neg(fp6 a) { fp6 r = {fp2_neg(a.c0), fp2_neg(a.c1), fp2_neg(a.c2)}; return r; }
static fp6 fp6_mul(fp6 a, fp6 b) {
  fp2 t0 = fp2_mul(a.c0, b.c0), t1 = fp2_mul(a.c1, b.c1), t2 = fp2_mul(a.c2, b.c2);
  fp6 r;
  r.c0 = fp2_add(t0, fp2_mul_xi(fp2_sub(fp2_mul(fp2_add(a.c1, a.c2), fp2_add(b.c1, b.c2)), fp2_add(t1, t2))));
  r.c1 = fp2_add(fp2_sub(fp2_mul(fp2_add(a.c0, a.c1), fp2_add(b.c0, b.c1)), fp2_add(t0, t1)), fp2_mul_xi(t2));
  r.c2 = fp2_add(fp2_sub(fp2_mul(fp2_add(a.c0, a.c2), fp2_add(b.c0, b.c2)), fp2_add(t0, t2)), t1);
  return r;
}
static fp6 fp6_mul_v(fp6 a) { fp6 r = {fp2_mul_xi(a.c2), a.c0, a.c1}; return r; }
static fp6 fp6_inv(fp6 a) {
  fp2 t0 = fp2_sub(fp2_sqr(a.c0), fp2_mul_xi(fp2_mul(a.c1, a.c2)));
  fp2 t1 = fp2_sub(fp2_mul_xi(fp2_sqr(a.c2)), fp2_mul(a.c0, a.c1));
  fp2 t2 = fp2_sub(fp2_sqr(a.c1), fp2_mul(a.c0, a.c2));
  fp2 den = fp2_add(fp2_mul(a.c0, t0), fp2_mul_xi(fp2_add(fp2_mul(a.c2, t1), fp2_mul(a.c1, t2))));
  fp2 di = fp2_inv(den);
  fp6 r = {fp2_mul(t0, di), fp2_mul(t1, di), fp2_mul(t2, di)};
  return r;
}
static fp12 fp12_one(void) {
  fp12 r;
  memset(&r, 0, sizeof r);
  r.c0.c0.c0 = ONE;
  return r;
}
static fp12 fp12_mul(fp12 a, fp12 b) {
  fp6 t0 = fp6_mul(a.c0, b.c0), t1 = fp6_mul(a.c1, b.c1);
  fp12 r;
  r.c1 = fp6_sub(fp6_mul(fp6_add(a.c0, a.c1), fp6_add(b.c0, b.c1)), fp6_add(t0, t1));
  r.c0 = fp6_add(t0, fp6_mul_v(t1));
  return r;
}
static fp12 fp12_sqr(fp12 a) { return fp12_mul(a, a); }
static fp12 fp12_conj(fp12 a) { a.c1 = fp6_neg(a.c1); return a; }
static fp12 fp12_inv(fp12 a) {
  fp6 den = fp6_sub(fp6_mul(a.c0, a.c0), fp6_mul_v(fp6_mul(a.c1, a.c1)));
  fp6 di = fp6_inv(den);
  fp12 r = {fp6_mul(a.c0, di), fp6_neg(fp6_mul(a.c1, di))};
  return r;
}

/* Frobenius coefficients gamma_{k,i} = (u+1)^(i (p^k-1)/6), computed once at load. */
static fp2 FROB[4][6];
static int frob_ready = 0;
static fp2 fp2_pow_big(fp2 a, const uint64_t* e, int nl) {
  fp2 r = {ONE, fp_zero()};
  for (int i = nl * 64 - 1; i >= 0; --i) {
    r = fp2_sqr(r);
    if ((e[i >> 6] >> (i & 63)) & 1) r = fp2_mul(r, a);
  }
  return r;
}
int oracle_set_frobenius(const uint64_t* exps /* 3 x 6 x 24 limbs: i (p^k - 1)/6 */) {
  fp2 xi = {ONE, ONE};
  for (int k = 1; k <= 3; ++k)
    for (int i = 0; i < 6; ++i) FROB[k][i] = fp2_pow_big(xi, exps + ((k - 1) * 6 + i) * 24, 24);
  frob_ready = 1;
  return 0;
}
static fp12 fp12_frob(fp12 a, int k) {
  fp2 g[6] = {a.c0.c0, a.c1.c0, a.c0.c1, a.c1.c1, a.c0.c2, a.c1.c2};
  for (int i = 0; i < 6; ++i) {
    if (k & 1) g[i] = fp2_conj(g[i]);
    if (i) g[i] = fp2_mul(g[i], FROB[k][i]);
  }
  fp12 r;
  r.c0.c0 = g[0]; r.c1.c0 = g[1]; r.c0.c1 = g[2]; r.c1.c1 = g[3]; r.c0.c2 = g[4]; r.c1.c2 = g[5];
  return r;
}
static fp12 cyc_sqr(fp12 a) { return fp12_sqr(a); }

/* line l0 + (cx xp) v + (cy yp) v w, multiplied in as a full Fp12 element */
static fp12 line_eval(fp2 l0, fp2 cx, fp2 cy, g1_aff P) {
  fp12 l;
  memset(&l, 0, sizeof l);
  l.c0.c0 = l0;
  l.c0.c1 = fp2_mul_fp(cx, P.x);
  l.c1.c1 = fp2_mul_fp(cy, P.y);
  return l;
}
static fp fp_half(fp a) {
  fp t = a;
  uint64_t c = 0;
  if (a.l[0] & 1) {
    for (int i = 0; i < 6; ++i) {
      u128 s2 = (u128)a.l[i] + P.l[i] + c;
      t.l[i] = (uint64_t)s2;
      c = (uint64_t)(s2 >> 64);
    }
  }
  fp r;
  for (int i = 0; i < 5; ++i) r.l[i] = (t.l[i] >> 1) | (t.l[i + 1] << 63);
  r.l[5] = (t.l[5] >> 1) | (c << 63);
  return r;
}
static fp2 fp2_half(fp2 a) { fp2 r = {fp_half(a.c0), fp_half(a.c1)}; return r; }
static fp2 fp2_mul3(fp2 a) { return fp2_add(fp2_dbl(a), a); }

static const uint64_t X_ABS = 0xd201000000010000ull;

/* f_{|x|,Q}(P) conjugated (homogeneous projective T on the twist). */
static fp12 miller(g1_aff P, g2_aff Q) {
  if (g1_aff_is_identity(&P) || g2_aff_is_identity(&Q)) return fp12_one();
  fp2 X = Q.x, Y = Q.y, Z = FP2_ONE_V();
  fp12 f = fp12_one();
  for (int b = 62; b >= 0; --b) {
    if (b != 62) f = fp12_sqr(f);
    {
      fp2 A = fp2_half(fp2_mul(X, Y)), B = fp2_sqr(Y), C = fp2_sqr(Z);
      fp2 E = fp2_dbl(fp2_dbl(fp2_mul3(fp2_mul_xi(C))));
      fp2 F = fp2_mul3(E), G = fp2_half(fp2_add(B, F));
      fp2 H = fp2_sub(fp2_sqr(fp2_add(Y, Z)), fp2_add(B, C));
      f = fp12_mul(f, line_eval(fp2_sub(B, E), fp2_neg(fp2_mul3(fp2_sqr(X))), H, P));
      X = fp2_mul(A, fp2_sub(B, F));
      Y = fp2_sub(fp2_sqr(G), fp2_mul3(fp2_sqr(E)));
      Z = fp2_mul(B, H);
    }
    if ((X_ABS >> b) & 1) {
      fp2 th = fp2_sub(Y, fp2_mul(Q.y, Z)), la = fp2_sub(X, fp2_mul(Q.x, Z));
      fp2 C = fp2_sqr(th), D = fp2_sqr(la), E = fp2_mul(la, D), F = fp2_mul(Z, C), G = fp2_mul(X, D);
      fp2 H = fp2_sub(fp2_add(E, F), fp2_dbl(G));
      f = fp12_mul(f, line_eval(fp2_sub(fp2_mul(th, Q.x), fp2_mul(la, Q.y)), fp2_neg(th), la, P));
      fp2 X3 = fp2_mul(la, H);
      Y = fp2_sub(fp2_mul(th, fp2_sub(G, H)), fp2_mul(Y, E));
      X = X3;
      Z = fp2_mul(Z, E);
    }
  }
  return fp12_conj(f);
}

/* Prepared form: the 68 (l0, cx, cy) line coefficient triples of Q, then evaluation at P. */
typedef struct { fp2 l0, cx, cy; } line3;
static void prepare_lines(g2_aff Q, line3* L) {
  fp2 X = Q.x, Y = Q.y, Z = FP2_ONE_V();
  int k = 0;
  for (int b = 62; b >= 0; --b) {
    fp2 A = fp2_half(fp2_mul(X, Y)), B = fp2_sqr(Y), C = fp2_sqr(Z);
    fp2 E = fp2_dbl(fp2_dbl(fp2_mul3(fp2_mul_xi(C))));
    fp2 F = fp2_mul3(E), G = fp2_half(fp2_add(B, F));
    fp2 H = fp2_sub(fp2_sqr(fp2_add(Y, Z)), fp2_add(B, C));
    L[k].l0 = fp2_sub(B, E); L[k].cx = fp2_neg(fp2_mul3(fp2_sqr(X))); L[k].cy = H; ++k;
    X = fp2_mul(A, fp2_sub(B, F));
    Y = fp2_sub(fp2_sqr(G), fp2_mul3(fp2_sqr(E)));
    Z = fp2_mul(B, H);
    if ((X_ABS >> b) & 1) {
      fp2 th = fp2_sub(Y, fp2_mul(Q.y, Z)), la = fp2_sub(X, fp2_mul(Q.x, Z));
      fp2 C2 = fp2_sqr(th), D = fp2_sqr(la), E2 = fp2_mul(la, D), F2 = fp2_mul(Z, C2), G2 = fp2_mul(X, D);
      fp2 H2 = fp2_sub(fp2_add(E2, F2), fp2_dbl(G2));
      L[k].l0 = fp2_sub(fp2_mul(th, Q.x), fp2_mul(la, Q.y)); L[k].cx = fp2_neg(th); L[k].cy = la; ++k;
      fp2 X3 = fp2_mul(la, H2);
      Y = fp2_sub(fp2_mul(th, fp2_sub(G2, H2)), fp2_mul(Y, E2));
      X = X3;
      Z = fp2_mul(Z, E2);
    }
  }
}
static fp12 miller_prepared(const line3* L, g1_aff P) {
  if (g1_aff_is_identity(&P)) return fp12_one();
  fp12 f = fp12_one();
  int k = 0;
  for (int b = 62; b >= 0; --b) {
    if (b != 62) f = fp12_sqr(f);
    f = fp12_mul(f, line_eval(L[k].l0, L[k].cx, L[k].cy, P)); ++k;
    if ((X_ABS >> b) & 1) { f = fp12_mul(f, line_eval(L[k].l0, L[k].cx, L[k].cy, P)); ++k; }
  }
  return fp12_conj(f);
}

static fp12 easy_part(fp12 f) {
  fp12 r = fp12_mul(fp12_conj(f), fp12_inv(f));
  return fp12_mul(r, fp12_frob(r, 2));
}
static fp12 mul_then_sqr(fp12 acc, fp12 m, int n) {
  fp12 r = fp12_mul(acc, m);
  for (int i = 0; i < n; ++i) r = cyc_sqr(r);
  return r;
}
static fp12 raise_z_half(fp12 a) {
  fp12 r = cyc_sqr(a);
  r = mul_then_sqr(r, a, 2);
  r = mul_then_sqr(r, a, 3);
  r = mul_then_sqr(r, a, 9);
  r = mul_then_sqr(r, a, 32);
  r = mul_then_sqr(r, a, 15);
  return fp12_conj(r);
}
static fp12 raise_z(fp12 a) { return cyc_sqr(raise_z_half(a)); }
static fp12 hard_part(fp12 f) {
  fp12 y0 = cyc_sqr(f), y1 = raise_z(y0), y2 = raise_z_half(y1), y3 = fp12_conj(f);
  y1 = fp12_conj(fp12_mul(y1, y3));
  y1 = fp12_mul(y1, y2);
  y2 = raise_z(y1);
  y3 = raise_z(y2);
  y1 = fp12_conj(y1);
  y3 = fp12_mul(y3, y1);
  y1 = fp12_conj(y1);
  y1 = fp12_frob(y1, 3);
  y2 = fp12_frob(y2, 2);
  y1 = fp12_mul(y1, y2);
  y2 = raise_z(y3);
  y2 = fp12_mul(fp12_mul(y2, y0), f);
  y1 = fp12_mul(y1, y2);
  return fp12_mul(y1, fp12_frob(y3, 1));
}

/* Decomposed cyclotomic power with a host-built plan (4 x 65 signed NAF digits + count). */
typedef struct { int8_t naf[4][65]; uint8_t count; uint8_t pad[3]; } cyc_plan_t;
static fp12 cyc_pow_plan(fp12 v, const cyc_plan_t* p) {
  if (!p->count) return fp12_one();
  fp12 odd[4];
  odd[0] = v;
  fp12 two = cyc_sqr(v);
  for (int i = 1; i < 4; ++i) odd[i] = fp12_mul(odd[i - 1], two);
  fp12 acc = fp12_one();
  int started = 0;
  for (int pos = p->count - 1; pos >= 0; --pos) {
    if (started) acc = cyc_sqr(acc);
    for (int c = 0; c < 4; ++c) {
      int d = p->naf[c][pos];
      if (!d) continue;
      fp12 m = odd[((d < 0 ? -d : d) - 1) >> 1];
      if (c) m = fp12_frob(m, c);
      if ((c & 1) != (d < 0)) m = fp12_conj(m);
      acc = started ? fp12_mul(acc, m) : m;
      started = 1;
    }
  }
  return acc;
}

static g1_jac g1_mul_scalar(g1_jac p, const uint64_t* k) {
  g1_jac acc = g1_identity();
  for (int i = 255; i >= 0; --i) {
    acc = g1_dbl(acc);
    if ((k[i >> 6] >> (i & 63)) & 1) acc = g1_add(acc, p);
  }
  return acc;
}

/* ---- exported pairing-path entry points (all parallel with OpenMP) ----------------------- */
int oracle_miller_pairs(const void* p, const void* q, size_t n, void* out) {
  #pragma omp parallel for schedule(dynamic, 4)
  for (size_t i = 0; i < n; ++i) ((fp12*)out)[i] = miller(((const g1_aff*)p)[i], ((const g2_aff*)q)[i]);
  return 0;
}
/* product of Miller loops per segment of `len` consecutive pairs */
int oracle_multi_miller(const void* p, const void* q, size_t nseg, size_t len, void* out) {
  #pragma omp parallel for schedule(dynamic, 1)
  for (size_t s2 = 0; s2 < nseg; ++s2) {
    fp12 acc = fp12_one();
    for (size_t i = 0; i < len; ++i)
      acc = fp12_mul(acc, miller(((const g1_aff*)p)[s2 * len + i], ((const g2_aff*)q)[s2 * len + i]));
    ((fp12*)out)[s2] = acc;
  }
  return 0;
}
int oracle_prepare_lines(const void* q, size_t n, void* lines) {
  #pragma omp parallel for schedule(dynamic, 4)
  for (size_t i = 0; i < n; ++i) prepare_lines(((const g2_aff*)q)[i], ((line3*)lines) + 68 * i);
  return 0;
}
int oracle_miller_prepared(const void* lines, const void* p, size_t n, void* out) {
  #pragma omp parallel for schedule(dynamic, 4)
  for (size_t i = 0; i < n; ++i)
    ((fp12*)out)[i] = miller_prepared(((const line3*)lines) + 68 * i, ((const g1_aff*)p)[i]);
  return 0;
}
int oracle_final_exp(const void* in, size_t n, int easy, int hard, void* out) {
  #pragma omp parallel for schedule(dynamic, 4)
  for (size_t i = 0; i < n; ++i) {
    fp12 f = ((const fp12*)in)[i];
    if (easy) f = easy_part(f);
    if (hard) f = hard_part(f);
    ((fp12*)out)[i] = f;
  }
  return 0;
}
int oracle_cyc_pow_plans(void* v, const void* plans, size_t n) {
  #pragma omp parallel for schedule(dynamic, 4)
  for (size_t i = 0; i < n; ++i) ((fp12*)v)[i] = cyc_pow_plan(((fp12*)v)[i], ((const cyc_plan_t*)plans) + i);
  return 0;
}
/* One radix-2 stage of a cyclotomic (mode 1) or G1 (mode 0) transform over n values, butterfly
 * size 2h, twiddle plans / scalars for k = 0..h-1 (index 0 unused). */
int oracle_cyc_stage(void* v, size_t n, size_t h, const void* plans) {
  fp12* a = (fp12*)v;
  #pragma omp parallel for schedule(dynamic, 1)
  for (size_t t = 0; t < n / 2; ++t) {
    size_t k = t % h, i = (t / h) * 2 * h + k;
    fp12 u = a[i], w = a[i + h];
    if (k) w = cyc_pow_plan(w, ((const cyc_plan_t*)plans) + k);
    a[i] = fp12_mul(u, w);
    a[i + h] = fp12_mul(u, fp12_conj(w));
  }
  return 0;
}
int oracle_g1_stage(void* v, size_t n, size_t h, const uint64_t* tw /* h x 4 limbs */) {
  g1_jac* a = (g1_jac*)v;
  #pragma omp parallel for schedule(dynamic, 1)
  for (size_t t = 0; t < n / 2; ++t) {
    size_t k = t % h, i = (t / h) * 2 * h + k;
    g1_jac u = a[i], w = a[i + h];
    if (k) w = g1_mul_scalar(w, tw + 4 * k);
    a[i] = g1_add(u, w);
    w.y = fp_neg(w.y);
    a[i + h] = g1_add(u, w);
  }
  return 0;
}

/* ---- exported (ctypes) ------------------------------------------------------------------ */
/* out[i] = (i + 1) * G1 for i < n, affine (inputs for timing samples and closed-form checks). */
int oracle_g1_consecutive_multiples(const void* gen_affine, size_t n, void* out) {
  g1_aff g = *(const g1_aff*)gen_affine;
  g1_jac* tmp = (g1_jac*)malloc(sizeof(g1_jac) * n);
  g1_jac acc = g1_identity();
  for (size_t i = 0; i < n; ++i) {
    acc = g1_add_aff(acc, g);
    tmp[i] = acc;
  }
  g1_aff* o = (g1_aff*)out;
  const size_t CH = 4096;
  #pragma omp parallel for schedule(static)
  for (size_t s = 0; s < n; s += CH) {
    size_t e = s + CH < n ? s + CH : n;
    fp* pre = (fp*)malloc(sizeof(fp) * (e - s));
    fp prod = ONE;
    for (size_t i = s; i < e; ++i) { pre[i - s] = prod; prod = fp_mul(prod, tmp[i].z); }
    fp inv = fp_inv(prod);
    for (size_t i = e; i-- > s;) {
      fp zi = fp_mul(inv, pre[i - s]);
      inv = fp_mul(inv, tmp[i].z);
      fp zi2 = fp_mul(zi, zi);
      o[i].x = fp_mul(tmp[i].x, zi2);
      o[i].y = fp_mul(tmp[i].y, fp_mul(zi2, zi));
    }
    free(pre);
  }
  free(tmp);
  return 0;
}

int oracle_g1_msm(const void* points, const uint8_t* scalars, size_t n, int c, void* out) {
  if (c <= 0) c = default_window(n);
  g1_msm((const g1_aff*)points, scalars, n, c, (g1_jac*)out);
  return 0;
}

int oracle_g2_msm(const void* points, const uint8_t* scalars, size_t n, int c, void* out) {
  if (c <= 0) c = default_window(n);
  g2_msm((const g2_aff*)points, scalars, n, c, (g2_jac*)out);
  return 0;
}

int oracle_fp_mul(const void* a, const void* b, void* out) {
  *(fp*)out = fp_mul(*(const fp*)a, *(const fp*)b);
  return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Thread count of the OpenMP regions above (0: all processors), for the one-thread rows of the
 * CPU baseline. */
int oracle_set_threads(int n) {
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  extern int omp_get_num_procs(void);
  omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
  (void)n;
#endif
  return 0;
}
