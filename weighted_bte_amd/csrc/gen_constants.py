"""Generates constants.hpp: Fp12 Frobenius coefficients for the BLS12-381 tower used by the
device code (Fp2 = Fp[u]/(u^2+1), Fp6 = Fp2[v]/(v^3-(u+1)), Fp12 = Fp6[w]/(w^2-v)).

Writing an Fp12 element as sum_i g_i w^i (g_i in Fp2), x^(p^k) = sum_i frob_k(g_i) * gamma_{k,i} w^i
with gamma_{k,i} = (u+1)^(i (p^k - 1)/6).  Values are emitted in Montgomery form (R = 2^384) as
12 little-endian 32-bit limbs.  Run: python gen_constants.py > constants.hpp
"""
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
MONT = pow(2, 384, P)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_pow(a, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = f2_mul(r, a)
        a = f2_mul(a, a)
        e >>= 1
    return r


def limbs(x):
    m = x * MONT % P
    return ", ".join("0x%08xu" % ((m >> (32 * i)) & 0xFFFFFFFF) for i in range(12))


print("// Generated by gen_constants.py; do not edit.")
print("#pragma once")
print("#include <stdint.h>")
print("namespace bte {")
print("// FROB_GAMMA[k-1][i] = (u+1)^(i (p^k - 1) / 6), Montgomery form, {c0, c1}.")
print("__device__ __constant__ const uint32_t FROB_GAMMA[3][6][2][12] = {")
for k in (1, 2, 3):
    print("  {")
    for i in range(6):
        g = f2_pow((1, 1), i * (P**k - 1) // 6)
        print("    {{%s},\n     {%s}}," % (limbs(g[0]), limbs(g[1])))
    print("  },")
print("};")
G1 = (0x17F1D3A73197D7942695638C4FA9AC0FC3688C4F9774B905A14E3A3F171BAC586C55E83FF97A1AEFFB3AF00ADB22C6BB,
      0x08B3F481E3AAA0F1A09E30ED741D8AE4FCF5E095D5D00AF600DB18CB2C04B3EDD03CC744A2888AE40CAA232946C5E7E1)
G2 = ((0x024AA2B2F08F0A91260805272DC51051C6E47AD4FA403B02B4510B647AE3D1770BAC0326A805BBEFD48056C8C121BDB8,
       0x13E02B6052719F607DACD3A088274F65596BD0D09920B61AB5DA61BBDC7F5049334CF11213945D57E5AC7D055D042B7E),
      (0x0CE5D527727D6E118CC9CDC6DA2E351AADFD9BAA8CBDD3A76D429A695160D12C923AC9CC3BACA289E193548608B82801,
       0x0606C4A02EA734CC32ACD2B02BC28B99CB3E287E85A763AF267492AB572E99AB3F370D275CEC1DA1AAA9075FF05F79BE))
R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
Z_ABS = 0xD201000000010000  # the BLS parameter is -Z_ABS
LAMBDA = Z_ABS * Z_ABS - 1  # lambda^2 + lambda + 1 = r


def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_inv(a):
    n = pow((a[0] * a[0] + a[1] * a[1]) % P, P - 2, P)
    return (a[0] * n % P, -a[1] * n % P)


def aff_mul(pt, k, add, sub, mul, inv, zero, three, two):
    """Affine double-and-add over a field given by its operations (None = identity)."""
    def padd(a, b):
        if a is None:
            return b
        if b is None:
            return a
        if a[0] == b[0]:
            if a[1] != b[1] or a[1] == zero:
                return None
            lam = mul(mul(three, mul(a[0], a[0])), inv(mul(two, a[1])))
        else:
            lam = mul(sub(b[1], a[1]), inv(sub(b[0], a[0])))
        x = sub(sub(mul(lam, lam), a[0]), b[0])
        return (x, sub(mul(lam, sub(a[0], x)), a[1]))
    acc = None
    while k:
        if k & 1:
            acc = padd(acc, pt)
        pt = padd(pt, pt)
        k >>= 1
    return acc


assert (LAMBDA * LAMBDA + LAMBDA + 1) == R
# G1 endomorphism (x, y) -> (beta x, y) acting as multiplication by LAMBDA
lg = aff_mul(G1, LAMBDA, lambda a, b: (a + b) % P, lambda a, b: (a - b) % P, lambda a, b: a * b % P,
             lambda a: pow(a, P - 2, P), 0, 3, 2)
BETA = None
for g in range(2, 50):
    b = pow(g, (P - 1) // 3, P)
    if b == 1:
        continue
    for cand in (b, b * b % P):
        if cand * G1[0] % P == lg[0] and G1[1] == lg[1]:
            BETA = cand
    if BETA:
        break
assert BETA is not None
# G2 endomorphism psi(x, y) = (conj(x) cx, conj(y) cy) on the twist, acting as multiplication by
# p mod r = -Z_ABS
CX = f2_inv(f2_pow((1, 1), (P - 1) // 3))
CY = f2_inv(f2_pow((1, 1), (P - 1) // 2))
psi_g2 = (f2_mul((G2[0][0], -G2[0][1] % P), CX), f2_mul((G2[1][0], -G2[1][1] % P), CY))
zg2 = aff_mul(G2, R - Z_ABS, f2_add, f2_sub, f2_mul, f2_inv, (0, 0), (3, 0), (2, 0))
assert psi_g2 == zg2 and P % R == R - Z_ABS
print("// G1 endomorphism: (x, y) -> (BETA x, y) is multiplication by lambda = |z|^2 - 1.")
print("__device__ __constant__ const uint32_t G1_BETA[12] = {%s};" % limbs(BETA))
print("// G2 endomorphism psi(x, y) = (conj(x) PSI[0], conj(y) PSI[1]) is multiplication by z = -|z|.")
print("__device__ __constant__ const uint32_t G2_PSI[2][2][12] = {{{%s},\n   {%s}},\n  {{%s},\n   {%s}}};" % (
    limbs(CX[0]), limbs(CX[1]), limbs(CY[0]), limbs(CY[1])))
print("// Standard generators (affine, Montgomery form): G1 {x, y}; G2 {x.c0, x.c1, y.c0, y.c1}.")
print("__device__ __constant__ const uint32_t G1_GEN[2][12] = {{%s},\n  {%s}};" % (limbs(G1[0]), limbs(G1[1])))
print("__device__ __constant__ const uint32_t G2_GEN[4][12] = {{%s},\n  {%s},\n  {%s},\n  {%s}};" % (
    limbs(G2[0][0]), limbs(G2[0][1]), limbs(G2[1][0]), limbs(G2[1][1])))
print("}  // namespace bte")
