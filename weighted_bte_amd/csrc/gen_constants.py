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
print("// Standard generators (affine, Montgomery form): G1 {x, y}; G2 {x.c0, x.c1, y.c0, y.c1}.")
print("__device__ __constant__ const uint32_t G1_GEN[2][12] = {{%s},\n  {%s}};" % (limbs(G1[0]), limbs(G1[1])))
print("__device__ __constant__ const uint32_t G2_GEN[4][12] = {{%s},\n  {%s},\n  {%s},\n  {%s}};" % (
    limbs(G2[0][0]), limbs(G2[0][1]), limbs(G2[1][0]), limbs(G2[1][1])))
print("}  // namespace bte")
