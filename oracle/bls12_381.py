"""CPU oracle: a from-scratch big-integer restatement of the BLS12-381 arithmetic that the
weighted_btx / weighted_pfe batch-decryption path runs through blst/blstrs.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (weighted_bte_amd/) may import or call this
module; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
the checker.

Every function below names the reference code whose semantics it restates.  The arithmetic of
blst 0.3.17 / blstrs 0.7.1 itself is not vendored in the reference, so the field tower, curve
and pairing formulas are the textbook BLS12-381 definitions; each is pinned in
tests/test_oracle.py by published constants (p, r, x, generator coordinates and their ZCash
compressed encodings), closed forms and invariants:

  Fp/Fp2/Fp6/Fp12 tower ........ pinned: Frobenius == x^p, inverses, generator on-curve checks
  G1/G2 group law, compress .... pinned: generator KATs, r*G == O, naive vs Jacobian
  MSM .......................... pinned: closed form sum(k_i s_i)*G, reference test shape
                                   (weighted_btx/src/blst_utils.rs:203-249)
  Radix2Domain FFT ............. pinned: naive O(n^2) DFT, ifft(fft(x)) == x
  Miller loop + final exp ...... pinned: bilinearity, e^r == 1, non-degeneracy, hard chain
                                   == f^(3*Lambda) with Lambda == HARD_EXPONENT
                                   (weighted_btx/src/fft.rs:660-681)
  Decomposed GT scalar action .. pinned: equals Gt^s on r-torsion inputs
                                   (weighted_btx/src/fft.rs:811-846)
  GT absolute value ............ pinned: e^c == t(Q,P)^(3L) against a reduced Tate pairing from
                                   an independent Miller loop over r (Hess-Smart-Vercauteren
                                   ate relation with T = x < 0; tests/test_oracle_gt_sign.py),
                                   which fixes the conjugation for negative x.  Still unpinned:
                                   that blst's un-vendored miller_loop_n makes the same (textbook)
                                   choice; no reference test holds a GT value.
"""

from __future__ import annotations

import hashlib
import struct

# ----------------------------------------------------------------------------------------------
# Parameters (published BLS12-381 constants).
# ----------------------------------------------------------------------------------------------
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
X_ABS = 0xD201000000010000  # weighted_btx/src/fft.rs:22 (BLS12_381_X_ABS); x = -X_ABS
X = -X_ABS
MONT_R = pow(2, 384, P)  # blst Montgomery radix for Fp (6x64 == 12x32 limbs)
MONT_R_INV = pow(MONT_R, -1, P)
FR_S = 32
FR_ROOT_OF_UNITY = 0x16A2A19EDFE81F20D09B681922C813B4B63683508C2280B93829971F439F0D2B  # 7^((r-1)/2^32)

# Lambda = (p^4 - p^2 + 1) / r, the hard exponent (weighted_btx/src/fft.rs:660-681).
HARD_EXPONENT = (P**4 - P**2 + 1) // R

G1_X = 0x17F1D3A73197D7942695638C4FA9AC0FC3688C4F9774B905A14E3A3F171BAC586C55E83FF97A1AEFFB3AF00ADB22C6BB
G1_Y = 0x08B3F481E3AAA0F1A09E30ED741D8AE4FCF5E095D5D00AF600DB18CB2C04B3EDD03CC744A2888AE40CAA232946C5E7E1
G2_X = (0x024AA2B2F08F0A91260805272DC51051C6E47AD4FA403B02B4510B647AE3D1770BAC0326A805BBEFD48056C8C121BDB8,
        0x13E02B6052719F607DACD3A088274F65596BD0D09920B61AB5DA61BBDC7F5049334CF11213945D57E5AC7D055D042B7E)
G2_Y = (0x0CE5D527727D6E118CC9CDC6DA2E351AADFD9BAA8CBDD3A76D429A695160D12C923AC9CC3BACA289E193548608B82801,
        0x0606C4A02EA734CC32ACD2B02BC28B99CB3E287E85A763AF267492AB572E99AB3F370D275CEC1DA1AAA9075FF05F79BE)

# Fp-multiplication counter for the algorithmic yardstick (counts only when enabled).
FP_MUL_COUNT = [0]


# ----------------------------------------------------------------------------------------------
# Fp2 = Fp[u]/(u^2+1); Fp6 = Fp2[v]/(v^3 - (u+1)); Fp12 = Fp6[w]/(w^2 - v).
# Elements are plain tuples of canonical ints.
# ----------------------------------------------------------------------------------------------
F2_ZERO = (0, 0)
F2_ONE = (1, 0)


def f2(a, b=0):
    return (a % P, b % P)


def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_neg(a):
    return ((-a[0]) % P, (-a[1]) % P)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_sqr(a):
    return f2_mul(a, a)


def f2_muls(a, s):
    return ((a[0] * s) % P, (a[1] * s) % P)


def f2_conj(a):
    return (a[0], (-a[1]) % P)


def f2_inv(a):
    n = (a[0] * a[0] + a[1] * a[1]) % P
    ni = pow(n, -1, P)
    return ((a[0] * ni) % P, (-a[1] * ni) % P)


def f2_mul_xi(a):  # multiply by xi = 1 + u
    return ((a[0] - a[1]) % P, (a[0] + a[1]) % P)


def f2_pow(a, e):
    r = F2_ONE
    while e:
        if e & 1:
            r = f2_mul(r, a)
        a = f2_sqr(a)
        e >>= 1
    return r


def f2_is_zero(a):
    return a[0] == 0 and a[1] == 0


def fp_sqrt(a):
    """Square root in Fp (p = 3 mod 4); returns None when a is a non-residue."""
    s = pow(a, (P + 1) // 4, P)
    return s if (s * s) % P == a % P else None


def f2_sqrt(a):
    """Square root in Fp2 (complex method); returns None for non-squares."""
    if f2_is_zero(a):
        return F2_ZERO
    a0, a1 = a
    if a1 == 0:
        s = fp_sqrt(a0)
        if s is not None:
            return (s, 0)
        s = fp_sqrt((-a0) % P)
        return (0, s)
    alpha = (a0 * a0 + a1 * a1) % P
    gamma = fp_sqrt(alpha)
    if gamma is None:
        return None
    inv2 = pow(2, -1, P)
    for g in (gamma, (-gamma) % P):
        delta = ((a0 + g) * inv2) % P
        x0 = fp_sqrt(delta)
        if x0 is not None and x0 != 0:
            x1 = (a1 * pow(2 * x0, -1, P)) % P
            cand = (x0, x1)
            if f2_sqr(cand) == (a0 % P, a1 % P):
                return cand
    return None


F6_ZERO = (F2_ZERO, F2_ZERO, F2_ZERO)
F6_ONE = (F2_ONE, F2_ZERO, F2_ZERO)


def f6_add(a, b):
    return (f2_add(a[0], b[0]), f2_add(a[1], b[1]), f2_add(a[2], b[2]))


def f6_sub(a, b):
    return (f2_sub(a[0], b[0]), f2_sub(a[1], b[1]), f2_sub(a[2], b[2]))


def f6_neg(a):
    return (f2_neg(a[0]), f2_neg(a[1]), f2_neg(a[2]))


def f6_mul(a, b):
    a0, a1, a2 = a
    b0, b1, b2 = b
    c0 = f2_add(f2_mul(a0, b0), f2_mul_xi(f2_add(f2_mul(a1, b2), f2_mul(a2, b1))))
    c1 = f2_add(f2_add(f2_mul(a0, b1), f2_mul(a1, b0)), f2_mul_xi(f2_mul(a2, b2)))
    c2 = f2_add(f2_add(f2_mul(a0, b2), f2_mul(a1, b1)), f2_mul(a2, b0))
    return (c0, c1, c2)


def f6_mul_v(a):  # multiply by v
    return (f2_mul_xi(a[2]), a[0], a[1])


def f6_inv(a):
    a0, a1, a2 = a
    t0 = f2_sub(f2_sqr(a0), f2_mul_xi(f2_mul(a1, a2)))
    t1 = f2_sub(f2_mul_xi(f2_sqr(a2)), f2_mul(a0, a1))
    t2 = f2_sub(f2_sqr(a1), f2_mul(a0, a2))
    den = f2_add(f2_mul(a0, t0), f2_mul_xi(f2_add(f2_mul(a2, t1), f2_mul(a1, t2))))
    di = f2_inv(den)
    return (f2_mul(t0, di), f2_mul(t1, di), f2_mul(t2, di))


F12_ONE = (F6_ONE, F6_ZERO)


def f12_add(a, b):
    return (f6_add(a[0], b[0]), f6_add(a[1], b[1]))


def f12_mul(a, b):
    a0, a1 = a
    b0, b1 = b
    t0 = f6_mul(a0, b0)
    t1 = f6_mul(a1, b1)
    c0 = f6_add(t0, f6_mul_v(t1))
    c1 = f6_add(f6_mul(a0, b1), f6_mul(a1, b0))
    return (c0, c1)


def f12_sqr(a):
    return f12_mul(a, a)


def f12_conj(a):
    return (a[0], f6_neg(a[1]))


def f12_inv(a):
    a0, a1 = a
    den = f6_sub(f6_mul(a0, a0), f6_mul_v(f6_mul(a1, a1)))
    di = f6_inv(den)
    return (f6_mul(a0, di), f6_neg(f6_mul(a1, di)))


def f12_pow(a, e):
    if e < 0:
        a = f12_inv(a)
        e = -e
    r = F12_ONE
    while e:
        if e & 1:
            r = f12_mul(r, a)
        a = f12_sqr(a)
        e >>= 1
    return r


def f12_is_one(a):
    return a == F12_ONE


def f12_eq(a, b):
    return a == b


# Frobenius: write x = sum_i g_i w^i (g_i in Fp2, i = 0..5) with the tower mapping
# c0 = (g0, g2, g4), c1 = (g1, g3, g5).  Then x^(p^k) = sum_i g_i^(p^k) * xi^(i (p^k - 1)/6) w^i.
def _frob_coeffs(k):
    e = (P**k - 1) // 6
    return [f2_pow((1, 1), i * e) for i in range(6)]


FROB_COEFFS = {k: _frob_coeffs(k) for k in (1, 2, 3)}


def _f12_to_w(a):
    (g0, g2, g4), (g1, g3, g5) = a
    return [g0, g1, g2, g3, g4, g5]


def _w_to_f12(g):
    return ((g[0], g[2], g[4]), (g[1], g[3], g[5]))


def f12_frobenius(a, k):
    """x -> x^(p^k) for k in 1..3 (blst_fp12_frobenius_map semantics; final_exponentiation.rs:418-428)."""
    assert 1 <= k <= 3
    g = _f12_to_w(a)
    c = FROB_COEFFS[k]
    out = []
    for i in range(6):
        gi = f2_conj(g[i]) if (k & 1) else g[i]
        out.append(f2_mul(gi, c[i]))
    return _w_to_f12(out)


def f12_cyclotomic_sqr(a):
    """blst_fp12_cyclotomic_sqr restated as the plain square (identical values on the cyclotomic
    subgroup, the only domain the reference calls it on)."""
    return f12_sqr(a)


# ----------------------------------------------------------------------------------------------
# Curves.  G1: y^2 = x^3 + 4 over Fp.  G2 (M-type twist): y^2 = x^3 + 4(u+1) over Fp2.
# Jacobian points are (X, Y, Z), Z == 0 means the identity.  Affine points are (x, y) or None.
# ----------------------------------------------------------------------------------------------
class _FpOps:
    zero = 0
    one = 1
    b = 4

    @staticmethod
    def add(a, b):
        return (a + b) % P

    @staticmethod
    def sub(a, b):
        return (a - b) % P

    @staticmethod
    def mul(a, b):
        return (a * b) % P

    @staticmethod
    def neg(a):
        return (-a) % P

    @staticmethod
    def inv(a):
        return pow(a, -1, P)

    @staticmethod
    def is_zero(a):
        return a % P == 0

    @staticmethod
    def muls(a, s):
        return (a * s) % P


class _Fp2Ops:
    zero = F2_ZERO
    one = F2_ONE
    b = (4, 4)
    add = staticmethod(f2_add)
    sub = staticmethod(f2_sub)
    mul = staticmethod(f2_mul)
    neg = staticmethod(f2_neg)
    inv = staticmethod(f2_inv)
    is_zero = staticmethod(f2_is_zero)
    muls = staticmethod(f2_muls)


class Curve:
    def __init__(self, F, name):
        self.F = F
        self.name = name

    # --- Jacobian arithmetic (dbl-2009-l / add-2007-bl style; results compared only in affine)
    def identity(self):
        F = self.F
        return (F.one, F.one, F.zero)

    def is_identity(self, p):
        return self.F.is_zero(p[2])

    def from_affine(self, a):
        if a is None:
            return self.identity()
        return (a[0], a[1], self.F.one)

    def to_affine(self, p):
        F = self.F
        if self.is_identity(p):
            return None
        zi = F.inv(p[2])
        zi2 = F.mul(zi, zi)
        return (F.mul(p[0], zi2), F.mul(p[1], F.mul(zi2, zi)))

    def double(self, p):
        F = self.F
        X1, Y1, Z1 = p
        if F.is_zero(Z1) or F.is_zero(Y1):
            return self.identity()
        A = F.mul(X1, X1)
        B = F.mul(Y1, Y1)
        C = F.mul(B, B)
        D = F.muls(F.sub(F.mul(F.add(X1, B), F.add(X1, B)), F.add(A, C)), 2)
        E = F.muls(A, 3)
        Fv = F.mul(E, E)
        X3 = F.sub(Fv, F.muls(D, 2))
        Y3 = F.sub(F.mul(E, F.sub(D, X3)), F.muls(C, 8))
        Z3 = F.muls(F.mul(Y1, Z1), 2)
        return (X3, Y3, Z3)

    def add(self, p, q):
        F = self.F
        if self.is_identity(p):
            return q
        if self.is_identity(q):
            return p
        X1, Y1, Z1 = p
        X2, Y2, Z2 = q
        Z1Z1 = F.mul(Z1, Z1)
        Z2Z2 = F.mul(Z2, Z2)
        U1 = F.mul(X1, Z2Z2)
        U2 = F.mul(X2, Z1Z1)
        S1 = F.mul(F.mul(Y1, Z2), Z2Z2)
        S2 = F.mul(F.mul(Y2, Z1), Z1Z1)
        H = F.sub(U2, U1)
        Rr = F.sub(S2, S1)
        if F.is_zero(H):
            if F.is_zero(Rr):
                return self.double(p)
            return self.identity()
        HH = F.mul(H, H)
        HHH = F.mul(H, HH)
        V = F.mul(U1, HH)
        X3 = F.sub(F.sub(F.mul(Rr, Rr), HHH), F.muls(V, 2))
        Y3 = F.sub(F.mul(Rr, F.sub(V, X3)), F.mul(S1, HHH))
        Z3 = F.mul(F.mul(Z1, Z2), H)
        return (X3, Y3, Z3)

    def neg(self, p):
        return (p[0], self.F.neg(p[1]), p[2])

    def sub(self, p, q):
        return self.add(p, self.neg(q))

    def mul(self, p, k):
        """Scalar multiplication by an integer (any sign)."""
        if k < 0:
            p = self.neg(p)
            k = -k
        acc = self.identity()
        for bit in bin(k)[2:] if k else "":
            acc = self.double(acc)
            if bit == "1":
                acc = self.add(acc, p)
        return acc

    def on_curve_affine(self, a):
        F = self.F
        if a is None:
            return True
        x, y = a
        return F.sub(F.mul(y, y), F.add(F.mul(F.mul(x, x), x), F.b)) == F.zero

    def eq(self, p, q):
        return self.to_affine(p) == self.to_affine(q)


G1 = Curve(_FpOps, "G1")
G2 = Curve(_Fp2Ops, "G2")
G1_GEN = (G1_X, G1_Y)
G2_GEN = (G2_X, G2_Y)


def g1_gen_jac():
    return G1.from_affine(G1_GEN)


def g2_gen_jac():
    return G2.from_affine(G2_GEN)


# ----------------------------------------------------------------------------------------------
# ZCash serialisation (G1 48 B, G2 96 B; flags 0x80 compressed, 0x40 infinity, 0x20 y-sign).
# ----------------------------------------------------------------------------------------------
def _fp_be(x):
    return x.to_bytes(48, "big")


def g1_compress(a):
    if a is None:
        return bytes([0xC0]) + bytes(47)
    x, y = a
    out = bytearray(_fp_be(x))
    out[0] |= 0x80
    if y > (P - 1) // 2:
        out[0] |= 0x20
    return bytes(out)


def g2_compress(a):
    if a is None:
        return bytes([0xC0]) + bytes(95)
    (x0, x1), (y0, y1) = a
    out = bytearray(_fp_be(x1) + _fp_be(x0))
    out[0] |= 0x80
    big = (y1 > (P - 1) // 2) if y1 != 0 else (y0 > (P - 1) // 2)
    if big:
        out[0] |= 0x20
    return bytes(out)


def g1_decompress(b):
    flags = b[0]
    assert flags & 0x80
    if flags & 0x40:
        return None
    x = int.from_bytes(bytes([flags & 0x1F]) + b[1:48], "big")
    y = fp_sqrt((x * x * x + 4) % P)
    assert y is not None
    if (y > (P - 1) // 2) != bool(flags & 0x20):
        y = (-y) % P
    return (x, y)


# ----------------------------------------------------------------------------------------------
# blst memory layouts (Montgomery form, little-endian 64-bit limbs), used at the C ABI.
# ----------------------------------------------------------------------------------------------
def fp_to_mont_bytes(x):
    return ((x * MONT_R) % P).to_bytes(48, "little")


def fp_from_mont_bytes(b):
    return (int.from_bytes(b[:48], "little") * MONT_R_INV) % P


def f2_to_bytes(a):
    return fp_to_mont_bytes(a[0]) + fp_to_mont_bytes(a[1])


def f2_from_bytes(b):
    return (fp_from_mont_bytes(b[0:48]), fp_from_mont_bytes(b[48:96]))


def g1_affine_to_bytes(a):
    """blst_p1_affine: x || y (the identity is all-zero)."""
    if a is None:
        return bytes(96)
    return fp_to_mont_bytes(a[0]) + fp_to_mont_bytes(a[1])


def g1_affine_from_bytes(b):
    if b[:96] == bytes(96):
        return None
    return (fp_from_mont_bytes(b[0:48]), fp_from_mont_bytes(b[48:96]))


def g1_jac_from_bytes(b):
    """blst_p1: X || Y || Z (Z == 0 is the identity)."""
    return (fp_from_mont_bytes(b[0:48]), fp_from_mont_bytes(b[48:96]), fp_from_mont_bytes(b[96:144]))


def g1_jac_to_bytes(p):
    return b"".join(fp_to_mont_bytes(c) for c in p)


def g2_affine_to_bytes(a):
    if a is None:
        return bytes(192)
    return f2_to_bytes(a[0]) + f2_to_bytes(a[1])


def g2_affine_from_bytes(b):
    if b[:192] == bytes(192):
        return None
    return (f2_from_bytes(b[0:96]), f2_from_bytes(b[96:192]))


def g2_jac_from_bytes(b):
    return (f2_from_bytes(b[0:96]), f2_from_bytes(b[96:192]), f2_from_bytes(b[192:288]))


def g2_jac_to_bytes(p):
    return b"".join(f2_to_bytes(c) for c in p)


def f12_to_bytes(a):
    """blst_fp12: fp6[2], fp6 = fp2[3], fp2 = fp[2] (Montgomery)."""
    return b"".join(f2_to_bytes(c) for f6 in a for c in f6)


def f12_from_bytes(b):
    cs = [f2_from_bytes(b[96 * i: 96 * (i + 1)]) for i in range(6)]
    return ((cs[0], cs[1], cs[2]), (cs[3], cs[4], cs[5]))


def scalar_to_le32(s):
    return (s % (1 << 256)).to_bytes(32, "little")


def scalar_from_le32(b):
    return int.from_bytes(b[:32], "little")


# ----------------------------------------------------------------------------------------------
# MSM (weighted_btx/src/blst_utils.rs:43-67, 88-152): sum s_i * P_i with the low 255 bits of each
# 32-byte little-endian scalar (blst_p{1,2}s_mult_pippenger with nbits = 255).  Empty -> identity.
# ----------------------------------------------------------------------------------------------
def msm(curve, points_affine, scalars):
    acc = curve.identity()
    mask = (1 << 255) - 1
    for pt, s in zip(points_affine, scalars):
        acc = curve.add(acc, curve.mul(curve.from_affine(pt), s & mask))
    return acc


def msm_indexed(curve, points_affine, indices, scalars):
    """g2_multi_exp_affine_indexed (blst_utils.rs:106-125): points gathered by index."""
    return msm(curve, [points_affine[i] for i in indices], scalars)


# ----------------------------------------------------------------------------------------------
# Fr radix-2 domain (weighted_btx/src/fft.rs:156-390).
# ----------------------------------------------------------------------------------------------
class Radix2Domain:
    """fft.rs:164-191: generator = ROOT_OF_UNITY^(2^(S - log n)); size must be a power of two."""

    def __init__(self, size):
        if size == 0 or size & (size - 1):
            raise ValueError("InvalidBatchSize(%d)" % size)
        log_size = size.bit_length() - 1
        if log_size > FR_S:
            raise ValueError("InvalidBatchSize(%d)" % size)
        g = FR_ROOT_OF_UNITY
        for _ in range(log_size, FR_S):
            g = g * g % R
        self.size = size
        self.log_size = log_size
        self.generator = g
        self.generator_inverse = pow(g, -1, R)
        self.size_inverse = pow(size, -1, R)


def bit_reverse_permute(values):
    """fft.rs:621-632."""
    n = len(values)
    if n <= 1:
        return list(values)
    bits = n.bit_length() - 1
    out = list(values)
    for i in range(n):
        j = int(bin(i)[2:].zfill(bits)[::-1], 2)
        if i < j:
            out[i], out[j] = out[j], out[i]
    return out


def stage_twiddles(n, generator):
    """Per-stage twiddle lists as fft.rs:357-390 walks them: stage of butterfly size 2h uses
    step = generator^(n/2h) and twiddles step^k for k = 0..h-1 (k = 0 is the trivial 1)."""
    stages = []
    size = 2
    while size <= n:
        h = size // 2
        step = pow(generator, n // size, R)
        tw = [1]
        t = 1
        for _ in range(1, h):
            t = t * step % R
            tw.append(t)
        stages.append(tw)
        size *= 2
    return stages


def fft_generic(values, generator, add, sub, smul):
    """fft_with_generator (fft.rs:357-390): bit reverse then radix-2 DIT."""
    vals = bit_reverse_permute(values)
    n = len(vals)
    for tw in stage_twiddles(n, generator):
        h = len(tw)
        size = 2 * h
        for start in range(0, n, size):
            for k in range(h):
                u = vals[start + k]
                v = vals[start + k + h] if k == 0 else smul(vals[start + k + h], tw[k])
                vals[start + k] = add(u, v)
                vals[start + k + h] = sub(u, v)
    return vals


def fr_fft(values, inverse=False, scaled=True):
    d = Radix2Domain(len(values))
    g = d.generator_inverse if inverse else d.generator
    out = fft_generic(values, g, lambda a, b: (a + b) % R, lambda a, b: (a - b) % R, lambda a, t: a * t % R)
    if inverse and scaled:
        out = [v * d.size_inverse % R for v in out]
    return out


def fr_dft_naive(values, inverse=False):
    d = Radix2Domain(len(values))
    g = d.generator_inverse if inverse else d.generator
    n = len(values)
    return [sum(values[j] * pow(g, j * k, R) for j in range(n)) % R for k in range(n)]


def group_fft(curve, values_jac, inverse=False, scaled=True):
    """Radix2Domain::fft / ifft over G1 or G2 (fft.rs:197-234): same DIT, scalar mult twiddles."""
    d = Radix2Domain(len(values_jac))
    g = d.generator_inverse if inverse else d.generator
    out = fft_generic(values_jac, g, curve.add, curve.sub, curve.mul)
    if inverse and scaled:
        out = [curve.mul(v, d.size_inverse) for v in out]
    return out


# ----------------------------------------------------------------------------------------------
# Pairing.  Optimal ate Miller loop over |x| with affine lines on the twist (values are only
# compared after the final exponentiation, which removes subfield factors), conjugated at the
# end because x < 0.  Final exponentiation split as in weighted_btx/src/final_exponentiation.rs.
# ----------------------------------------------------------------------------------------------
def _line(lam, xt, yt, p):
    """Line with slope lam (Fp2, on the twist) through (xt, yt), evaluated at P = (xp, yp) in G1,
    scaled by w^3: (lam*xt - yt) + (-lam*xp) v + yp v w."""
    xp, yp = p
    c00 = f2_sub(f2_mul(lam, xt), yt)
    c01 = f2_muls(f2_neg(lam), xp)
    return ((c00, c01, F2_ZERO), (F2_ZERO, (yp % P, 0), F2_ZERO))


def miller_loop(pairs):
    """Product of Miller loops f_{|x|,Q}(P) over (P in G1 affine, Q in G2 affine) pairs, then
    conjugated (x < 0).  Pairs containing an identity are skipped and an empty product is one,
    as unprepared_multi_miller_loop does (final_exponentiation.rs:229-273)."""
    pairs = [(p, q) for p, q in pairs if p is not None and q is not None]
    f = F12_ONE
    if not pairs:
        return f
    ts = [q for _, q in pairs]
    bits = bin(X_ABS)[3:]
    for bit in bits:
        f = f12_sqr(f)
        for i, (p, q) in enumerate(pairs):
            xt, yt = ts[i]
            lam = f2_mul(f2_muls(f2_sqr(xt), 3), f2_inv(f2_muls(yt, 2)))
            f = f12_mul(f, _line(lam, xt, yt, p))
            x3 = f2_sub(f2_sqr(lam), f2_muls(xt, 2))
            y3 = f2_sub(f2_mul(lam, f2_sub(xt, x3)), yt)
            ts[i] = (x3, y3)
        if bit == "1":
            for i, (p, q) in enumerate(pairs):
                xt, yt = ts[i]
                xq, yq = q
                lam = f2_mul(f2_sub(yq, yt), f2_inv(f2_sub(xq, xt)))
                f = f12_mul(f, _line(lam, xt, yt, p))
                x3 = f2_sub(f2_sub(f2_sqr(lam), xt), xq)
                y3 = f2_sub(f2_mul(lam, f2_sub(xt, x3)), yt)
                ts[i] = (x3, y3)
    return f12_conj(f)


def easy_final_exponentiation(f):
    """final_exponentiation.rs:356-371: f^((p^6-1)(p^2+1))."""
    r = f12_mul(f12_conj(f), f12_inv(f))
    return f12_mul(r, f12_frobenius(r, 2))


def batch_easy_final_exponentiation(fs):
    """final_exponentiation.rs:281-305 (one shared inversion; values equal the per-element form)."""
    return [easy_final_exponentiation(f) for f in fs]


def _multiply_then_square(acc, m, n):
    r = f12_mul(acc, m)
    for _ in range(n):
        r = f12_cyclotomic_sqr(r)
    return r


def raise_to_z_div_by_2(a):
    """final_exponentiation.rs:452-463."""
    r = f12_cyclotomic_sqr(a)
    r = _multiply_then_square(r, a, 2)
    r = _multiply_then_square(r, a, 3)
    r = _multiply_then_square(r, a, 9)
    r = _multiply_then_square(r, a, 32)
    r = _multiply_then_square(r, a, 15)
    return f12_conj(r)


def raise_to_z(a):
    """final_exponentiation.rs:465-467."""
    return f12_cyclotomic_sqr(raise_to_z_div_by_2(a))


def hard_final_exponentiation(f):
    """final_exponentiation.rs:308-336 (the blst chain, exponent 3*Lambda)."""
    y0 = f12_cyclotomic_sqr(f)
    y1 = raise_to_z(y0)
    y2 = raise_to_z_div_by_2(y1)
    y3 = f12_conj(f)
    y1 = f12_mul(y1, y3)
    y1 = f12_conj(y1)
    y1 = f12_mul(y1, y2)
    y2 = raise_to_z(y1)
    y3 = raise_to_z(y2)
    y1 = f12_conj(y1)
    y3 = f12_mul(y3, y1)
    y1 = f12_conj(y1)
    y1 = f12_frobenius(y1, 3)
    y2 = f12_frobenius(y2, 2)
    y1 = f12_mul(y1, y2)
    y2 = raise_to_z(y3)
    y2 = f12_mul(y2, y0)
    y2 = f12_mul(y2, f)
    y1 = f12_mul(y1, y2)
    y2 = f12_frobenius(y3, 1)
    return f12_mul(y1, y2)


def final_exponentiation(f):
    return hard_final_exponentiation(easy_final_exponentiation(f))


def pairing(p_aff, q_aff):
    return final_exponentiation(miller_loop([(p_aff, q_aff)]))


def multi_pairing(pairs):
    return final_exponentiation(miller_loop(pairs))


def gt_generator():
    return pairing(G1_GEN, G2_GEN)


def gt_pow(a, s):
    return f12_pow(a, s)


# ----------------------------------------------------------------------------------------------
# Decomposed cyclotomic scalar action (weighted_btx/src/fft.rs:33-154, 543-609).
# ----------------------------------------------------------------------------------------------
GT_WNAF_WIDTH = 4
GT_WNAF_DIGITS = 65


def decompose_scalar_base_x(s):
    """fft.rs:543-557: canonical scalar = sum d_i |x|^i, four 64-bit components."""
    s = s % R
    comps = []
    for _ in range(4):
        comps.append(s % X_ABS)
        s //= X_ABS
    assert s == 0
    return comps


def width_w_naf_u64(value):
    """fft.rs:569-594: width-4 NAF, at most 65 digits; returns (digits, count)."""
    digits = [0] * GT_WNAF_DIGITS
    pos = 0
    while value:
        if value & 1:
            d = value & ((1 << GT_WNAF_WIDTH) - 1)
            if d >= (1 << (GT_WNAF_WIDTH - 1)):
                d -= 1 << GT_WNAF_WIDTH
            digits[pos] = d
            value -= d
        value >>= 1
        pos += 1
    return digits, pos


def cyclotomic_scalar_plan(s):
    comps = decompose_scalar_base_x(s)
    nafs = []
    count = 0
    for c in comps:
        d, n = width_w_naf_u64(c)
        nafs.append(d)
        count = max(count, n)
    return nafs, count


def cyclotomic_mul_scalar_with_plan(value, plan):
    """fft.rs:74-154, restated step for step (table of odd multiples, Frobenius/conjugate tables,
    interleaved NAF evaluation sharing one squaring chain)."""
    nafs, count = plan
    if count == 0:
        return F12_ONE
    two = f12_cyclotomic_sqr(value)
    base_table = [value]
    for _ in range(1, 1 << (GT_WNAF_WIDTH - 2)):
        base_table.append(f12_mul(base_table[-1], two))
    tables = [base_table]
    for power in range(1, 4):
        t = []
        for e in base_table:
            fe = f12_frobenius(e, power)
            if power & 1:
                fe = f12_conj(fe)
            t.append(fe)
        tables.append(t)
    acc = F12_ONE
    started = False
    for pos in range(count - 1, -1, -1):
        if started:
            acc = f12_cyclotomic_sqr(acc)
        for comp in range(4):
            d = nafs[comp][pos]
            if d == 0:
                continue
            m = tables[comp][(abs(d) - 1) // 2]
            if d < 0:
                m = f12_conj(m)
            if started:
                acc = f12_mul(acc, m)
            else:
                acc = m
                started = True
    return acc


def cyclotomic_mul_scalar_decomposed(value, s):
    return cyclotomic_mul_scalar_with_plan(value, cyclotomic_scalar_plan(s))


def gt_fft(values, inverse=False, unscaled=True):
    """Radix2Domain over Gt (fft.rs:197-234 generic path; 242-274 for ifft_gt_unscaled): group
    operation is Fp12 multiplication, group 'scalar mult' is exponentiation."""
    d = Radix2Domain(len(values))
    g = d.generator_inverse if inverse else d.generator
    out = fft_generic(values, g, f12_mul, lambda a, b: f12_mul(a, f12_conj(b)), gt_pow)
    if inverse and not unscaled:
        out = [gt_pow(v, d.size_inverse) for v in out]
    return out


def ifft_cyclotomic_prefix_unscaled(values, output_count):
    """fft.rs:284-350: unscaled inverse transform over the cyclotomic subgroup keeping only the
    first output_count outputs of the last stage, twiddles applied by the decomposed action."""
    d = Radix2Domain(len(values))
    n = len(values)
    assert output_count <= n // 2
    vals = bit_reverse_permute(values)
    stages = stage_twiddles(n, d.generator_inverse)
    for tw in stages[:-1]:
        h = len(tw)
        for start in range(0, n, 2 * h):
            for k in range(h):
                u = vals[start + k]
                v = vals[start + k + h]
                if k:
                    v = cyclotomic_mul_scalar_decomposed(v, tw[k])
                vals[start + k] = f12_mul(u, v)
                vals[start + k + h] = f12_mul(u, f12_conj(v))
    half = n // 2
    out = vals[:half]
    if output_count == 0:
        return out
    out[0] = f12_mul(vals[0], vals[half])
    tw = stages[-1]
    for k in range(1, output_count):
        v = cyclotomic_mul_scalar_decomposed(vals[half + k], tw[k])
        out[k] = f12_mul(vals[k], v)
    return out[:output_count]


# ----------------------------------------------------------------------------------------------
# Seeded input generation shared by the oracle, the tests and the bench (the build's own rule;
# the reference samples from OsRng).  ChaCha20 keystream (RFC 8439 block function) per
# (seed, stream id); scalars are 64-byte little-endian wide reductions mod r.
# ----------------------------------------------------------------------------------------------
def _rotl(v, c):
    return ((v << c) & 0xFFFFFFFF) | (v >> (32 - c))


def _qr(s, a, b, c, d):
    s[a] = (s[a] + s[b]) & 0xFFFFFFFF; s[d] ^= s[a]; s[d] = _rotl(s[d], 16)
    s[c] = (s[c] + s[d]) & 0xFFFFFFFF; s[b] ^= s[c]; s[b] = _rotl(s[b], 12)
    s[a] = (s[a] + s[b]) & 0xFFFFFFFF; s[d] ^= s[a]; s[d] = _rotl(s[d], 8)
    s[c] = (s[c] + s[d]) & 0xFFFFFFFF; s[b] ^= s[c]; s[b] = _rotl(s[b], 7)


def chacha20_block(key32, counter, nonce12):
    const = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    k = list(struct.unpack("<8I", key32))
    n = list(struct.unpack("<3I", nonce12))
    st = const + k + [counter & 0xFFFFFFFF] + n
    w = list(st)
    for _ in range(10):
        _qr(w, 0, 4, 8, 12); _qr(w, 1, 5, 9, 13); _qr(w, 2, 6, 10, 14); _qr(w, 3, 7, 11, 15)
        _qr(w, 0, 5, 10, 15); _qr(w, 1, 6, 11, 12); _qr(w, 2, 7, 8, 13); _qr(w, 3, 4, 9, 14)
    return struct.pack("<16I", *[(w[i] + st[i]) & 0xFFFFFFFF for i in range(16)])


class SeededRng:
    """ChaCha20 keystream keyed by SHA-256(seed || stream id)."""

    def __init__(self, seed, stream=0):
        self.key = hashlib.sha256(b"weighted-bte-amd/%d/%d" % (seed, stream)).digest()
        self.nonce = bytes(12)
        self.counter = 0
        self.buf = b""

    def bytes(self, n):
        while len(self.buf) < n:
            self.buf += chacha20_block(self.key, self.counter, self.nonce)
            self.counter += 1
        out, self.buf = self.buf[:n], self.buf[n:]
        return out

    def scalar(self):
        return int.from_bytes(self.bytes(64), "little") % R

    def u64(self):
        return int.from_bytes(self.bytes(8), "little")

    def below(self, n):
        return self.u64() % n

    def sparse_scalar(self, max_bits=16):
        bits = set()
        k = 1 + self.below(max_bits)
        while len(bits) < k:
            bits.add(self.below(255))
        return sum(1 << b for b in bits) % R


def msm_config_scalars(rng, n, sparse_fraction=0.1):
    """Scalars for the standalone MSM workload: 90% uniform, 10% sparse (<= 16 set bits)."""
    n_sparse = int(n * sparse_fraction)
    out = [rng.scalar() for _ in range(n - n_sparse)] + [rng.sparse_scalar() for _ in range(n_sparse)]
    return out


def fft_cyclotomic(values, inverse=False):
    """weighted_pfe/src/fft.rs:324-365: full unscaled transform over the cyclotomic subgroup,
    twiddles applied by the decomposed action."""
    d = Radix2Domain(len(values))
    g = d.generator_inverse if inverse else d.generator
    return fft_generic(values, g, f12_mul, lambda a, b: f12_mul(a, f12_conj(b)),
                       cyclotomic_mul_scalar_decomposed)


def mul_cyclotomic_signed_small(v, c):
    """weighted_pfe/src/fft.rs:671-697 (an exact power; negative -> conjugate)."""
    r = f12_pow(v, abs(c))
    return f12_conj(r) if c < 0 else r
