"""Pins the CPU oracle (oracle/bls12_381.py) against published constants, the reference's own
constants and test shapes, closed forms and invariants.  CPU only."""
import random

import pytest

from helpers import O

# HARD_EXPONENT limbs as written in weighted_btx/src/fft.rs:660-681 (little-endian u64).
REF_HARD_EXPONENT = [
    0xE516C3F438E3BA79, 0xFA9912AAE208CCF1, 0x905CE937335D5B68, 0xC71A2629B0DEA236,
    0x83774940996754C8, 0x21D160AEB6A1E799, 0x2ED0B283ED237DB4, 0x915C97F36C6F1821,
    0x67F17FCBDE783765, 0x2378B9039096D1B7, 0x7988F8761BDC51DC, 0x2076995003FC77A1,
    0x827ECA0BA621315B, 0xE5A72BCE8D63CB9F, 0xF68F7764C28B6F8A, 0x2F230063CF081517,
    0x94506632528D6A9A, 0xD3CDE88EEB996CA3, 0xC0BD38C3195C899E, 0x000F686B3D807D01,
]

# ZCash-format compressed generators (published test vectors).
G1_GEN_COMPRESSED = (
    "97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb")
G2_GEN_COMPRESSED = (
    "93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e"
    "024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8")


def rand_f12(rnd):
    return tuple(tuple((rnd.randrange(O.P), rnd.randrange(O.P)) for _ in range(3)) for _ in range(2))


def test_parameters():
    x = O.X
    assert O.R == x**4 - x**2 + 1
    assert O.P == (x - 1) ** 2 * O.R // 3 + x
    assert (x - 1) ** 2 * O.R % 3 == 0
    assert O.HARD_EXPONENT == sum(l << (64 * i) for i, l in enumerate(REF_HARD_EXPONENT))
    assert (O.P**4 - O.P**2 + 1) % O.R == 0
    # Fr two-adicity and root of unity (ff PrimeField::S / ROOT_OF_UNITY)
    assert (O.R - 1) % (1 << 32) == 0 and ((O.R - 1) >> 32) & 1
    w = O.FR_ROOT_OF_UNITY
    assert pow(7, (O.R - 1) >> 32, O.R) == w
    assert pow(w, 1 << 31, O.R) == O.R - 1


def test_generators_kat():
    assert O.G1.on_curve_affine(O.G1_GEN)
    assert O.G2.on_curve_affine(O.G2_GEN)
    assert O.g1_compress(O.G1_GEN).hex() == G1_GEN_COMPRESSED
    assert O.g2_compress(O.G2_GEN).hex() == G2_GEN_COMPRESSED
    assert O.G1.is_identity(O.G1.mul(O.g1_gen_jac(), O.R))
    assert O.G2.is_identity(O.G2.mul(O.g2_gen_jac(), O.R))
    assert O.g1_decompress(bytes.fromhex(G1_GEN_COMPRESSED)) == O.G1_GEN


def test_group_law_against_affine_naive():
    rnd = random.Random(1)
    for C, gen in ((O.G1, O.g1_gen_jac()), (O.G2, O.g2_gen_jac())):
        a, b = rnd.randrange(1, 1000), rnd.randrange(1, 1000)
        pa, pb = C.mul(gen, a), C.mul(gen, b)
        assert C.eq(C.add(pa, pb), C.mul(gen, a + b))
        assert C.eq(C.double(pa), C.mul(gen, 2 * a))
        assert C.eq(C.sub(pa, pa), C.identity())
        # repeated addition brute force
        acc = C.identity()
        for _ in range(a % 37 + 1):
            acc = C.add(acc, gen)
        assert C.eq(acc, C.mul(gen, a % 37 + 1))
        assert C.on_curve_affine(C.to_affine(pa))


def test_frobenius_is_p_power():
    rnd = random.Random(2)
    a = rand_f12(rnd)
    for k in (1, 2, 3):
        assert O.f12_frobenius(a, k) == O.f12_pow(a, O.P**k)


def test_fp12_inverse_and_tower():
    rnd = random.Random(3)
    a = rand_f12(rnd)
    assert O.f12_mul(a, O.f12_inv(a)) == O.F12_ONE
    # w^6 == xi
    w = ((O.F2_ZERO,) * 3, (O.F2_ONE, O.F2_ZERO, O.F2_ZERO))
    w6 = O.f12_pow(w, 6)
    assert w6 == (((1, 1), O.F2_ZERO, O.F2_ZERO), (O.F2_ZERO,) * 3)


def test_hard_chain_is_three_lambda():
    """final_exponentiation.rs:308-336 computes f^(3*Lambda) on the cyclotomic subgroup."""
    rnd = random.Random(4)
    f = O.easy_final_exponentiation(rand_f12(rnd))
    assert O.hard_final_exponentiation(f) == O.f12_pow(f, 3 * O.HARD_EXPONENT)
    # and the full exponent is 3 (p^12 - 1) / r
    g = rand_f12(rnd)
    assert O.final_exponentiation(g) == O.f12_pow(g, 3 * (O.P**12 - 1) // O.R)


def test_easy_part_lands_in_cyclotomic_subgroup():
    rnd = random.Random(5)
    f = O.easy_final_exponentiation(rand_f12(rnd))
    phi6 = O.P**4 - O.P**2 + 1
    assert O.f12_pow(f, phi6) == O.F12_ONE
    assert O.f12_mul(f, O.f12_conj(f)) == O.F12_ONE  # inverse = conjugate


def test_pairing_bilinear_nondegenerate():
    rnd = random.Random(6)
    g = O.gt_generator()
    assert g != O.F12_ONE
    assert O.f12_pow(g, O.R) == O.F12_ONE
    a, b = rnd.randrange(O.R), rnd.randrange(O.R)
    pa = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), a))
    qb = O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), b))
    assert O.pairing(pa, qb) == O.f12_pow(g, a * b % O.R)
    # multi-pairing = product, identities are dropped, empty = one
    c = rnd.randrange(O.R)
    pc = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), c))
    assert O.multi_pairing([(pa, qb), (pc, O.G2_GEN), (None, qb), (pa, None)]) == \
        O.f12_pow(g, (a * b + c) % O.R)
    assert O.miller_loop([]) == O.F12_ONE


def test_decomposed_gt_action_matches_pow():
    """fft.rs:811-846: decomposed action equals Gt * scalar on r-torsion inputs."""
    rnd = random.Random(7)
    g = O.gt_generator()
    for s in [0, 1, O.R - 1, O.X_ABS, O.X_ABS**2 + 5, rnd.randrange(O.R), rnd.randrange(O.R)]:
        assert O.cyclotomic_mul_scalar_decomposed(g, s) == O.f12_pow(g, s % O.R)
    comps = O.decompose_scalar_base_x(O.R - 1)
    assert sum(c * O.X_ABS**i for i, c in enumerate(comps)) == O.R - 1
    d, n = O.width_w_naf_u64(2**64 - 1)
    assert n <= 65 and sum(v << i for i, v in enumerate(d)) == 2**64 - 1


def test_fr_fft_matches_naive_dft_and_roundtrip():
    rnd = random.Random(8)
    for n in (1, 2, 4, 8, 16, 32):
        v = [rnd.randrange(O.R) for _ in range(n)]
        assert O.fr_fft(v) == O.fr_dft_naive(v)
        assert O.fr_fft(O.fr_fft(v), inverse=True) == v
    with pytest.raises(ValueError):
        O.Radix2Domain(3)
    with pytest.raises(ValueError):
        O.Radix2Domain(0)


def test_group_fft_closed_form():
    """FFT(s_i * G) == DFT_Fr(s) * G (G1 and G2), ifft inverts."""
    rnd = random.Random(9)
    n = 8
    s = [rnd.randrange(O.R) for _ in range(n)]
    for C, gen in ((O.G1, O.g1_gen_jac()), (O.G2, O.g2_gen_jac())):
        vals = [C.mul(gen, k) for k in s]
        out = O.group_fft(C, vals)
        want = O.fr_fft(s)
        for o, k in zip(out, want):
            assert C.eq(o, C.mul(gen, k))
        back = O.group_fft(C, out, inverse=True)
        for b, k in zip(back, s):
            assert C.eq(b, C.mul(gen, k))


def test_gt_fft_and_cyclotomic_prefix():
    rnd = random.Random(10)
    g = O.gt_generator()
    n = 8
    s = [rnd.randrange(O.R) for _ in range(n)]
    vals = [O.f12_pow(g, k) for k in s]
    out = O.gt_fft(vals)
    for o, k in zip(out, O.fr_fft(s)):
        assert o == O.f12_pow(g, k)
    # prefix IFFT (unscaled) keeps the first outputs of n * IFFT
    pref = O.ifft_cyclotomic_prefix_unscaled(vals, n // 2)
    full = O.fr_fft(s, inverse=True, scaled=False)
    for o, k in zip(pref, full[: n // 2]):
        assert O.hard_final_exponentiation(o) == O.hard_final_exponentiation(O.f12_pow(g, k))


def test_msm_reference_test_shape():
    """blst_utils.rs:203-249 shape: bases (3i+1)G with identities at i = 0 mod 17, scalars 5i+2."""
    n = 40
    ks = [0 if i % 17 == 0 else 3 * i + 1 for i in range(n)]
    pts = [None if k == 0 else O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), k)) for k in ks]
    sc = [5 * i + 2 for i in range(n)]
    got = O.msm(O.G1, pts, sc)
    want = sum(k * s for k, s in zip(ks, sc)) % O.R
    assert O.G1.eq(got, O.G1.mul(O.g1_gen_jac(), want))
    assert O.G1.is_identity(O.msm(O.G1, [], []))


def test_msm_ignores_bit_255():
    s = (1 << 255) | 12345
    got = O.msm(O.G1, [O.G1_GEN], [s])
    assert O.G1.eq(got, O.G1.mul(O.g1_gen_jac(), 12345))


def test_layout_roundtrip():
    rnd = random.Random(11)
    k = rnd.randrange(O.R)
    p = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), k))
    assert O.g1_affine_from_bytes(O.g1_affine_to_bytes(p)) == p
    q = O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), k))
    assert O.g2_affine_from_bytes(O.g2_affine_to_bytes(q)) == q
    f = rand_f12(rnd)
    assert O.f12_from_bytes(O.f12_to_bytes(f)) == f
    assert O.g1_affine_to_bytes(None) == bytes(96)


def test_chacha20_rfc8439_block():
    """RFC 8439 section 2.3.2 test vector for the seeded input generator."""
    key = bytes(range(32))
    nonce = bytes.fromhex("000000090000004a00000000")
    blk = O.chacha20_block(key, 1, nonce)
    assert blk[:16].hex() == "10f1e7e4d13b5915500fdd1fa32071c4"
