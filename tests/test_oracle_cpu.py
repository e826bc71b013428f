"""Pins the oracle's C restatement (oracle/csrc/bls_cpu.c) against the Python oracle."""
import random

import pytest

from helpers import O, g1_points_bytes, g2_points_bytes, pack_scalars


@pytest.fixture(scope="module")
def clib():
    from oracle import cpu

    lib = cpu.load()
    if lib is None:
        pytest.skip("C oracle not built")
    return lib


def test_fp_mul_matches(clib):
    import ctypes
    rnd = random.Random(1)
    for _ in range(20):
        a, b = rnd.randrange(O.P), rnd.randrange(O.P)
        out = ctypes.create_string_buffer(48)
        clib.oracle_fp_mul(O.fp_to_mont_bytes(a), O.fp_to_mont_bytes(b), out)
        assert O.fp_from_mont_bytes(out.raw) == a * b % O.P


@pytest.mark.parametrize("n,c", [(1, 0), (37, 0), (100, 4), (300, 7)])
def test_c_g1_msm_matches_python(clib, n, c):
    from oracle import cpu
    rnd = random.Random(n)
    ks = [rnd.randrange(1, O.R) for _ in range(n)]
    pts = [O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), k)) for k in ks]
    pts[0] = None
    sc = [rnd.randrange(O.R) for _ in range(n)]
    got = O.G1.to_affine(O.g1_jac_from_bytes(cpu.g1_msm(clib, g1_points_bytes(pts), pack_scalars(sc), c)))
    assert got == O.G1.to_affine(O.msm(O.G1, pts, sc))


def test_c_g2_msm_matches_python(clib):
    from oracle import cpu
    rnd = random.Random(5)
    n = 20
    pts = [O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), rnd.randrange(1, O.R))) for _ in range(n)]
    sc = [rnd.randrange(O.R) for _ in range(n)]
    got = O.G2.to_affine(O.g2_jac_from_bytes(cpu.g2_msm(clib, g2_points_bytes(pts), pack_scalars(sc))))
    assert got == O.G2.to_affine(O.msm(O.G2, pts, sc))


def test_c_consecutive_multiples_closed_form(clib):
    from oracle import cpu
    n = 5000
    pts = cpu.g1_consecutive_multiples(clib, O.g1_affine_to_bytes(O.G1_GEN), n)
    for i in (0, 1, 2, 777, n - 1):
        assert O.g1_affine_from_bytes(pts[96 * i: 96 * i + 96]) == \
            O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), i + 1))
    rnd = random.Random(9)
    sc = [rnd.randrange(O.R) for _ in range(n)]
    got = O.G1.to_affine(O.g1_jac_from_bytes(cpu.g1_msm(clib, pts, pack_scalars(sc))))
    want = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), sum((i + 1) * s for i, s in enumerate(sc)) % O.R))
    assert got == want


def test_c_pairing_path_matches_python(clib):
    import ctypes
    from oracle import cpu
    rnd = random.Random(21)
    a, b = rnd.randrange(1, O.R), rnd.randrange(1, O.R)
    P = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), a))
    Q = O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), b))
    out = ctypes.create_string_buffer(576)
    clib.oracle_miller_pairs(O.g1_affine_to_bytes(P), O.g2_affine_to_bytes(Q), 1, out)
    fe = ctypes.create_string_buffer(576)
    clib.oracle_final_exp(out.raw, 1, 1, 1, fe)
    assert O.f12_from_bytes(fe.raw) == O.pairing(P, Q)
    # easy alone is exact
    f = tuple(tuple((rnd.randrange(O.P), rnd.randrange(O.P)) for _ in range(3)) for _ in range(2))
    clib.oracle_final_exp(O.f12_to_bytes(f), 1, 1, 0, fe)
    e = O.f12_from_bytes(fe.raw)
    assert e == O.easy_final_exponentiation(f)
    # decomposed power with a plan
    s = rnd.randrange(O.R)
    buf = ctypes.create_string_buffer(O.f12_to_bytes(e), 576)
    clib.oracle_cyc_pow_plans(buf, cpu.plan_bytes(s), 1)
    assert O.f12_from_bytes(buf.raw) == O.cyclotomic_mul_scalar_decomposed(e, s)
