"""Shared test helpers: seeded inputs from the oracle, blst-layout packing, device round trips."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import bls12_381 as O  # noqa: E402


def pack_scalars(scalars):
    return b"".join(O.scalar_to_le32(s) for s in scalars)


def g1_points_bytes(points_aff):
    return b"".join(O.g1_affine_to_bytes(p) for p in points_aff)


def g2_points_bytes(points_aff):
    return b"".join(O.g2_affine_to_bytes(p) for p in points_aff)


def g1_jac_affine(b):
    return O.G1.to_affine(O.g1_jac_from_bytes(b))


def g2_jac_affine(b):
    return O.G2.to_affine(O.g2_jac_from_bytes(b))


def g1_mul_gen(k):
    return O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), k % O.R))


def g2_mul_gen(k):
    return O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), k % O.R))
