"""GPU compressed encodings of G1/G2 against the oracle's ZCash-format compressors (pinned by
the generator KATs in test_oracle.py), round trips, and rejection of malformed encodings: no
compression flag, bad infinity encodings, x >= p, x off the curve and points outside the
order-r subgroup."""
import random

import pytest

from helpers import O, g1_mul_gen, g2_mul_gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def test_g1_compress_roundtrip_and_rejects(api):
    rnd = random.Random(1)
    pts = [g1_mul_gen(rnd.randrange(1, O.R)) for _ in range(20)] + [None]
    dev = api.to_device(b"".join(O.g1_affine_to_bytes(p) for p in pts))
    enc = api.to_host(api.g1_compress(dev))
    assert enc == b"".join(O.g1_compress(p) for p in pts)
    back, ok = api.g1_decompress(api.to_device(enc))
    assert api.to_host(ok) == bytes([1]) * len(pts)
    assert api.to_host(back) == api.to_host(dev)
    bad = []
    e0 = bytearray(enc[:48])
    e0[0] &= 0x7f
    bad.append(bytes(e0))                                   # no compression flag
    bad.append(bytes([0xc0]) + bytes(46) + b"\x01")         # infinity with a nonzero byte
    bad.append(bytes([0xe0]) + bytes(47))                   # infinity with the sign flag
    bad.append(bytes([0x80 | 0x1a]) + bytes([0xff]) * 47)   # x >= p
    # x off the curve, and on the curve but outside the subgroup (cofactor h != 1)
    off, outside = None, None
    while off is None or outside is None:
        x = rnd.randrange(O.P)
        rhs = (x * x * x + 4) % O.P
        if O.fp_sqrt(rhs) is None:
            off = off or x
        elif not O.G1.is_identity(O.G1.mul(O.G1.from_affine((x, O.fp_sqrt(rhs))), O.R)):
            outside = outside or x
    for x in (off, outside):
        b = bytearray(x.to_bytes(48, "big"))
        b[0] |= 0x80
        bad.append(bytes(b))
    _, okb = api.g1_decompress(api.to_device(b"".join(bad)))
    assert api.to_host(okb) == bytes(len(bad))


def test_g2_compress_roundtrip_and_rejects(api):
    rnd = random.Random(2)
    pts = [g2_mul_gen(rnd.randrange(1, O.R)) for _ in range(8)] + [None]
    dev = api.to_device(b"".join(O.g2_affine_to_bytes(p) for p in pts))
    enc = api.to_host(api.g2_compress(dev))
    assert enc == b"".join(O.g2_compress(p) for p in pts)
    back, ok = api.g2_decompress(api.to_device(enc))
    assert api.to_host(ok) == bytes([1]) * len(pts)
    assert api.to_host(back) == api.to_host(dev)
    bad = [bytes([0xc0]) + bytes(94) + b"\x02", bytes([0x80 | 0x1b]) + bytes([0xff]) * 95]
    outside = None
    while outside is None:
        x = (rnd.randrange(O.P), rnd.randrange(O.P))
        rhs = O.f2_add(O.f2_mul(O.f2_sqr(x), x), (4, 4))
        y = O.f2_sqrt(rhs)
        if y is not None and not O.G2.is_identity(O.G2.mul(O.G2.from_affine((x, y)), O.R)):
            outside = x
    b = bytearray(outside[1].to_bytes(48, "big") + outside[0].to_bytes(48, "big"))
    b[0] |= 0x80
    bad.append(bytes(b))
    _, okb = api.g2_decompress(api.to_device(b"".join(bad)))
    assert api.to_host(okb) == bytes(len(bad))
