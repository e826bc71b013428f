"""Device Fr transforms and the Lagrange-at-zero fast path (weighted_btx/src/interpolation.rs:
101-197: product tree of Fr transforms, P' evaluated on the whole domain) against the oracle and
the definition, at the reference's own test sizes (interpolation.rs:275-338) and at W = 2^14."""
import random

import pytest

from helpers import O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def definition(xs, i):
    num, den = 1, 1
    for j, xj in enumerate(xs):
        if j != i:
            num = num * xj % O.R
            den = den * (xj - xs[i]) % O.R
    return num * pow(den, -1, O.R) % O.R


def domain_points(W, sel):
    dom = O.Radix2Domain(1 << (W - 1).bit_length() if W > 1 else 1)
    return [pow(dom.generator, v, O.R) for v in sel]


@pytest.mark.parametrize("logn,count", [(1, 3), (4, 2), (11, 1), (12, 2), (14, 1)])
def test_fft_fr_matches_oracle_and_roundtrips(api, logn, count):
    """Radix2Domain::fft over Scalar (fft.rs:197-234): equal to the oracle's transform (naive DFT
    at small sizes, which the oracle's radix-2 FFT is pinned to), inverse scaled returns the input."""
    n = 1 << logn
    rnd = random.Random(logn)
    vals = [rnd.randrange(O.R) for _ in range(n * count)]
    raw = b"".join(v.to_bytes(32, "little") for v in vals)
    t = api.to_device(raw)
    y = api.fft_fr(t, count=count)
    got = api.to_host(y)
    for c in range(count):
        want = O.fr_fft(vals[c * n:(c + 1) * n])
        if n <= 16:
            assert want == O.fr_dft_naive(vals[c * n:(c + 1) * n])
        out = [int.from_bytes(got[32 * (c * n + k):32 * (c * n + k + 1)], "little") for k in range(n)]
        assert out == want
    assert api.to_host(api.fft_fr(y, inverse=True, count=count)) == raw
    # unscaled inverse = n * input
    un = api.to_host(api.fft_fr(y, inverse=True, scaled=False, count=count))
    assert int.from_bytes(un[:32], "little") == vals[0] * n % O.R


def test_lagrange_fast_path_sizes_65_to_191(api):
    """The reference's fast-path sizes (interpolation.rs:294-323): every coefficient equals the
    definition, across the direct / product-tree cutoff (64)."""
    rnd = random.Random(9)
    for W, n in ((70, 65), (100, 64), (200, 100), (191, 191), (256, 129)):
        sel = rnd.sample(range(W), n)
        got = api.lagrange_at_zero(W, sel)
        xs = domain_points(W, sel)
        assert got == [definition(xs, i) for i in range(n)], (W, n)


def test_lagrange_complete_and_invalid_sets(api):
    from weighted_bte_amd import BteError

    assert api.lagrange_at_zero(256, list(range(256))) == [pow(256, -1, O.R)] * 256
    # complete set of a non-power-of-two prefix goes through the tree
    sel = list(range(200))
    got = api.lagrange_at_zero(200, sel)
    xs = domain_points(200, sel)
    for i in (0, 99, 199):
        assert got[i] == definition(xs, i)
    for bad in ([1, 1] + list(range(2, 100)), [0, 300] + list(range(1, 99)), []):
        with pytest.raises(BteError):
            api.lagrange_at_zero(200, bad)


def test_lagrange_w_2p14_half_selected(api):
    """W = 2^14 with about W/2 points selected (the serving committees' scale): the
    coefficients reproduce f(0) from f(x_i) for every monomial f = X^k, k < n (sum L_i x_i^k =
    [k == 0]) at several k, and spot coefficients equal the definition."""
    rnd = random.Random(14)
    W = 1 << 14
    sel = sorted(rnd.sample(range(W), W // 2 + 37))
    n = len(sel)
    got = api.lagrange_at_zero(W, sel)
    xs = domain_points(W, sel)
    for k in (0, 1, 2, 3, 1000, n // 2, n - 1):
        acc = sum(l * pow(x, k, O.R) for l, x in zip(got, xs)) % O.R
        assert acc == (1 if k == 0 else 0), "monomial %d" % k
    # degree n: sum L_i x_i^n = (-1)^(n+1) prod x_i (the identity fails for the wrong set)
    prod = 1
    for x in xs:
        prod = prod * x % O.R
    acc = sum(l * pow(x, n, O.R) for l, x in zip(got, xs)) % O.R
    assert acc == ((-1) ** (n + 1) * prod) % O.R
    for i in (0, n // 3, n - 1):
        assert got[i] == definition(xs, i)


def test_device_lagrange_matches_host_interface(api):
    import ctypes

    import torch

    from weighted_bte_amd import call

    rnd = random.Random(15)
    W, n = 3000, 1700
    sel = rnd.sample(range(W), n)
    host = api.lagrange_at_zero(W, sel)
    out = torch.empty(32 * n, dtype=torch.uint8, device="cuda")
    arr = (ctypes.c_uint64 * n)(*sel)
    call("bte_lagrange_at_zero_device", W, arr, n, ctypes.c_void_p(out.data_ptr()),
         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    raw = api.to_host(out)
    assert [int.from_bytes(raw[32 * i:32 * i + 32], "little") for i in range(n)] == host


def test_failed_table_build_leaves_no_partial_entry(api):
    """An allocation failure while a transform size's tables are built is reported, and a later
    call of that size builds them completely (nothing half-made stays cached)."""
    from weighted_bte_amd import BteError, call

    rnd = random.Random(16)
    n = 1 << 13  # a size no other test transforms in the group FFT
    s = [rnd.randrange(O.R) for _ in range(n)]
    pts = api.g1_generator_mul(api.to_device(b"".join(v.to_bytes(32, "little") for v in s)))
    import torch
    one = ((1 << 384) % O.P).to_bytes(48, "little")
    jac = torch.cat([pts.view(n, 96), api.to_device(one).view(1, -1).expand(n, -1)], dim=1).reshape(-1).contiguous()
    d = api.Radix2Domain(n)
    for fail_at in (1, 2, 3, 4):
        call("bte_test_fail_alloc", fail_at)
        with pytest.raises(BteError):
            # tables of the inverse direction of this size are not built yet
            d.ifft_g1(jac)
        call("bte_test_fail_alloc", 0)
    y = d.ifft_g1(jac)
    back = api.batch_normalize_g1(d.fft_g1(y))
    assert torch.equal(back, pts)
    k = 123
    want = O.fr_fft(s, inverse=True)[k]
    got = api.to_host(api.batch_normalize_g1(y[144 * k:144 * (k + 1)]))
    assert O.g1_affine_from_bytes(got) == O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), want))
