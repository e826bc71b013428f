"""GPU parity for the Radix2Domain transforms (weighted_btx/src/fft.rs, weighted_pfe/src/fft.rs):
G1/G2 group FFTs against the closed form DFT_Fr(s) * G, GT transforms against gT^DFT(s), and
cyclotomic transforms against the oracle's restatement bit-exactly (raw values, before any
hard exponentiation, are determined by the base-|x| decomposition both sides use)."""
import random

import pytest

from helpers import O, g1_mul_gen, g2_mul_gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def g1_jac_bytes(aff_list):
    return b"".join(O.g1_jac_to_bytes(O.G1.from_affine(a)) for a in aff_list)


def g2_jac_bytes(aff_list):
    return b"".join(O.g2_jac_to_bytes(O.G2.from_affine(a)) for a in aff_list)


def read_g1(api, t, n):
    aff = api.to_host(api.batch_normalize_g1(t))
    return [O.g1_affine_from_bytes(aff[96 * i: 96 * i + 96]) for i in range(n)]


def read_g2(api, t, n):
    aff = api.to_host(api.batch_normalize_g2(t))
    return [O.g2_affine_from_bytes(aff[192 * i: 192 * i + 192]) for i in range(n)]


def read_f12(api, t, n):
    b = api.to_host(t)
    return [O.f12_from_bytes(b[576 * i: 576 * i + 576]) for i in range(n)]


def f12_bytes(vals):
    return b"".join(O.f12_to_bytes(v) for v in vals)


def rand_cyclotomic(rnd):
    f = tuple(tuple((rnd.randrange(O.P), rnd.randrange(O.P)) for _ in range(3)) for _ in range(2))
    return O.easy_final_exponentiation(f)


@pytest.mark.parametrize("n", [1, 2, 8, 32])
def test_g1_fft_closed_form_and_roundtrip(api, n):
    rnd = random.Random(10 + n)
    s = [rnd.randrange(O.R) for _ in range(n)]
    s[0] = 0  # an identity input
    x = api.to_device(g1_jac_bytes([g1_mul_gen(k) for k in s]))
    d = api.Radix2Domain(n)
    y = d.fft_g1(x)
    want = O.fr_fft(s)
    assert read_g1(api, y, n) == [g1_mul_gen(k) for k in want]
    back = d.ifft_g1(y)
    assert read_g1(api, back, n) == [g1_mul_gen(k) for k in s]
    unscaled = d.ifft_g1(y, scaled=False)
    assert read_g1(api, unscaled, n) == [g1_mul_gen(k * n) for k in s]


def test_g1_fft_matches_literal_oracle_fft(api):
    rnd = random.Random(3)
    n = 8
    pts = [O.G1.mul(O.g1_gen_jac(), rnd.randrange(O.R)) for _ in range(n)]
    x = api.to_device(b"".join(O.g1_jac_to_bytes(p) for p in pts))
    got = read_g1(api, api.Radix2Domain(n).fft_g1(x), n)
    assert got == [O.G1.to_affine(v) for v in O.group_fft(O.G1, pts)]


@pytest.mark.parametrize("n", [2, 8])
def test_g2_fft_closed_form_and_roundtrip(api, n):
    rnd = random.Random(20 + n)
    s = [rnd.randrange(O.R) for _ in range(n)]
    x = api.to_device(g2_jac_bytes([g2_mul_gen(k) for k in s]))
    d = api.Radix2Domain(n)
    y = d.fft_g2(x)
    assert read_g2(api, y, n) == [g2_mul_gen(k) for k in O.fr_fft(s)]
    assert read_g2(api, d.ifft_g2(y), n) == [g2_mul_gen(k) for k in s]


def test_fft_invalid_sizes(api):
    with pytest.raises(ValueError):
        api.Radix2Domain(3)
    from weighted_bte_amd import BteError, call
    import ctypes
    buf = api.to_device(bytes(144 * 3))
    with pytest.raises(BteError):
        call("bte_fft_g1", ctypes.c_void_p(buf.data_ptr()), 3, 0, 0, ctypes.c_void_p(0))


def test_g1_fft_2_8_closed_form(api):
    """Config-3 G1 size 2^8: inputs s_i * G, outputs DFT_Fr(s) * G."""
    n = 256
    rng = O.SeededRng(3, 8)
    s = [rng.scalar() for _ in range(n)]
    x = api.g1_generator_mul(api.to_device(b"".join(O.scalar_to_le32(k) for k in s)))
    # affine -> Jacobian with z = 1 on the host side of the test
    xa = api.to_host(x)
    pts = [O.g1_affine_from_bytes(xa[96 * i: 96 * i + 96]) for i in range(n)]
    d = api.Radix2Domain(n)
    y = d.fft_g1(api.to_device(g1_jac_bytes(pts)))
    got = read_g1(api, y, n)
    want = O.fr_fft(s)
    for k in (0, 1, 2, 77, 128, 255):
        assert got[k] == g1_mul_gen(want[k])
    assert read_g1(api, d.ifft_g1(y), n) == pts


def test_gt_fft_closed_form(api):
    rnd = random.Random(30)
    n = 4
    g = O.gt_generator()
    s = [rnd.randrange(O.R) for _ in range(n)]
    vals = [O.f12_pow(g, k) for k in s]
    d = api.Radix2Domain(n)
    y = d.fft_gt(api.to_device(f12_bytes(vals)))
    assert read_f12(api, y, n) == [O.f12_pow(g, k) for k in O.fr_fft(s)]
    assert read_f12(api, d.ifft_gt(y), n) == vals
    unscaled = read_f12(api, d.ifft_gt(y, scaled=False), n)
    assert unscaled == [O.f12_pow(g, k * n % O.R) for k in s]


def test_cyclotomic_fft_raw_bit_exact(api):
    rnd = random.Random(31)
    n = 4
    vals = [rand_cyclotomic(rnd) for _ in range(n)]
    d = api.Radix2Domain(n)
    x = api.to_device(f12_bytes(vals))
    assert read_f12(api, d.fft_cyclotomic(x), n) == O.fft_cyclotomic(vals)
    assert read_f12(api, d.ifft_cyclotomic_unscaled(x), n) == O.fft_cyclotomic(vals, inverse=True)


@pytest.mark.parametrize("n,count", [(4, 2), (8, 3), (8, 4), (8, 0)])
def test_cyclotomic_prefix_ifft(api, n, count):
    rnd = random.Random(40 + n + count)
    vals = [rand_cyclotomic(rnd) for _ in range(n)]
    got = read_f12(api, api.Radix2Domain(n).ifft_cyclotomic_prefix_unscaled(
        api.to_device(f12_bytes(vals)), count), count)
    want = O.ifft_cyclotomic_prefix_unscaled(vals, count)
    assert got == want[:count]
    # after the hard part the prefix equals the unscaled plain inverse transform (fft.rs:898-957)
    for a, b in zip(got, O.fft_cyclotomic(vals, inverse=True)[:count]):
        assert O.hard_final_exponentiation(a) == O.hard_final_exponentiation(b)


def test_cyclotomic_scalar_actions(api):
    rnd = random.Random(50)
    g = O.gt_generator()
    scal = [0, 1, O.R - 1, O.X_ABS, rnd.randrange(O.R), rnd.randrange(O.R)]
    vals = [rand_cyclotomic(rnd) for _ in scal[:3]] + [g, g, O.f12_pow(g, 5)]
    got = read_f12(api, api.cyclotomic_mul_scalar_decomposed(api.to_device(f12_bytes(vals)), scal),
                   len(vals))
    for v, s, r in zip(vals, scal, got):
        assert r == O.cyclotomic_mul_scalar_decomposed(v, s)
    for i in (3, 4, 5):
        assert got[i] == O.f12_pow(vals[i], scal[i])
    coeffs = [0, 1, -1, 7, -1023, 511]
    got = read_f12(api, api.mul_cyclotomic_signed_small(api.to_device(f12_bytes(vals)), coeffs),
                   len(vals))
    assert got == [O.mul_cyclotomic_signed_small(v, c) for v, c in zip(vals, coeffs)]


def test_g1_fft_segmented_uneven_sizes(api):
    """Uneven transforms in one launch per stage level: each equals its own DFT closed form."""
    rnd = random.Random(60)
    logs = [1, 3, 2, 4, 0, 3]
    s = [[rnd.randrange(O.R) for _ in range(1 << l)] for l in logs]
    flat = [k for seg in s for k in seg]
    x = api.to_device(g1_jac_bytes([g1_mul_gen(k) for k in flat]))
    y = api.fft_g1_segmented(x, logs)
    got = read_g1(api, y, len(flat))
    want = [k for seg in s for k in O.fr_fft(seg)]
    assert got == [g1_mul_gen(k) for k in want]
    back = read_g1(api, api.fft_g1_segmented(y, logs, inverse=True), len(flat))
    assert back == [g1_mul_gen(k) for k in flat]


def test_batched_64_ffts_log_uniform_sizes(api):
    """64 independent G1 FFTs of log-uniform sizes 2^6..2^12 in one launch per stage: round trip
    and closed-form spot checks of forward outputs."""
    import math
    rng = O.SeededRng(3, 64)
    logs = [6 + rng.below(7) for _ in range(64)]
    total = sum(1 << l for l in logs)
    s = [rng.scalar() for _ in range(total)]
    pts = api.g1_generator_mul(api.to_device(b"".join(O.scalar_to_le32(k) for k in s)))
    # affine -> Jacobian (z = 1) on the device side of the test via a round trip through bytes
    aff = api.to_host(pts)
    one = O.fp_to_mont_bytes(1)
    jac = b"".join(aff[96 * i:96 * i + 96] + one for i in range(total))
    x = api.to_device(jac)
    y = api.fft_g1_segmented(x, logs)
    back = api.fft_g1_segmented(y, logs, inverse=True)
    assert api.to_host(api.batch_normalize_g1(back)) == aff
    ya = api.to_host(api.batch_normalize_g1(y))
    off = 0
    for f in (0, 17, 63):
        n = 1 << logs[f]
        seg = s[off_f(logs, f): off_f(logs, f) + n]
        d = O.Radix2Domain(n)
        for k in (0, 1, n - 1):
            v = sum(seg[j] * pow(d.generator, j * k, O.R) for j in range(n)) % O.R
            i = off_f(logs, f) + k
            assert O.g1_affine_from_bytes(ya[96 * i:96 * i + 96]) == g1_mul_gen(v)


def off_f(logs, f):
    return sum(1 << l for l in logs[:f])


@pytest.mark.parametrize("group,logn", [(1, 16), (2, 12)])
def test_fft_config_sizes_roundtrip(api, group, logn):
    """Config-3 largest sizes: G1 2^16 and G2 2^12 forward then inverse is the identity, and
    forward outputs match the closed form at spot indices."""
    n = 1 << logn
    rng = O.SeededRng(4, logn + 100 * group)
    s = [rng.scalar() for _ in range(n)]
    gen = api.g1_generator_mul if group == 1 else api.g2_generator_mul
    width = 96 if group == 1 else 192
    aff = api.to_host(gen(api.to_device(b"".join(O.scalar_to_le32(k) for k in s))))
    one = O.fp_to_mont_bytes(1) if group == 1 else O.f2_to_bytes((1, 0))
    x = api.to_device(b"".join(aff[width * i:width * (i + 1)] + one for i in range(n)))
    d = api.Radix2Domain(n)
    fwd = d.fft_g1 if group == 1 else d.fft_g2
    inv = d.ifft_g1 if group == 1 else d.ifft_g2
    norm = api.batch_normalize_g1 if group == 1 else api.batch_normalize_g2
    y = fwd(x)
    assert api.to_host(norm(inv(y))) == aff
    ya = api.to_host(norm(y))
    od = O.Radix2Domain(n)
    mul = g1_mul_gen if group == 1 else g2_mul_gen
    frm = O.g1_affine_from_bytes if group == 1 else O.g2_affine_from_bytes
    for k in (0, 1, n // 2 + 3, n - 1):
        v = sum(s[j] * pow(od.generator, j * k, O.R) for j in range(n)) % O.R
        assert frm(ya[width * k:width * (k + 1)]) == mul(v)


def test_cyclotomic_prefix_ifft_segmented(api):
    """Uneven prefix transforms in one launch per stage level equal the per-transform oracle."""
    rnd = random.Random(70)
    logs, counts = [1, 3, 2, 3, 4], [1, 2, 0, 4, 5]
    segs = [[rand_cyclotomic(rnd) for _ in range(1 << l)] for l in logs]
    x = api.to_device(f12_bytes([v for seg in segs for v in seg]))
    got = read_f12(api, api.ifft_cyclotomic_prefix_unscaled_segmented(x, logs, counts), sum(counts))
    want = [v for seg, c in zip(segs, counts) for v in O.ifft_cyclotomic_prefix_unscaled(seg, c)[:c]]
    assert got == want


def test_cyclotomic_prefix_one_lane_equals_three_lane(api):
    """64 prefix transforms of size 1024 in one call (32768 butterflies per stage: one thread per
    butterfly) equal the same transforms one call each (512 butterflies: three lanes per element),
    bit for bit."""
    import ctypes

    import torch

    from weighted_bte_amd import call, workloads

    n, count, outs = 1024, 64, 512
    rnd = random.Random(90)
    gt = workloads.gt_generator_bytes()
    x = api.cyclotomic_mul_scalar_decomposed(api.to_device(gt * (n * count)),
                                             [rnd.randrange(O.R) for _ in range(n * count)])
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    a = x.clone()
    call("bte_ifft_cyclotomic_prefix_unscaled_batched", ctypes.c_void_p(a.data_ptr()), n, count, outs, stream)
    b = x.clone()
    for f in range(count):
        call("bte_ifft_cyclotomic_prefix_unscaled", ctypes.c_void_p(b.data_ptr() + f * n * 576), n, outs, stream)
    a = a.view(count, n * 576)[:, :outs * 576]
    b = b.view(count, n * 576)[:, :outs * 576]
    assert torch.equal(a, b)


def _jac_from_affine_device(api, aff, group):
    import torch

    w = 96 if group == 1 else 192
    n = aff.numel() // w
    one = ((1 << 384) % O.P).to_bytes(48, "little") + (bytes(48) if group == 2 else b"")
    rows = aff.view(n, w)
    z = api.to_device(one).view(1, -1).repeat(n, 1)
    z[(rows == 0).all(dim=1)] = 0  # the all-zero affine encoding is the identity: Z = 0
    return torch.cat([rows, z], dim=1).reshape(-1).contiguous()


@pytest.mark.parametrize("group", [1, 2])
def test_small_transforms_by_msm_equal_staged_and_closed_form(api, group):
    """Transforms up to 2^10 (n^2 x count <= 2^20) run as one batched MSM (X_k = sum_j w^(jk)
    x_j): the same points as the staged radix-2 transform for forward, inverse and scaled inverse,
    single and batched up to the 2^10 cutoff, with an identity input; and the closed form
    DFT(s) * G at 2^8."""
    import ctypes

    import torch

    from weighted_bte_amd import call

    rnd = random.Random(80 + group)
    gen = api.g1_generator_mul if group == 1 else api.g2_generator_mul
    norm = api.batch_normalize_g1 if group == 1 else api.batch_normalize_g2
    fn = "bte_fft_g1_batched" if group == 1 else "bte_fft_g2_batched"
    width = 144 if group == 1 else 288
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for logn, count in ((2, 1), (6, 3), (8, 1), (9, 2), (10, 1)):
        n = 1 << logn
        s = [rnd.randrange(O.R) for _ in range(n * count)]
        s[1] = 0  # an identity input
        x = _jac_from_affine_device(api, gen(api.to_device(b"".join(v.to_bytes(32, "little") for v in s))), group)
        for inverse, scaled in ((0, 0), (1, 1), (1, 0)):
            outs = []
            for max_log in (10, 0):
                call("bte_fft_set_dft_max_log", max_log)
                y = x.clone()
                call(fn, ctypes.c_void_p(y.data_ptr()), n, count, inverse, scaled, st)
                outs.append(norm(y))
            call("bte_fft_set_dft_max_log", 10)
            assert torch.equal(outs[0], outs[1]), (logn, count, inverse, scaled)
            assert outs[0].numel() == n * count * width // 3 * 2
        if logn == 8:
            y = x.clone()
            call(fn, ctypes.c_void_p(y.data_ptr()), n, 1, 0, 0, st)
            got = api.to_host(norm(y))
            want = O.fr_fft(s)
            aw = 96 if group == 1 else 192
            for k in (0, 1, 77, 255):
                pt = (O.g1_affine_from_bytes if group == 1 else O.g2_affine_from_bytes)(got[aw * k:aw * (k + 1)])
                mul = g1_mul_gen if group == 1 else g2_mul_gen
                assert pt == (mul(want[k]) if want[k] else None), "closed form k=%d" % k


def test_dft_cutoff_tuner_keeps_result(api, tmp_path):
    """The DFT-cutoff tuner times a G1 transform of 2^9 points in the MSM form and staged,
    installs the faster cutoff (9 or 8), and the transform's points do not change; the cutoff
    round-trips through the autotuner table."""
    import torch

    from weighted_bte_amd import autotune

    rnd = random.Random(91)
    n = 1 << 9
    s = [rnd.randrange(O.R) for _ in range(n)]
    x = _jac_from_affine_device(api, api.g1_generator_mul(api.to_device(b"".join(v.to_bytes(32, "little") for v in s))), 1)
    d = api.Radix2Domain(n)
    before = api.batch_normalize_g1(d.fft_g1(x))
    try:
        cut, times = autotune.tune_dft_cutoff(1, x, reps=1)
        assert cut in (8, 9) and set(times) == {"msm_ms", "staged_ms"}
        assert torch.equal(api.batch_normalize_g1(d.fft_g1(x)), before)
        path = str(tmp_path / "t.json")
        autotune.save_table({}, path, knobs={"dft_max_log": 8})
        assert autotune.apply_table(path) == 1
        assert torch.equal(api.batch_normalize_g1(d.fft_g1(x)), before)
    finally:
        autotune.set_dft_max_log(autotune.DFT_MAX_LOG_DEFAULT)
