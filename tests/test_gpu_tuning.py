"""The launch-shape knobs the autotuner sets (bte_tune_set) change how work is split across
threads, never a result: MSMs with other accumulation chunk lengths give the same points, the
cyclotomic prefix transform on one lane or three lanes per element gives the same raw values,
and multi-Miller loops with other accumulator chunks give the same raw Fp12 (the shared
squarings distribute over the product, final_exponentiation.rs:229-273)."""
import ctypes
import random

import pytest

from helpers import O, g1_mul_gen, g1_points_bytes, g2_mul_gen, g2_points_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    yield A
    for knob, default in A.TUNE_DEFAULTS.items():
        A.tune_set(knob, default)


def test_msm_same_points_for_every_accumulation_round_count(api):
    from weighted_bte_amd import inputs

    n = 1 << 14
    st = inputs.Stream(31, 14)
    ks, sc = st.scalars(n), inputs.msm_scalars(st, n)
    bases = api.g1_generator_mul(api.to_device(inputs.pack_scalars(ks)))
    dsc = api.to_device(inputs.pack_scalars(sc))
    want = api.to_host(api.g1_generator_mul(api.to_device((sum(k * s for k, s in zip(ks, sc)) % O.R).to_bytes(32, "little"))))
    for rounds in (1, 2, 4, 16):
        api.tune_set(api.TUNE_MSM_ACC_ROUNDS, rounds)
        got = api.to_host(api.batch_normalize_g1(api.g1_multi_exp_affine_bytes(bases, dsc)))
        assert got == want, rounds
    api.tune_set(api.TUNE_MSM_ACC_ROUNDS, api.TUNE_DEFAULTS[api.TUNE_MSM_ACC_ROUNDS])


def test_cyclotomic_prefix_same_raw_values_on_one_or_three_lanes(api):
    import torch

    from weighted_bte_amd import call, workloads

    n, count, outs = 64, 12, 32
    rnd = random.Random(17)
    gt = workloads.gt_generator_bytes()
    x = api.cyclotomic_mul_scalar_decomposed(api.to_device(gt * (n * count)),
                                             [rnd.randrange(O.R) for _ in range(n * count)])
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    results = []
    for thr in (0, 1 << 30, 100):  # every stage one lane, every stage three lanes, a split
        api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, thr)
        a = x.clone()
        call("bte_ifft_cyclotomic_prefix_unscaled_batched", ctypes.c_void_p(a.data_ptr()), n, count, outs, stream)
        results.append(a.view(count, n * 576)[:, :outs * 576].clone())
    api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, api.TUNE_DEFAULTS[api.TUNE_CYC_THREE_LANE_MAX])
    assert torch.equal(results[0], results[1]) and torch.equal(results[0], results[2])
    # and the values are the oracle's prefix transform of the first element
    first = [O.f12_from_bytes(api.to_host(x[576 * i:576 * (i + 1)])) for i in range(n)]
    want = O.ifft_cyclotomic_prefix_unscaled(first, outs)[:outs]
    got = api.to_host(results[0][0])
    assert [O.f12_from_bytes(got[576 * i:576 * (i + 1)]) for i in range(outs)] == want


def test_multi_miller_same_raw_values_for_every_chunk(api):
    import torch

    rnd = random.Random(23)
    arity = [1, 5, 0, 9, 2, 17]
    off = [0]
    for a in arity:
        off.append(off[-1] + a)
    npairs = off[-1]
    P = api.to_device(b"".join(O.g1_affine_to_bytes(g1_mul_gen(rnd.randrange(1, O.R))) for _ in range(npairs)))
    Q = api.to_device(b"".join(O.g2_affine_to_bytes(g2_mul_gen(rnd.randrange(1, O.R))) for _ in range(npairs)))
    lines = api.prepare_g2_lines(Q)
    outs = []
    for chunk in (1, 3, 8, 64):
        api.tune_set(api.TUNE_MM_CHUNK_MAX, chunk)
        api.tune_set(api.TUNE_MM_MIN_THREADS, 1)  # let the chunk reach its maximum at this size
        outs.append(api.multi_miller_loop_prepared(lines, P, off).clone())
    for knob in (api.TUNE_MM_CHUNK_MAX, api.TUNE_MM_MIN_THREADS):
        api.tune_set(knob, api.TUNE_DEFAULTS[knob])
    assert all(torch.equal(outs[0], o) for o in outs[1:])


def test_final_exponentiation_same_values_on_one_or_three_lanes(api):
    """The hard part of the final exponentiation with one lane per element and with three lanes
    per element (below the three-lane switch) gives the same values: 50 elements (workgroups of
    21, the last one partial), through the hard part alone and through the full final
    exponentiation (in place), equal to the oracle's final exponentiation on a few
    (final_exponentiation.rs:308-354)."""
    import torch

    rnd = random.Random(23)
    n = 50
    P = [g1_mul_gen(rnd.randrange(1, O.R)) for _ in range(n)]
    Q = [g2_mul_gen(rnd.randrange(1, O.R)) for _ in range(n)]
    ml = api.miller_loop_pairs(api.to_device(g1_points_bytes(P)), api.to_device(g2_points_bytes(Q)))
    easy = api.easy_final_exponentiation(ml)
    hard, full = [], []
    try:
        for thr in (0, 1 << 30):  # every size on one lane, every size on three lanes
            api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, thr)
            hard.append(api.hard_final_exponentiation(easy))
            full.append(api.full_final_exponentiation(ml))
    finally:
        api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, api.TUNE_DEFAULTS[api.TUNE_CYC_THREE_LANE_MAX])
    assert torch.equal(hard[0], hard[1]) and torch.equal(full[0], full[1]) and torch.equal(hard[0], full[0])
    raw = api.to_host(full[1])
    for i in (0, 20, 21, 49):
        want = O.final_exponentiation(O.miller_loop([(P[i], Q[i])]))
        assert O.f12_from_bytes(raw[576 * i:576 * (i + 1)]) == want, i


def test_easy_part_same_values_for_every_chunk_length(api):
    """The easy part shares one Fp12 inversion over 1, 2 or 4 elements depending on n (one per
    thread up to 65,536 elements): 131,077 elements (4 per inversion) equal the first 65,536
    (1 per inversion) and the remaining 65,541 (2 per inversion) run alone, and two of them equal
    the oracle's easy part (final_exponentiation.rs:276-305)."""
    import torch

    n1, n = 65536, 131077
    g = torch.Generator(device="cpu").manual_seed(9)
    v = torch.randint(0, 1 << 31, (n * 144,), dtype=torch.int64, generator=g).to(torch.int32).view(n, 12, 12)
    v[:, :, 11] &= 0x0FFFFFFF  # every limb string below p
    x = v.reshape(-1).view(torch.uint8).to("cuda")
    whole = api.easy_final_exponentiation(x)
    first = api.easy_final_exponentiation(x[:576 * n1].clone())
    rest = api.easy_final_exponentiation(x[576 * n1:].clone())
    assert torch.equal(whole, torch.cat([first, rest]))
    raw_in, raw_out = api.to_host(x[:576 * 2]), api.to_host(whole[:576 * 2])
    for i in range(2):
        f = O.f12_from_bytes(raw_in[576 * i:576 * (i + 1)])
        assert O.f12_from_bytes(raw_out[576 * i:576 * (i + 1)]) == O.easy_final_exponentiation(f)
