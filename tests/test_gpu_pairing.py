"""GPU parity for Miller loops and the split final exponentiation
(weighted_btx/src/final_exponentiation.rs) against the oracle.  Raw Miller values depend on
line scaling, so comparisons are made after the final exponentiation (and, for the easy and
hard parts alone, on their exactly defined exponents)."""
import random

import pytest

from helpers import O, g1_mul_gen, g1_points_bytes, g2_mul_gen, g2_points_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def read_f12(api, t, n):
    b = api.to_host(t)
    return [O.f12_from_bytes(b[576 * i: 576 * i + 576]) for i in range(n)]


def f12_bytes(vals):
    return b"".join(O.f12_to_bytes(v) for v in vals)


def rand_f12(rnd):
    return tuple(tuple((rnd.randrange(O.P), rnd.randrange(O.P)) for _ in range(3)) for _ in range(2))


def test_easy_hard_full_exact(api):
    rnd = random.Random(1)
    fs = [rand_f12(rnd) for _ in range(7)]
    x = api.to_device(f12_bytes(fs))
    easy = api.easy_final_exponentiation(x)
    assert read_f12(api, easy, 7) == [O.easy_final_exponentiation(f) for f in fs]
    hard = read_f12(api, api.hard_final_exponentiation(easy), 7)
    assert hard == [O.final_exponentiation(f) for f in fs]
    assert read_f12(api, api.full_final_exponentiation(x), 7) == hard


def test_prepared_miller_loop_pairing(api):
    rnd = random.Random(2)
    a = [rnd.randrange(1, O.R) for _ in range(5)]
    b = [rnd.randrange(1, O.R) for _ in range(5)]
    P = [g1_mul_gen(k) for k in a] + [None, g1_mul_gen(3)]
    Q = [g2_mul_gen(k) for k in b] + [g2_mul_gen(9), None]
    lines = api.prepare_g2_lines(api.to_device(g2_points_bytes(Q)))
    ml = api.miller_loop_prepared(lines, api.to_device(g1_points_bytes(P)))
    got = read_f12(api, api.full_final_exponentiation(ml), len(P))
    g = O.gt_generator()
    for i in range(5):
        assert got[i] == O.f12_pow(g, a[i] * b[i] % O.R)
    assert got[5] == O.F12_ONE and got[6] == O.F12_ONE


def test_unprepared_pairs_match_prepared(api):
    rnd = random.Random(3)
    P = [g1_mul_gen(rnd.randrange(1, O.R)) for _ in range(4)]
    Q = [g2_mul_gen(rnd.randrange(1, O.R)) for _ in range(4)]
    pd, qd = api.to_device(g1_points_bytes(P)), api.to_device(g2_points_bytes(Q))
    a = read_f12(api, api.full_final_exponentiation(api.miller_loop_pairs(pd, qd)), 4)
    b = read_f12(api, api.full_final_exponentiation(
        api.miller_loop_prepared(api.prepare_g2_lines(qd), pd)), 4)
    assert a == b
    assert a == [O.pairing(p, q) for p, q in zip(P, Q)]


def test_multi_miller_uneven_segments(api):
    """unprepared_multi_miller_loop over segments of arity 3, 0, 1, 5 (identities dropped)."""
    rnd = random.Random(4)
    sizes = [3, 0, 1, 5]
    P, Q, segs = [], [], []
    for sz in sizes:
        seg = []
        for _ in range(sz):
            a, b = rnd.randrange(1, O.R), rnd.randrange(1, O.R)
            seg.append((a, b))
        segs.append(seg)
    flat = [ab for seg in segs for ab in seg]
    P = [g1_mul_gen(a) for a, _ in flat]
    Q = [g2_mul_gen(b) for _, b in flat]
    P[1] = None  # identity inside segment 0
    Q[6] = None  # identity inside segment 3
    offs = [0]
    for sz in sizes:
        offs.append(offs[-1] + sz)
    out = api.unprepared_multi_miller_loop(api.to_device(g1_points_bytes(P)),
                                           api.to_device(g2_points_bytes(Q)), offs)
    got = read_f12(api, api.full_final_exponentiation(out), len(sizes))
    for s in range(len(sizes)):
        pairs = [(P[i], Q[i]) for i in range(offs[s], offs[s + 1])]
        assert got[s] == O.multi_pairing(pairs)


def test_multi_pairing_fused_one_launch(api):
    """The one-launch multi-pairing (Miller chunks, per-segment product and final exponentiation in
    one kernel) equals the staged path bit for bit and the oracle's multi-pairing, over segments of
    arity 3, 0, 1, 5, 2 and 12 with identity pairs (final_exponentiation.rs:229-273, :348-354),
    followed by 30 segments of arity 1 or 2, so that one wavefront closes more segments than one
    round of three-lane hard parts takes; with the chunk knob at 1 and at 16 pairs per
    accumulator."""
    import torch

    rnd = random.Random(41)
    sizes = [3, 0, 1, 5, 2, 12] + [1, 2, 1] * 10
    n = sum(sizes)
    ab = [(rnd.randrange(1, O.R), rnd.randrange(1, O.R)) for _ in range(n)]
    P = [g1_mul_gen(a) for a, _ in ab]
    Q = [g2_mul_gen(b) for _, b in ab]
    P[1] = None  # identity inside segment 0
    Q[7] = None  # identity inside segment 3
    offs = [0]
    for sz in sizes:
        offs.append(offs[-1] + sz)
    pd, qd = api.to_device(g1_points_bytes(P)), api.to_device(g2_points_bytes(Q))
    lines = api.prepare_g2_lines(qd)
    staged = api.full_final_exponentiation(api.multi_miller_loop_prepared(lines, pd, offs))
    try:
        for chunk in (1, 16):
            api.tune_set(api.TUNE_MM_CHUNK_MAX, chunk)
            api.tune_set(api.TUNE_MM_MIN_THREADS, 1)
            fused = api.multi_pairing_prepared_fused(lines, pd, offs)
            assert torch.equal(fused, staged), chunk
    finally:
        api.tune_set(api.TUNE_MM_CHUNK_MAX, api.TUNE_DEFAULTS[api.TUNE_MM_CHUNK_MAX])
        api.tune_set(api.TUNE_MM_MIN_THREADS, api.TUNE_DEFAULTS[api.TUNE_MM_MIN_THREADS])
    got = read_f12(api, fused, len(sizes))
    for s_ in range(6):
        pairs = [(P[i], Q[i]) for i in range(offs[s_], offs[s_ + 1])]
        assert got[s_] == O.multi_pairing(pairs)


def test_bilinearity_on_gpu(api):
    rnd = random.Random(5)
    a, b = rnd.randrange(1, O.R), rnd.randrange(1, O.R)
    P = [g1_mul_gen(a), O.G1_GEN]
    Q = [g2_mul_gen(b), g2_mul_gen(a * b)]
    e = read_f12(api, api.full_final_exponentiation(
        api.miller_loop_pairs(api.to_device(g1_points_bytes(P)), api.to_device(g2_points_bytes(Q)))), 2)
    assert e[0] == e[1]
    assert O.f12_pow(e[0], O.R) == O.F12_ONE
