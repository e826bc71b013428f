"""GPU parity for the Pippenger MSM, batch normalisation and fixed-base generator multiples
against the oracle (weighted_btx/src/blst_utils.rs semantics).  Bit-exact on affine points."""
import random

import pytest
import torch

from helpers import O, g1_jac_affine, g1_mul_gen, g1_points_bytes, g2_jac_affine, g2_mul_gen, \
    g2_points_bytes, pack_scalars

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def _rand_bases(rnd, n, mul):
    ks = [rnd.randrange(1, O.R) for _ in range(n)]
    return ks, [mul(k) for k in ks]


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 200])
def test_g1_msm_small_vs_oracle(api, n):
    rnd = random.Random(100 + n)
    ks, pts = _rand_bases(rnd, n, g1_mul_gen)
    sc = [rnd.randrange(O.R) for _ in range(n)]
    out = api.g1_multi_exp_affine_bytes(api.to_device(g1_points_bytes(pts)),
                                        api.to_device(pack_scalars(sc)))
    got = g1_jac_affine(api.to_host(out))
    want = O.G1.to_affine(O.msm(O.G1, pts, sc))
    assert got == want
    assert O.g1_compress(got) == O.g1_compress(g1_mul_gen(sum(k * s for k, s in zip(ks, sc))))


def test_g1_msm_reference_shape_with_identities(api):
    """blst_utils.rs:203-249: bases (3i+1)G, identity when i = 0 mod 17, scalars 5i+2."""
    n = 100
    ks = [0 if i % 17 == 0 else 3 * i + 1 for i in range(n)]
    pts = [None if k == 0 else g1_mul_gen(k) for k in ks]
    sc = [5 * i + 2 for i in range(n)]
    out = api.g1_multi_exp_affine_bytes(api.to_device(g1_points_bytes(pts)),
                                        api.to_device(pack_scalars(sc)))
    assert g1_jac_affine(api.to_host(out)) == g1_mul_gen(sum(k * s for k, s in zip(ks, sc)))


def test_g1_msm_edge_cases(api):
    rnd = random.Random(7)
    # empty -> identity
    out = api.g1_multi_exp_affine_bytes(api.to_device(b""), api.to_device(b""))
    assert g1_jac_affine(api.to_host(out)) is None
    # duplicate points (forces the doubling branch), zero and r-1 scalars, bit 255 ignored,
    # a point and its negation cancelling
    p = g1_mul_gen(12345)
    pn = (p[0], (-p[1]) % O.P)
    pts = [p, p, p, pn, p, g1_mul_gen(7), g1_mul_gen(9)]
    sc = [5, 5, 0, 3, 3, O.R - 1, (1 << 255) | 11]
    out = api.g1_multi_exp_affine_bytes(api.to_device(g1_points_bytes(pts)),
                                        api.to_device(pack_scalars(sc)))
    want = g1_mul_gen(12345 * (5 + 5 + 0 - 3 + 3) + 7 * (O.R - 1) + 9 * 11)
    assert g1_jac_affine(api.to_host(out)) == want
    # all scalars identical (every entry in one bucket per window) and sparse scalars
    n = 300
    ks, pts = _rand_bases(rnd, n, g1_mul_gen)
    for sc in ([0xDEADBEEF] * n, [1 << rnd.randrange(255) for _ in range(n)]):
        out = api.g1_multi_exp_affine_bytes(api.to_device(g1_points_bytes(pts)),
                                            api.to_device(pack_scalars(sc)))
        assert g1_jac_affine(api.to_host(out)) == g1_mul_gen(sum(k * s for k, s in zip(ks, sc)))


def test_g1_msm_length_mismatch_raises(api):
    with pytest.raises(ValueError):
        api.g1_multi_exp_affine_bytes(api.to_device(bytes(96)), api.to_device(bytes(31)))


@pytest.mark.parametrize("n", [1, 5, 40])
def test_g2_msm_small_vs_oracle(api, n):
    rnd = random.Random(200 + n)
    ks, pts = _rand_bases(rnd, n, g2_mul_gen)
    sc = [rnd.randrange(O.R) for _ in range(n)]
    out = api.g2_multi_exp_affine_bytes(api.to_device(g2_points_bytes(pts)),
                                        api.to_device(pack_scalars(sc)))
    got = g2_jac_affine(api.to_host(out))
    assert got == O.G2.to_affine(O.msm(O.G2, pts, sc))
    assert O.g2_compress(got) == O.g2_compress(g2_mul_gen(sum(k * s for k, s in zip(ks, sc))))


def test_g2_msm_indexed(api):
    rnd = random.Random(300)
    npts = 50
    ks, pts = _rand_bases(rnd, npts, g2_mul_gen)
    idx = [rnd.randrange(npts) for _ in range(30)] + [3, 3, 3]
    sc = [rnd.randrange(O.R) for _ in idx]
    it = torch.tensor(idx, dtype=torch.int32, device="cuda")
    out = api.g2_multi_exp_affine_indexed(api.to_device(g2_points_bytes(pts)), it,
                                          api.to_device(pack_scalars(sc)))
    want = O.msm_indexed(O.G2, pts, idx, sc)
    assert g2_jac_affine(api.to_host(out)) == O.G2.to_affine(want)
    from weighted_bte_amd import BteError
    bad = torch.tensor([0, npts], dtype=torch.int32, device="cuda")
    with pytest.raises(BteError):
        api.g2_multi_exp_affine_indexed(api.to_device(g2_points_bytes(pts)), bad,
                                        api.to_device(pack_scalars([1, 2])))


def test_generator_mul_and_batch_normalize(api):
    rnd = random.Random(400)
    ks = [0, 1, 2, O.R - 1] + [rnd.randrange(O.R) for _ in range(60)]
    a1 = api.to_host(api.g1_generator_mul(api.to_device(pack_scalars(ks))))
    for i, k in enumerate(ks):
        assert O.g1_affine_from_bytes(a1[96 * i: 96 * i + 96]) == (g1_mul_gen(k) if k % O.R else None)
    ks2 = ks[:12]
    a2 = api.to_host(api.g2_generator_mul(api.to_device(pack_scalars(ks2))))
    for i, k in enumerate(ks2):
        assert O.g2_affine_from_bytes(a2[192 * i: 192 * i + 192]) == (g2_mul_gen(k) if k % O.R else None)
    # batch normalisation of non-normalised Jacobian inputs (random Z) and identities
    jacs = []
    for k in ks[:40]:
        p = O.G1.mul(O.g1_gen_jac(), k)
        z = rnd.randrange(1, O.P)
        if not O.G1.is_identity(p):
            x, y, zz = p
            p = (x * z * z % O.P, y * z * z * z % O.P, zz * z % O.P)
        jacs.append(p)
    raw = b"".join(O.g1_jac_to_bytes(p) for p in jacs)
    aff = api.to_host(api.batch_normalize_g1(api.to_device(raw)))
    for i, p in enumerate(jacs):
        assert O.g1_affine_from_bytes(aff[96 * i: 96 * i + 96]) == O.G1.to_affine(p)


@pytest.mark.parametrize("logn", [12, 16])
def test_g1_msm_closed_form_bench_sizes(api, logn):
    """The standalone-MSM benchmark shape: bases k_i*G from the seed, 90% uniform + 10% sparse scalars;
    MSM == (sum k_i s_i mod r) * G, compared as compressed bytes."""
    n = 1 << logn
    rng = O.SeededRng(2, logn)
    ks = [rng.scalar() for _ in range(n)]
    sc = O.msm_config_scalars(rng, n)
    bases = api.g1_generator_mul(api.to_device(pack_scalars(ks)))
    # spot-check the generated bases against the oracle
    hb = api.to_host(bases[: 96 * 8])
    for i in range(8):
        assert O.g1_affine_from_bytes(hb[96 * i: 96 * i + 96]) == g1_mul_gen(ks[i])
    out = api.g1_multi_exp_affine_bytes(bases, api.to_device(pack_scalars(sc)))
    want = g1_mul_gen(sum(k * s for k, s in zip(ks, sc)))
    assert O.g1_compress(g1_jac_affine(api.to_host(out))) == O.g1_compress(want)


def test_g2_msm_closed_form_2_12(api):
    n = 1 << 12
    rng = O.SeededRng(3, 12)
    ks = [rng.scalar() for _ in range(n)]
    sc = O.msm_config_scalars(rng, n)
    bases = api.g2_generator_mul(api.to_device(pack_scalars(ks)))
    out = api.g2_multi_exp_affine_bytes(bases, api.to_device(pack_scalars(sc)))
    want = g2_mul_gen(sum(k * s for k, s in zip(ks, sc)))
    assert O.g2_compress(g2_jac_affine(api.to_host(out))) == O.g2_compress(want)


@pytest.mark.parametrize("c", [5, 9, 13])
def test_g1_msm_window_override_same_result(api, c):
    rnd = random.Random(500 + c)
    n = 1000
    ks, pts = _rand_bases(rnd, 40, g1_mul_gen)
    pts = [pts[i % 40] for i in range(n)]
    ks = [ks[i % 40] for i in range(n)]
    sc = [rnd.randrange(O.R) for _ in range(n)]
    api.set_msm_window(1, n, c)
    try:
        out = api.g1_multi_exp_affine_bytes(api.to_device(g1_points_bytes(pts)),
                                            api.to_device(pack_scalars(sc)))
    finally:
        api.set_msm_window(1, n, 0)
    assert g1_jac_affine(api.to_host(out)) == g1_mul_gen(sum(k * s for k, s in zip(ks, sc)))


@pytest.mark.parametrize("c,n", [(12, 5000), (16, 20001), (17, 3000)])
def test_g1_msm_sort_paths(api, c, n):
    """The LDS-histogram sort at its 8192-counter tier (c = 12) and its 32768-counter tier with
    several point slices and a short last slice (c = 16), and the global-counter sort of windows
    wider than 16 bits (c = 17); each with uniform scalars, identical scalars (one bucket per
    window spanning every accumulation chunk) and one-bit scalars (pile-ups on power-of-two
    buckets).  Closed form (sum k_i s_i) G, compared as compressed bytes."""
    rng = random.Random(900 + c)
    ks = [rng.randrange(1, O.R) for _ in range(n)]
    bases = api.g1_generator_mul(api.to_device(pack_scalars(ks)))
    api.set_msm_window(1, n, c)
    try:
        for sc in ([rng.randrange(O.R) for _ in range(n)], [0xDEADBEEF] * n,
                   [1 << rng.randrange(255) for _ in range(n)]):
            out = api.g1_multi_exp_affine_bytes(bases, api.to_device(pack_scalars(sc)))
            want = g1_mul_gen(sum(k * s for k, s in zip(ks, sc)) % O.R)
            assert O.g1_compress(g1_jac_affine(api.to_host(out))) == O.g1_compress(want)
    finally:
        api.set_msm_window(1, n, 0)


@pytest.mark.parametrize("group", [1, 2])
def test_batched_msm_uneven_segments(api, group):
    """Many MSMs of uneven sizes (0, 1, 5, 37, 200, 3 terms) in one launch sequence, each equal
    to its closed form; G2 also with gathered (indexed) points."""
    rnd = random.Random(600 + group)
    sizes = [0, 1, 5, 37, 200, 3] if group == 1 else [0, 1, 5, 37, 3]
    mul = g1_mul_gen if group == 1 else g2_mul_gen
    pack = g1_points_bytes if group == 1 else g2_points_bytes
    read = g1_jac_affine if group == 1 else g2_jac_affine
    width = 144 if group == 1 else 288
    npts = 40
    ks, pts = _rand_bases(rnd, npts, mul)
    nterms = sum(sizes)
    idx = [rnd.randrange(npts) for _ in range(nterms)]
    sc = [rnd.randrange(O.R) if rnd.random() < 0.8 else 1 << rnd.randrange(255) for _ in range(nterms)]
    offs = [0]
    for s in sizes:
        offs.append(offs[-1] + s)
    fn = api.g1_multi_exp_batched if group == 1 else api.g2_multi_exp_batched
    out = api.to_host(fn(api.to_device(pack(pts)), api.to_device(pack_scalars(sc)), offs,
                         torch.tensor(idx, dtype=torch.int32, device="cuda")))
    for sgm in range(len(sizes)):
        want = sum(ks[idx[t]] * sc[t] for t in range(offs[sgm], offs[sgm + 1]))
        got = read(out[width * sgm: width * (sgm + 1)])
        assert got == (mul(want) if want % O.R else None), "segment %d" % sgm
    if group == 1:
        # contiguous (non-gathered) points
        pts2 = [pts[i] for i in idx]
        out = api.to_host(fn(api.to_device(pack(pts2)), api.to_device(pack_scalars(sc)), offs))
        for sgm in range(len(sizes)):
            want = sum(ks[idx[t]] * sc[t] for t in range(offs[sgm], offs[sgm + 1]))
            assert read(out[width * sgm: width * (sgm + 1)]) == (mul(want) if want % O.R else None)


@pytest.mark.parametrize("group", [1, 2])
def test_batched_msm_reports_out_of_range_index(api, group):
    """An index at or beyond n_points in a batched, gathered MSM is skipped on the device and
    reported through the status word (the wrapper raises); a valid call reports BTE_OK."""
    import ctypes

    from weighted_bte_amd import call

    rnd = random.Random(700 + group)
    mul = g1_mul_gen if group == 1 else g2_mul_gen
    pack = g1_points_bytes if group == 1 else g2_points_bytes
    fn = api.g1_multi_exp_batched if group == 1 else api.g2_multi_exp_batched
    ks, pts = _rand_bases(rnd, 8, mul)
    sc = [rnd.randrange(O.R) for _ in range(10)]
    offs = [0, 4, 10]
    good = [rnd.randrange(8) for _ in range(10)]
    fn(api.to_device(pack(pts)), api.to_device(pack_scalars(sc)), offs,
       torch.tensor(good, dtype=torch.int32, device="cuda"))
    for bad_at, bad in ((2, 8), (9, 1 << 20)):
        idx = list(good)
        idx[bad_at] = bad
        with pytest.raises(ValueError):
            fn(api.to_device(pack(pts)), api.to_device(pack_scalars(sc)), offs,
               torch.tensor(idx, dtype=torch.int32, device="cuda"))
    # through the C ABI directly: status is BTE_EINVAL (1)
    name = "bte_g1_msm_batched" if group == 1 else "bte_g2_msm_batched"
    P = api.to_device(pack(pts))
    S = api.to_device(pack_scalars(sc))
    I = torch.tensor(good[:9] + [8], dtype=torch.int32, device="cuda")
    off = torch.tensor(offs, dtype=torch.int64, device="cuda")
    out = torch.empty(2 * (144 if group == 1 else 288), dtype=torch.uint8, device="cuda")
    status = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    call(name, ptr(P), 8, ptr(I), ptr(S), 10, ptr(off), 2, ptr(out), ptr(status),
         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert int(status.item()) == 1


def test_g2_msm_closed_form_2_16(api):
    """Config-2 G2 at 2^16 (bases k_i * G2, 10% sparse scalars), compressed-byte equality."""
    n = 1 << 16
    rng = O.SeededRng(3, 16)
    ks = [rng.scalar() for _ in range(n)]
    sc = O.msm_config_scalars(rng, n)
    bases = api.g2_generator_mul(api.to_device(pack_scalars(ks)))
    out = api.g2_multi_exp_affine_bytes(bases, api.to_device(pack_scalars(sc)))
    want = g2_mul_gen(sum(k * s for k, s in zip(ks, sc)))
    assert O.g2_compress(g2_jac_affine(api.to_host(out))) == O.g2_compress(want)


def test_autotune_window_keeps_result(api, tmp_path):
    """The autotuner sweeps windows, installs the fastest and the MSM result is unchanged; its
    table round-trips through save/apply."""
    from weighted_bte_amd import autotune

    n = 1 << 12
    rng = O.SeededRng(9, 12)
    ks = [rng.scalar() for _ in range(n)]
    sc = O.msm_config_scalars(rng, n)
    bases = api.g1_generator_mul(api.to_device(pack_scalars(ks)))
    dsc = api.to_device(pack_scalars(sc))
    want = O.g1_compress(g1_mul_gen(sum(k * s for k, s in zip(ks, sc))))
    try:
        best, times = autotune.tune_msm(1, bases, dsc, candidates=[6, 8, 10], reps=1)
        assert set(times) == {6, 8, 10} and best in times
        assert O.g1_compress(g1_jac_affine(api.to_host(api.g1_multi_exp_affine_bytes(bases, dsc)))) == want
        path = str(tmp_path / "w.json")
        autotune.save_table({("g1", 12): 7}, path)
        assert autotune.apply_table(path) == 1
        assert O.g1_compress(g1_jac_affine(api.to_host(api.g1_multi_exp_affine_bytes(bases, dsc)))) == want
    finally:
        api.set_msm_window(1, n, 0)


@pytest.mark.parametrize("logn", [14, 18, 19])
def test_g1_msm_built_in_window_table(api, logn):
    """The built-in window table at sizes where it picks 8, 13 and 16 bits (2^14, 2^18, 2^19;
    the former log2 n - 4 rule gave 10, 14 and 15): closed form (sum k_i s_i) G with the bench's
    10%-sparse scalars, compared as compressed bytes (blst_utils.rs:43-67)."""
    from weighted_bte_amd import autotune, inputs

    n = 1 << logn
    assert autotune.default_window(n) == {14: 8, 18: 13, 19: 16}[logn]
    st = inputs.Stream(77, logn)
    ks, sc = st.scalars(n), inputs.msm_scalars(st, n)
    bases = api.g1_generator_mul(api.to_device(inputs.pack_scalars(ks)))
    out = api.g1_multi_exp_affine_bytes(bases, api.to_device(inputs.pack_scalars(sc)))
    want = api.to_host(api.g1_generator_mul(api.to_device((sum(k * s for k, s in zip(ks, sc)) % O.R).to_bytes(32, "little"))))
    assert api.to_host(api.batch_normalize_g1(out)) == want
