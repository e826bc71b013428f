"""End-to-end weighted_btx batch decryption on the GPU (weighted_btx/src/decryption.rs):
decrypt(encrypt(m)) == m bit-exactly, and beta equals the reference's literal weighted
middle-product formula (decryption.rs:806-842) in closed form."""
import random

import pytest

from btx_setup import BtxSetup, heaviest_first_committee
from helpers import O, g1_mul_gen, g2_mul_gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def read_f12(api, t, n):
    b = api.to_host(t)
    return [O.f12_from_bytes(b[576 * i: 576 * i + 576]) for i in range(n)]


def run_btx(api, st, B, invalid=()):
    ct1, ct2 = st.encrypt_batch(B)
    valid = [i not in invalid for i in range(B)]
    if invalid:
        raw = bytearray(api.to_host(ct1))
        for i in invalid:
            raw[96 * i: 96 * i + 96] = bytes(96)
        ct1 = api.to_device(bytes(raw))
    committee = heaviest_first_committee(st.weights, st.t)
    sigma_all = api.btx_partial_decrypt(ct1, st.q)
    shares = b"".join(api.to_host(sigma_all[144 * j: 144 * j + 144]) for j in committee)
    lines, keys = api.btx_prepare(st.material, st.W, st.L, st.offsets, committee, B)
    beta = api.btx_precompute(ct1, valid, lines)
    msgs = api.btx_open(api.to_device(shares), keys, ct2, beta, valid)
    return committee, valid, beta, msgs


def test_lagrange_matches_oracle_definition(api):
    rnd = random.Random(3)
    W = 45
    dom = O.Radix2Domain(64)
    for sel in ([0], [3, 1, 7], sorted(rnd.sample(range(W), 20)), list(range(W))):
        got = api.lagrange_at_zero(W, sel)
        xs = [pow(dom.generator, v, O.R) for v in sel]
        for i, xi in enumerate(xs):
            num, den = 1, 1
            for j, xj in enumerate(xs):
                if j != i:
                    num = num * xj % O.R
                    den = den * (xj - xi) % O.R
            assert got[i] == num * pow(den, -1, O.R) % O.R
    # complete power-of-two domain: every coefficient is 1/M (interpolation.rs:325-338)
    assert api.lagrange_at_zero(16, list(range(16))) == [pow(16, -1, O.R)] * 16


def test_lagrange_large_subset_interpolates(api):
    """700 of the first 1000 points of the size-1024 domain (the device path for the n^2
    denominators): the coefficients reproduce f(0) from f(x_v) for a random f of degree 699, and
    spot coefficients equal the definition."""
    rnd = random.Random(8)
    W = 1000
    sel = sorted(rnd.sample(range(W), 700))
    got = api.lagrange_at_zero(W, sel)
    dom = O.Radix2Domain(1024)
    xs = [pow(dom.generator, v, O.R) for v in sel]
    f = [rnd.randrange(O.R) for _ in range(700)]

    def ev(x):
        acc = 0
        for c in reversed(f):
            acc = (acc * x + c) % O.R
        return acc
    assert sum(l * ev(x) for l, x in zip(got, xs)) % O.R == f[0]
    for i in (0, 349, 699):
        num, den = 1, 1
        for j, xj in enumerate(xs):
            if j != i:
                num = num * xj % O.R
                den = den * (xj - xs[i]) % O.R
        assert got[i] == num * pow(den, -1, O.R) % O.R


@pytest.mark.parametrize("B", [1, 2, 3, 7, 8])
def test_btx_beta_matches_literal_formula_and_decrypts(api, B):
    st = BtxSetup(api, [1, 2, 3], threshold=3, max_batch=8, seed=B)
    committee, valid, beta, msgs = run_btx(api, st, B)
    sel = [v for j in committee for v in range(st.offsets[j], st.offsets[j + 1])]
    lam = api.lagrange_at_zero(st.W, sel)
    g = O.gt_generator()
    got_beta = read_f12(api, beta, B)
    for i in range(B):
        e = sum(st.r[k] * st.kappa(sel, lam, k - i) for k in range(B) if k != i) % O.R
        assert got_beta[i] == O.f12_pow(g, e), "beta slot %d" % i
    got = read_f12(api, msgs, B)
    assert got == [O.f12_pow(g, s) for s in st.s]


def test_btx_small_committee_with_invalid_slot(api):
    """n = 8 parties, weights U[1,8] from the seed, threshold = W // 2, B = L = 32; slot 5 is
    invalid (identity first component) and opens to None (one)."""
    rng = O.SeededRng(11, 1)
    weights = [1 + rng.below(8) for _ in range(8)]
    W = sum(weights)
    st = BtxSetup(api, weights, threshold=W // 2, max_batch=32, seed=11)
    committee, valid, beta, msgs = run_btx(api, st, 32, invalid=(5,))
    got = read_f12(api, msgs, 32)
    g = O.gt_generator()
    for i in range(32):
        if i == 5:
            assert got[i] == O.F12_ONE
        else:
            assert got[i] == O.f12_pow(g, st.s[i]), "slot %d" % i


def test_btx_fused_batches_equal_alone(api):
    """Each result is the same alone (plain per-call path) and in the fused multi-batch path."""
    from weighted_bte_amd import workloads

    weights = [3, 1, 4, 1, 5]
    wl = workloads.BtxWorkload(weights, threshold=7, max_batch=4, seed=9)
    committee = workloads.heaviest_first(weights, wl.t)
    T, B, K = len(committee), 4, 3
    ct1, ct2, msgs = wl.encrypt(B, K)
    valid = [True] * (B * K)
    valid[5] = False  # batch 1, slot 1
    raw = bytearray(api.to_host(ct1))
    raw[96 * 5: 96 * 6] = bytes(96)
    ct1 = api.to_device(bytes(raw))
    lines, keys = api.btx_prepare(wl.material, wl.W, wl.L, wl.offsets, committee, B)
    shares = []
    for b in range(K):
        sig = api.btx_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.q)
        shares.append(b"".join(api.to_host(sig[144 * j:144 * (j + 1)]) for j in committee))
    sh = api.to_device(b"".join(shares))
    beta = api.btx_precompute_batched(ct1, valid, lines, B)
    fused = api.to_host(api.btx_open_batched(sh, T, keys, ct2, beta, valid, B))
    want = api.to_host(msgs)
    for b in range(K):
        sl = slice(b * B, (b + 1) * B)
        be = api.btx_precompute(ct1[b * B * 96:(b + 1) * B * 96], valid[sl], lines)
        alone = api.to_host(api.btx_open(sh[b * T * 144:(b + 1) * T * 144], keys,
                                         ct2[b * B * 576:(b + 1) * B * 576], be, valid[sl]))
        assert alone == fused[b * B * 576:(b + 1) * B * 576]
    one = O.f12_to_bytes(O.F12_ONE)
    for i in range(B * K):
        assert fused[576 * i:576 * (i + 1)] == (want[576 * i:576 * (i + 1)] if valid[i] else one)


def test_btx_open_prepared_equals_unprepared(api):
    """Opening against the keys' line tables (all slots valid, and with invalid slots) gives the
    same messages as the unprepared multi-Miller opening."""
    from weighted_bte_amd import workloads

    weights = [2, 3, 1, 4, 2, 5]
    wl = workloads.BtxWorkload(weights, threshold=8, max_batch=8, seed=13)
    committee = workloads.heaviest_first(weights, wl.t)
    T, B, K = len(committee), 8, 2
    ct1, ct2, msgs = wl.encrypt(B, K)
    lines, keys = api.btx_prepare(wl.material, wl.W, wl.L, wl.offsets, committee, B)
    key_lines = api.prepare_g2_lines(keys)
    sh = []
    for b in range(K):
        sig = api.btx_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.q)
        sh.append(b"".join(api.to_host(sig[144 * j:144 * (j + 1)]) for j in committee))
    sh = api.to_device(b"".join(sh))
    results = []
    for valid in ([True] * (B * K), [i % 5 != 2 for i in range(B * K)]):
        beta = api.btx_precompute_batched(ct1, valid, lines, B)
        a = api.to_host(api.btx_open_batched(sh, T, keys, ct2, beta, valid, B))
        p = api.to_host(api.btx_open_prepared_batched(sh, T, key_lines, ct2, beta, valid, B))
        assert a == p
        results.append(p)
    # with every slot valid the shares match the batch, so the plaintexts come back
    assert results[0] == api.to_host(msgs)
    one = O.f12_to_bytes(O.F12_ONE)
    assert all(results[1][576 * i:576 * (i + 1)] == one for i in range(B * K) if i % 5 == 2)


def test_multi_miller_prepared_uneven_segments(api):
    """Prepared multi-Miller loops over uneven segments (0, 1, 3, 9, 17 pairs, with identity
    points) equal the unprepared ones bit for bit (same lines, products reordered)."""
    import ctypes

    import torch

    from weighted_bte_amd import call

    rnd = random.Random(77)
    arities = [0, 1, 3, 9, 17]
    n = sum(arities)
    ps = [g1_mul_gen(rnd.randrange(1, O.R)) for _ in range(n)]
    qs = [g2_mul_gen(rnd.randrange(1, O.R)) for _ in range(n)]
    ps[4] = None  # identity G1 point in segment 3
    qs[20] = None  # identity G2 point in segment 4
    P = api.to_device(b"".join(O.g1_affine_to_bytes(x) for x in ps))
    Q = api.to_device(b"".join(O.g2_affine_to_bytes(x) for x in qs))
    off = [0]
    for a in arities:
        off.append(off[-1] + a)
    want = api.unprepared_multi_miller_loop(P, Q, off)
    lines = api.prepare_g2_lines(Q)
    base = lines.data_ptr()
    tabs = torch.tensor([base + 19584 * g for g in range(n)], dtype=torch.int64).cuda()
    out = torch.empty(576 * len(arities), dtype=torch.uint8, device="cuda")
    hoff = (ctypes.c_uint64 * len(off))(*off)
    call("bte_multi_miller_loop_prepared", ctypes.c_void_p(tabs.data_ptr()), ctypes.c_void_p(P.data_ptr()), n,
         hoff, len(arities), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert api.to_host(out) == api.to_host(want)


@pytest.mark.parametrize("B,K", [(8, 3), (64, 8)])
def test_btx_plan_run_and_graph_replays(api, B, K):
    """The launch-only plan gives the fused entry points' messages; its HIP graph, replayed three
    times with eager work and synchronisation between replays, gives them every time."""
    import torch

    from weighted_bte_amd import workloads

    weights = [2, 3, 1, 4]
    wl = workloads.BtxWorkload(weights, threshold=5, max_batch=B, seed=31 + B)
    committee = workloads.heaviest_first(weights, wl.t)
    T = len(committee)
    ct1, ct2, msgs = wl.encrypt(B, K)
    valid = [True] * (B * K)
    valid[1] = False
    raw = bytearray(api.to_host(ct1))
    raw[96:192] = bytes(96)
    ct1 = api.to_device(bytes(raw))
    lines, keys = api.btx_prepare(wl.material, wl.W, wl.L, wl.offsets, committee, B)
    key_lines = api.prepare_g2_lines(keys)
    sh = []
    for b in range(K):
        sig = api.btx_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.q)
        sh.append(torch.cat([sig[144 * j:144 * (j + 1)] for j in committee]))
    shares = torch.cat(sh)
    beta = api.btx_precompute_batched(ct1, valid, lines, B)
    want = api.to_host(api.btx_open_prepared_batched(shares, T, key_lines, ct2, beta, valid, B))
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        plan = api.BtxPlan(B, K, T, lines, key_lines, valid)
        assert api.to_host(plan.run(ct1, shares, ct2)) == want
        out = torch.empty(B * K * 576, dtype=torch.uint8, device="cuda")
        for rep in range(3):
            out.zero_()
            plan.launch_graph(ct1, shares, ct2, out)
            stream.synchronize()
            assert api.to_host(out) == want, "replay %d" % rep
            api.g1_generator_mul(api.to_device((rep + 5).to_bytes(32, "little")))  # eager work between
            torch.cuda.synchronize()
        plan.close()
    one = O.f12_to_bytes(O.F12_ONE)
    assert want[576:1152] == one


def test_gpu_keygen_equals_host_tables(api):
    """Device key generation (Horner evaluation, power tables, fixed-base G2 multiplications)
    gives byte-identical material to the host-computed scalar tables, and the verification keys
    are [q_j^(i+1)]_2."""
    import torch

    from weighted_bte_amd import inputs, workloads

    weights, L = [2, 1, 3, 1], 4
    wl = workloads.BtxWorkload(weights, threshold=3, max_batch=L, seed=41)
    poly = inputs.Stream(41, 4000).scalars(4)
    host = wl.host_material(poly)
    assert torch.equal(wl.material, host)
    mat, vk = api.btx_keygen(poly, wl.q, wl.offsets, L, verification_keys=True)
    assert torch.equal(mat, host)
    pows = [pow(q, i + 1, O.R) for i in range(L) for q in wl.q]
    want = api.g2_generator_mul(api.to_device(b"".join(p.to_bytes(32, "little") for p in pows)))
    assert torch.equal(vk, want)


def test_share_verification_each_and_batched(api):
    """Valid shares pass both checks; a share moved by G1 fails, the batched check fails and
    the per-share check blames exactly that party (decryption.rs:236-422)."""
    import random as rnd_mod

    import torch

    from weighted_bte_amd import inputs, workloads

    weights, L, B = [3, 1, 2, 4, 2], 8, 8
    wl = workloads.BtxWorkload(weights, threshold=6, max_batch=L, seed=51)
    poly = inputs.Stream(51, 4000).scalars(7)
    _, vk = api.btx_keygen(poly, wl.q, wl.offsets, L, verification_keys=True)
    ct1, ct2, msgs = wl.encrypt(B)
    valid = [True] * B
    valid[3] = False
    raw = bytearray(api.to_host(ct1))
    raw[3 * 96:4 * 96] = bytes(96)
    ct1 = api.to_device(bytes(raw))
    sig = api.btx_partial_decrypt(ct1, wl.q)
    parties = [0, 2, 3, 4]
    shares = torch.cat([sig[144 * j:144 * (j + 1)] for j in parties])
    rnd = rnd_mod.Random(5)
    rho = [rnd.randrange(1, O.R) for _ in parties]
    assert api.btx_verify_shares(shares, parties, vk, len(weights), L, ct1, valid) == [True] * 4
    assert api.btx_verify_shares_batched(shares, parties, rho, vk, len(weights), L, ct1, valid)
    # tamper with party 3's share: sigma + G1
    bad = api.to_host(api.batch_normalize_g1(shares[2 * 144:3 * 144]))
    bad_pt = O.G1.add(O.G1.from_affine(O.g1_affine_from_bytes(bad)), O.g1_gen_jac())
    tampered = shares.clone()
    tampered[2 * 144:3 * 144] = api.to_device(O.g1_jac_to_bytes(bad_pt))
    assert api.btx_verify_shares(tampered, parties, vk, len(weights), L, ct1, valid) == [True, True, False, True]
    assert not api.btx_verify_shares_batched(tampered, parties, rho, vk, len(weights), L, ct1, valid)
    # a zero challenge would drop the tampered share from the folded check: refused, as is a
    # non-canonical one (decryption.rs:363-370 draws nonzero challenges)
    from weighted_bte_amd import BteError
    for bad in (0, O.R):
        rho_bad = list(rho)
        rho_bad[2] = bad
        with pytest.raises(BteError):
            api.btx_verify_shares_batched(tampered, parties, rho_bad, vk, len(weights), L, ct1, valid)
