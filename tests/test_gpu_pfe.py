"""End-to-end weighted_pfe batch decryption on the GPU (weighted_pfe/src/decryption.rs):
decrypt(encrypt(m)) == m bit-exactly, and both Cauchy transforms equal the reference's literal
quadratic sums (decryption.rs:981-1039) in closed form."""
import pytest

from helpers import O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def read_f12(api, t, n):
    b = api.to_host(t)
    return [O.f12_from_bytes(b[576 * i: 576 * i + 576]) for i in range(n)]


@pytest.mark.parametrize("B", [1, 2, 4, 8])
def test_pfe_decrypts_and_cauchy_matches_literal(api, B):
    from weighted_bte_amd import workloads

    weights = [1, 2, 3]
    wl = workloads.PfeWorkload(weights, threshold=3, batch_size=B, seed=B)
    committee = workloads.heaviest_first(weights, wl.t)
    ct1, ct2, msgs = wl.encrypt()
    sigma = api.pfe_partial_decrypt(ct1, wl.g)
    shares = api.to_device(b"".join(api.to_host(sigma[144 * j:144 * (j + 1)]) for j in committee))
    d1p, tl, ml = api.pfe_prepare(wl.material, wl.W, B, wl.offsets, committee, wl.alpha, wl.global_key)
    fc, pc = api.pfe_precompute(ct1, tl)
    out = api.pfe_open(shares, d1p, ml, ct1, ct2, fc, pc, wl.alpha)
    assert api.to_host(out) == api.to_host(msgs)
    # literal Cauchy sums: first_cauchy_i = 2B sum_{k != i} r_k / (a_k - a_i) * G1
    fca = api.to_host(api.batch_normalize_g1(fc))
    for i in range(B):
        e = sum(2 * B * wl.r[k] * pow(wl.alpha[k] - wl.alpha[i], -1, O.R) for k in range(B) if k != i) % O.R
        want = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), e)) if e else None
        assert O.g1_affine_from_bytes(fca[96 * i:96 * i + 96]) == want
    # pairing Cauchy after the hard part: gT^(sum_{k != i} r_k kappa_k / (a_k - a_i)),
    # kappa_k = sum_{v in S} lambda_v Z(w_v) g_{owner(v), k}
    sel = [v for j in committee for v in range(wl.offsets[j], wl.offsets[j + 1])]
    lam = api.lagrange_at_zero(wl.W, sel)
    kappa = [sum(l * wl.zv[v] * wl.g[wl.owner[v]][k] for v, l in zip(sel, lam)) % O.R for k in range(B)]
    g = O.gt_generator()
    got = read_f12(api, api.hard_final_exponentiation(pc), B)
    for i in range(B):
        e = sum(wl.r[k] * kappa[k] * pow(wl.alpha[k] - wl.alpha[i], -1, O.R) for k in range(B) if k != i) % O.R
        assert got[i] == O.f12_pow(g, e), "pairing cauchy slot %d" % i


def test_pfe_fused_batches_equal_alone(api):
    """Three batches through the fused precompute / prepared opening give the plain per-batch
    results and the plaintexts."""
    import torch

    from weighted_bte_amd import workloads

    weights = [2, 1, 3, 2]
    B, K = 4, 3
    wl = workloads.PfeWorkload(weights, threshold=4, batch_size=B, seed=21)
    committee = workloads.heaviest_first(weights, wl.t)
    T = len(committee)
    ct1, ct2, msgs = wl.encrypt(K)
    d1p, tl, ml = api.pfe_prepare(wl.material, wl.W, B, wl.offsets, committee, wl.alpha, wl.global_key)
    shares = []
    for b in range(K):
        sig = api.pfe_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.g)
        shares.append(torch.cat([sig[144 * j:144 * (j + 1)] for j in committee]))
    shares = torch.cat(shares)
    fc, pc = api.pfe_precompute_batched(ct1, tl, B)
    fused = api.pfe_open_prepared_batched(shares, T, api.prepare_g2_lines(d1p), ml, ct1, ct2, fc, pc, wl.alpha, B)
    assert api.to_host(fused) == api.to_host(msgs)
    for b in range(K):
        c1 = ct1[b * B * 96:(b + 1) * B * 96]
        f1, p1 = api.pfe_precompute(c1, tl)
        assert api.to_host(f1) == api.to_host(fc[b * B * 144:(b + 1) * B * 144])
        assert api.to_host(p1) == api.to_host(pc[b * B * 576:(b + 1) * B * 576])
        alone = api.pfe_open(shares[b * T * 144:(b + 1) * T * 144], d1p, ml, c1, ct2[b * B * 576:(b + 1) * B * 576],
                             f1, p1, wl.alpha)
        assert api.to_host(alone) == api.to_host(fused[b * B * 576:(b + 1) * B * 576])


def test_pfe_proof_batch_verification(api):
    """Schnorr proofs s_i G = A_i + e_i C_i pass the folded one-MSM check; one wrong response
    fails it (weighted_pfe/src/proof.rs:74-122)."""
    import random

    rnd = random.Random(12)
    B = 16
    r = [rnd.randrange(1, O.R) for _ in range(B)]
    k = [rnd.randrange(1, O.R) for _ in range(B)]
    e = [rnd.randrange(O.R) for _ in range(B)]
    resp = [(ki + ei * ri) % O.R for ki, ei, ri in zip(k, e, r)]
    coef = [rnd.randrange(1, O.R) for _ in range(B)]
    pack = lambda v: api.to_device(b"".join(x.to_bytes(32, "little") for x in v))
    ct1 = api.g1_generator_mul(pack(r))
    commit = api.g1_generator_mul(pack(k))
    assert api.pfe_verify_proofs_batched(commit, ct1, resp, e, coef)
    bad = list(resp)
    bad[5] = (bad[5] + 1) % O.R
    assert not api.pfe_verify_proofs_batched(commit, ct1, bad, e, coef)


def test_pfe_gpu_keygen_matches_host(api):
    """bte_pfe_keygen (weighted_pfe/src/setup.rs:149-354 on the device) equals the host-computed
    scalar tables pushed through the generator: d1 || d2 byte for byte, the global key, the
    fractions g_{j,i} = 1/(rho_j + alpha_i), and the slot-major verification keys [g_{j,i}]_2."""
    from weighted_bte_amd import workloads
    weights = [3, 1, 4, 2, 5]
    B = 8
    wl = workloads.PfeWorkload(weights, threshold=6, batch_size=B, seed=33, gpu_keygen=False)
    mat, gk, g, vk = api.pfe_keygen(wl.poly, wl.rho, wl.alpha, (wl.beta_m1, wl.beta_0), wl.offsets, B,
                                    verification_keys=True)
    assert api.to_host(mat) == api.to_host(wl.material)
    assert api.to_host(gk) == api.to_host(wl.global_key)
    assert g == wl.g
    N = len(weights)
    want_vk = api.g2_generator_mul(api.to_device(b"".join(wl.g[j][i].to_bytes(32, "little")
                                                          for i in range(B) for j in range(N))))
    assert api.to_host(vk) == api.to_host(want_vk)
    for j in range(N):
        for i in range(B):
            assert g[j][i] * (wl.rho[j] + wl.alpha[i]) % O.R == 1


@pytest.mark.parametrize("B,K", [(8, 3), (64, 4)])
def test_pfe_plan_run_and_graph_replays(api, B, K):
    """The weighted_pfe plan gives the fused entry points' messages; its HIP graph, replayed three
    times with eager work and synchronisation between replays, gives them every time."""
    import torch

    from weighted_bte_amd import workloads

    weights = [2, 3, 1, 4]
    wl = workloads.PfeWorkload(weights, threshold=5, batch_size=B, seed=90 + B)
    committee = workloads.heaviest_first(weights, wl.t)
    T = len(committee)
    ct1, ct2, msgs = wl.encrypt(K)
    sh = []
    for b in range(K):
        sigma = api.pfe_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.g)
        sh.append(torch.cat([sigma[144 * j:144 * (j + 1)] for j in committee]))
    shares = torch.cat(sh)
    d1p, tl, ml = api.pfe_prepare(wl.material, wl.W, B, wl.offsets, committee, wl.alpha, wl.global_key)
    d1l = api.prepare_g2_lines(d1p)
    fc, pc = api.pfe_precompute_batched(ct1, tl, B)
    want = api.to_host(api.pfe_open_prepared_batched(shares, T, d1l, ml, ct1, ct2, fc, pc, wl.alpha, B))
    assert want == api.to_host(msgs)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        plan = api.PfePlan(B, K, T, tl, d1l, ml, wl.alpha)
        assert api.to_host(plan.run(ct1, shares, ct2)) == want
        out = torch.empty(B * K * 576, dtype=torch.uint8, device="cuda")
        for rep in range(3):
            out.zero_()
            plan.launch_graph(ct1, shares, ct2, out)
            stream.synchronize()
            assert api.to_host(out) == want, "replay %d" % rep
            api.pfe_precompute_batched(ct1, tl, B)  # eager work between replays
            torch.cuda.synchronize()
        plan.close()
