"""GPU parity at the full workload shapes, checked against the CPU oracle.

  * batched weighted_btx and weighted_pfe decryption at B = 512 with n = 128 Pareto(1.16)
    parties and W = 4096 (fused over two batches, and btx also as a captured HIP graph);
  * 512 multi-pairings of log-uniform arity 2..512 in one call;
  * the serving loop with batch sizes spread over 16..2048 and a committee of about 1000
    parties with total weight near 2^14;
  * the GT FFT at 2^8 and the G1 MSM at 2^20.

Every plaintext is compared with the plain per-call path, and sampled slots with gT^s computed by
the oracle (O.f12_pow of the oracle's own pairing of the generators) from exponents kept on the
host; group outputs are compared with closed forms evaluated by the oracle.
"""
import math
import random

import pytest

from helpers import O

pytestmark = pytest.mark.gpu

N_SAMPLE = 32


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


@pytest.fixture(scope="module")
def gt():
    return O.gt_generator()


def f12_at(raw, i):
    return O.f12_from_bytes(raw[576 * i:576 * (i + 1)])


def check_sampled(raw, exps, gt, rnd, n=N_SAMPLE, what="slot"):
    """Sampled Fp12 slots of `raw` against gT^e by the oracle (always slots 0 and the last)."""
    idx = sorted(set([0, len(exps) - 1] + rnd.sample(range(len(exps)), min(n, len(exps)))))
    for i in idx:
        assert f12_at(raw, i) == O.f12_pow(gt, exps[i] % O.R), "%s %d differs from the oracle" % (what, i)
    return len(idx)


def test_btx_config4_fused_graph_and_plain_vs_oracle(api, gt):
    """weighted_btx at B = 512, n = 128 Pareto(1.16), W = 4096, two batches: fused precompute +
    prepared opening, the launch-only plan replayed as a HIP graph, and the plain per-call path
    agree slot for slot; sampled plaintexts equal the oracle's gT^s."""
    import torch

    from weighted_bte_amd import workloads

    B, K = 512, 2
    weights = workloads.pareto_weights(128, 4096)
    assert sum(weights) == 4096 and len(weights) == 128 and min(weights) >= 1
    wl = workloads.BtxWorkload(weights, 2048, B, seed=5)
    committee = workloads.heaviest_first(weights, wl.t)
    T = len(committee)
    assert sum(weights[j] for j in committee) > wl.t
    ct1, ct2, msgs = wl.encrypt(B, K)
    s = list(wl.s)
    valid = [True] * (B * K)
    sh = []
    for b in range(K):
        sig = api.btx_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.q)
        sh.append(torch.cat([sig[144 * j:144 * (j + 1)] for j in committee]))
    shares = torch.cat(sh)
    lines, keys = api.btx_prepare(wl.material, wl.W, wl.L, wl.offsets, committee, B)
    key_lines = api.prepare_g2_lines(keys)
    beta = api.btx_precompute_batched(ct1, valid, lines, B)
    fused = api.btx_open_prepared_batched(shares, T, key_lines, ct2, beta, valid, B)
    plain = []
    for b in range(K):
        sl = slice(b * B, (b + 1) * B)
        be = api.btx_precompute(ct1[b * B * 96:(b + 1) * B * 96], valid[sl], lines)
        plain.append(api.btx_open(shares[b * T * 144:(b + 1) * T * 144], keys, ct2[b * B * 576:(b + 1) * B * 576],
                                  be, valid[sl]))
    plain = torch.cat(plain)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        plan = api.BtxPlan(B, K, T, lines, key_lines, valid)
        out = torch.zeros(B * K * 576, dtype=torch.uint8, device="cuda")
        for rep in range(2):
            out.zero_()
            plan.launch_graph(ct1, shares, ct2, out)
            stream.synchronize()
            assert torch.equal(out, fused), "graph replay %d" % rep
        plan.close()
    assert torch.equal(plain, fused)
    raw = api.to_host(fused)
    check_sampled(raw, s, gt, random.Random(40))
    assert raw == api.to_host(msgs)


def test_pfe_config4_fused_and_plain_vs_oracle(api, gt):
    """weighted_pfe at B = 512, n = 128 Pareto(1.16), W = 4096, two batches: the fused path
    equals the plain path slot for slot; sampled plaintexts equal the oracle's gT^s."""
    import torch

    from weighted_bte_amd import workloads

    B, K = 512, 2
    weights = workloads.pareto_weights(128, 4096)
    wl = workloads.PfeWorkload(weights, 2048, B, seed=6)
    committee = workloads.heaviest_first(weights, wl.t)
    T = len(committee)
    ct1, ct2, msgs = wl.encrypt(K)
    s = list(wl.s)
    sh = []
    for b in range(K):
        sigma = api.pfe_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.g)
        sh.append(torch.cat([sigma[144 * j:144 * (j + 1)] for j in committee]))
    shares = torch.cat(sh)
    d1p, tl, ml = api.pfe_prepare(wl.material, wl.W, B, wl.offsets, committee, wl.alpha, wl.global_key)
    d1l = api.prepare_g2_lines(d1p)
    fc, pc = api.pfe_precompute_batched(ct1, tl, B)
    fused = api.pfe_open_prepared_batched(shares, T, d1l, ml, ct1, ct2, fc, pc, wl.alpha, B)
    plain = []
    for b in range(K):
        c1 = ct1[b * B * 96:(b + 1) * B * 96]
        f1, p1 = api.pfe_precompute(c1, tl)
        plain.append(api.pfe_open(shares[b * T * 144:(b + 1) * T * 144], d1p, ml, c1,
                                  ct2[b * B * 576:(b + 1) * B * 576], f1, p1, wl.alpha))
    assert torch.equal(torch.cat(plain), fused)
    raw = api.to_host(fused)
    check_sampled(raw, s, gt, random.Random(41))
    assert raw == api.to_host(msgs)


def test_multi_pairing_512_uneven_slots_vs_oracle(api, gt):
    """512 multi-pairings of log-uniform arity 2..512 (about 50k pairs) in one call, unprepared
    and prepared, with one batched final exponentiation: slot s equals gT^(sum a_i b_i) over its
    pairs (a_i G1, b_i G2) by the oracle on sampled slots, prepared == unprepared == the
    one-launch prepared multi-pairing everywhere."""
    import torch

    from weighted_bte_amd import inputs

    rnd = random.Random(11)
    ar = [int(round(math.exp(rnd.uniform(math.log(2), math.log(512))))) for _ in range(512)]
    assert min(ar) >= 2 and max(ar) <= 512
    off = [0]
    for a in ar:
        off.append(off[-1] + a)
    npairs = off[-1]
    st = inputs.Stream(11, 4100)
    a_s, b_s = st.scalars(npairs), st.scalars(npairs)
    P = api.g1_generator_mul(api.to_device(inputs.pack_scalars(a_s)))
    Q = api.g2_generator_mul(api.to_device(inputs.pack_scalars(b_s)))
    # the bases themselves against the oracle at a few places
    for i in (0, npairs // 3, npairs - 1):
        assert O.g1_affine_from_bytes(api.to_host(P[96 * i:96 * (i + 1)])) == \
            O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), a_s[i]))
        assert O.g2_affine_from_bytes(api.to_host(Q[192 * i:192 * (i + 1)])) == \
            O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), b_s[i]))
    lines = api.prepare_g2_lines(Q)
    un = api.full_final_exponentiation(api.unprepared_multi_miller_loop(P, Q, off))
    pr = api.full_final_exponentiation(api.multi_miller_loop_prepared(lines, P, off))
    assert torch.equal(un, pr)
    # the same 512 pairings with the Miller loops, slot products and final exponentiations in one
    # kernel launch
    assert torch.equal(api.multi_pairing_prepared_fused(lines, P, off), pr)
    exps = [sum(a * b for a, b in zip(a_s[off[k]:off[k + 1]], b_s[off[k]:off[k + 1]])) for k in range(512)]
    check_sampled(api.to_host(un), exps, gt, random.Random(42))


def test_serving_config5_batches_16_to_2048_large_committee(api, gt):
    """Serving with batch sizes spread over 16..2048 against two resident committees: one of
    about 1000 Pareto(1.16) parties with total weight near 2^14 (batches up to 64) and one with
    L = 2048 (batches up to 2048).  Each request's plaintexts equal the plain path alone (no
    padding) and, on sampled slots, the oracle's gT^s."""
    import torch

    from weighted_bte_amd import serving, workloads

    rnd = random.Random(5)
    big_n, big_w = 1000, 16000
    w_big = workloads.pareto_weights(big_n, big_w, seed=21)
    w_small = workloads.pareto_weights(24, 96, seed=22)
    specs = [(w_big, 64, 31), (w_small, 2048, 32)]
    wls, coms = [], []
    for weights, L, seed in specs:
        wl = workloads.BtxWorkload(weights, sum(weights) // 2, L, seed=seed)
        committee = workloads.heaviest_first(weights, wl.t)
        wls.append((wl, committee))
        coms.append(serving.Committee(wl.material, wl.W, wl.L, wl.offsets, committee))
    assert len(wls[0][1]) >= 40  # the big committee: 44 parties, about 8000 points for the Lagrange step
    sizes = [(0, 16), (0, 40), (0, 64), (1, 16), (1, 100), (1, 300), (1, 700), (1, 1500), (1, 2048), (0, 23)]
    reqs, exps = [], []
    for rid, (cid, B) in enumerate(sizes):
        wl, committee = wls[cid]
        ct1, ct2, _ = wl.encrypt(B)
        exps.append(list(wl.s))
        sig = api.btx_partial_decrypt(ct1, wl.q)
        shares = torch.cat([sig[144 * j:144 * (j + 1)] for j in committee])
        reqs.append(serving.Request(rid, cid, B, ct1, ct2, shares))
    srv = serving.BtxServer(coms, hbm_budget=6 << 30, max_fuse=16)
    for r in reqs:
        srv.submit(r)
    srv.drain()
    checked = 0
    for r, e in zip(reqs, exps):
        mixed = api.to_host(r.messages)
        assert api.to_host(srv.decrypt_alone(r)) == mixed, "request %d (alone != mixed)" % r.rid
        checked += check_sampled(mixed, e, gt, rnd, n=4, what="request %d slot" % r.rid)
    assert checked >= N_SAMPLE


def test_gt_fft_2p8_vs_oracle(api, gt):
    """GT FFT of size 2^8 over gT^(s_i): output k equals gT^(DFT(s)_k) by the oracle on sampled
    k; the scaled inverse returns the input exactly."""
    from weighted_bte_amd import inputs

    n = 256
    st = inputs.Stream(8, 4200)
    s = st.scalars(n)
    x = api.cyclotomic_mul_scalar_decomposed(api.to_device(O.f12_to_bytes(gt) * n), s)
    raw_x = api.to_host(x)
    for i in (0, 77, 255):
        assert f12_at(raw_x, i) == O.f12_pow(gt, s[i])
    d = api.Radix2Domain(n)
    y = d.fft_gt(x)
    check_sampled(api.to_host(y), O.fr_fft(s), gt, random.Random(43))
    assert api.to_host(d.ifft_gt(y)) == raw_x


def test_g1_msm_2p20_vs_oracle_closed_form(api):
    """G1 MSM at 2^20 over bases k_i G (10% sparse scalars): the affine result equals
    (sum k_i s_i mod r) G computed by the oracle; sampled bases equal the oracle's k_i G."""
    from weighted_bte_amd import inputs

    n = 1 << 20
    st = inputs.Stream(20, 4300)
    ks = st.scalars(n)
    sc = inputs.msm_scalars(st, n)
    P = api.g1_generator_mul(api.to_device(inputs.pack_scalars(ks)))
    for i in (0, 12345, n - 1):
        assert O.g1_affine_from_bytes(api.to_host(P[96 * i:96 * (i + 1)])) == \
            O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), ks[i]))
    out = api.g1_multi_exp_affine_bytes(P, api.to_device(inputs.pack_scalars(sc)))
    want = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), sum(k * (s & ((1 << 255) - 1)) for k, s in zip(ks, sc)) % O.R))
    assert O.G1.to_affine(O.g1_jac_from_bytes(api.to_host(out))) == want
