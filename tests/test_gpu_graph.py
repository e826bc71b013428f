"""Graph capture straight through the C ABI: bte_btx_precompute_batched, then
bte_btx_open_prepared_batched, then bte_g1_msm captured into one HIP graph (torch.cuda.CUDAGraph,
global capture mode) and replayed several times with eager work on the same device between
replays; every replay gives the eager results bit for bit.  The per-call entry points allocate
no memory and wait for nothing inside the capture (their scratch and staging for captured work
are owned by it; the MSM's last window combination is a host callback node)."""
import ctypes

import pytest

from helpers import O

pytestmark = pytest.mark.gpu


def test_capture_precompute_open_and_msm(dev):
    import torch

    from weighted_bte_amd import api, call, inputs, workloads

    B, K = 16, 2
    weights = [2, 3, 1, 4]
    wl = workloads.BtxWorkload(weights, threshold=5, max_batch=B, seed=61)
    committee = workloads.heaviest_first(weights, wl.t)
    T = len(committee)
    ct1, ct2, msgs = wl.encrypt(B, K)
    valid = [True] * (B * K)
    valid[3] = False
    raw = bytearray(api.to_host(ct1))
    raw[3 * 96:4 * 96] = bytes(96)
    ct1 = api.to_device(bytes(raw))
    sh = []
    for b in range(K):
        sig = api.btx_partial_decrypt(ct1[b * B * 96:(b + 1) * B * 96], wl.q)
        sh.append(torch.cat([sig[144 * j:144 * (j + 1)] for j in committee]))
    shares = torch.cat(sh)
    lines, keys = api.btx_prepare(wl.material, wl.W, wl.L, wl.offsets, committee, B)
    key_lines = api.prepare_g2_lines(keys)
    n = 4096
    st = inputs.Stream(61, 4400)
    ks, sc = st.scalars(n), inputs.msm_scalars(st, n)
    P = api.g1_generator_mul(api.to_device(inputs.pack_scalars(ks)))
    S = api.to_device(inputs.pack_scalars(sc))

    # eager results (these calls also build the domain tables and grow the workspaces)
    beta_e = api.btx_precompute_batched(ct1, valid, lines, B)
    msg_e = api.btx_open_prepared_batched(shares, T, key_lines, ct2, beta_e, valid, B)
    msm_e = api.g1_multi_exp_affine_bytes(P, S)
    torch.cuda.synchronize()
    one = O.f12_to_bytes(O.F12_ONE)
    want_msgs = bytearray(api.to_host(msgs))
    want_msgs[3 * 576:4 * 576] = one
    assert api.to_host(msg_e) == bytes(want_msgs)
    want_msm = O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), sum(k * (s & ((1 << 255) - 1)) for k, s in zip(ks, sc)) % O.R))
    assert O.G1.to_affine(O.g1_jac_from_bytes(api.to_host(msm_e))) == want_msm

    beta_g = torch.empty_like(beta_e)
    msg_g = torch.empty_like(msg_e)
    msm_g = torch.empty_like(msm_e)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        vflags = (ctypes.c_uint8 * (B * K))(*[1 if v else 0 for v in valid])
        call("bte_btx_precompute_batched", ptr(ct1), vflags, B, K, ptr(lines), ptr(beta_g), s)
        call("bte_btx_open_prepared_batched", ptr(shares), T, ptr(key_lines), ptr(ct2), ptr(beta_g), vflags, B, K,
             ptr(msg_g), s)
        call("bte_g1_msm", ptr(P), ptr(S), n, ptr(msm_g), s)
        del vflags  # staged at capture time: the graph does not read this host array
    for rep in range(3):
        beta_g.zero_()
        msg_g.zero_()
        msm_g.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(beta_g, beta_e), "beta, replay %d" % rep
        assert torch.equal(msg_g, msg_e), "messages, replay %d" % rep
        # the MSM's Jacobian coordinates depend on the order of additions inside a bucket, which
        # the counting sort's LDS atomics leave open: compare the affine (canonical) values
        assert torch.equal(api.batch_normalize_g1(msm_g), api.batch_normalize_g1(msm_e)), "msm, replay %d" % rep
        # eager work between replays: the same entry points on the current stream
        other = api.g1_multi_exp_affine_bytes(P[:96 * 1000], S[:32 * 1000])
        api.btx_precompute_batched(ct1, valid, lines, B)
        torch.cuda.synchronize()
        assert other.numel() == 144
    del g
