"""The continuous-batching service: every request's plaintexts are bit-exact and identical
alone (plain per-call path, no padding) and in the fused, padded mix; the HBM budget splits the
admissions into several rounds."""
import pytest

pytestmark = pytest.mark.gpu


def make_requests(api, workloads, serving, wls, sizes):
    reqs, expected = [], []
    for rid, (cid, B) in enumerate(sizes):
        wl, committee = wls[cid]
        ct1, ct2, msgs = wl.encrypt(B)
        sig = api.btx_partial_decrypt(ct1, wl.q)
        import torch
        shares = torch.cat([sig[144 * j:144 * (j + 1)] for j in committee])
        reqs.append(serving.Request(rid, cid, B, ct1, ct2, shares))
        expected.append(api.to_host(msgs))
    return reqs, expected


def test_serving_alone_equals_mixed(dev):
    import torch

    from weighted_bte_amd import api, serving, workloads

    wls = []
    for seed, weights in ((3, [2, 1, 3, 1]), (4, [1, 1, 1, 2, 4])):
        wl = workloads.BtxWorkload(weights, threshold=sum(weights) // 2, max_batch=16, seed=seed)
        wls.append((wl, workloads.heaviest_first(weights, wl.t)))
    coms = [serving.Committee(wl.material, wl.W, wl.L, wl.offsets, c) for wl, c in wls]
    sizes = [(0, 3), (1, 5), (0, 8), (0, 2), (1, 16), (1, 7), (0, 3), (1, 1)]
    reqs, expected = make_requests(api, workloads, serving, wls, sizes)
    # an invalid slot: identity first component in the validated batch (so in the parties'
    # shares too) and it opens to one (the reference's None)
    r = reqs[3]
    wl, committee = wls[r.committee]
    raw = bytearray(api.to_host(r.ct1))
    raw[96:192] = bytes(96)
    r.ct1 = api.to_device(bytes(raw))
    sig = api.btx_partial_decrypt(r.ct1, wl.q)
    r.shares = torch.cat([sig[144 * j:144 * (j + 1)] for j in committee])
    r.valid = [True, False]
    small = max(r.footprint() for r in reqs) * 3  # forces several admissions
    srv = serving.BtxServer(coms, hbm_budget=small, max_fuse=4)
    for r in reqs:
        srv.submit(r)
    srv.drain()
    assert srv.launches >= 2
    from helpers import O
    one = O.f12_to_bytes(O.F12_ONE)
    for r, want in zip(reqs, expected):
        mixed = api.to_host(r.messages)
        if r.valid:
            want = b"".join(want[576 * i:576 * (i + 1)] if v else one for i, v in enumerate(r.valid))
        assert mixed == want, "request %d (mixed)" % r.rid
        assert api.to_host(srv.decrypt_alone(r)) == mixed, "request %d (alone)" % r.rid


def test_decrypt_requests_two_in_flight(dev):
    """With max_fuse = 1 every request is its own launch, at most two in flight on two streams;
    each result equals the plain path.  (That the two streams' kernels overlap on the device is
    measured by the bench: throughput at max_inflight 1 against 2, and a kernel trace.)"""
    from weighted_bte_amd import api, serving, workloads

    weights = [1, 2, 2, 1]
    wl = workloads.BtxWorkload(weights, threshold=3, max_batch=8, seed=9)
    com = serving.Committee(wl.material, wl.W, wl.L, wl.offsets, workloads.heaviest_first(weights, wl.t))
    reqs, expected = make_requests(api, workloads, serving, [(wl, com.accepted)], [(0, 5), (0, 8), (0, 2)])
    srv = serving.BtxServer([com], max_fuse=1, max_inflight=2)
    for r in reqs:
        srv.submit(r)
    srv.step()
    srv.step()
    assert len(srv.inflight) <= 2
    srv.drain()
    assert srv.launches == 3
    for r, want in zip(reqs, expected):
        assert api.to_host(r.messages) == want


def test_serving_accepts_any_batch_up_to_max(dev):
    """validate_batch_size (weighted_btx/src/decryption.rs:711-722) accepts any 1 <= B <= L: with
    L = 100 a request of B = 100 runs against the state prepared for L (same FFT size as 128);
    B = 101 is refused."""
    from weighted_bte_amd import api, serving, workloads

    weights = [1, 2, 1, 3]
    wl = workloads.BtxWorkload(weights, threshold=3, max_batch=100, seed=19)
    com = serving.Committee(wl.material, wl.W, wl.L, wl.offsets, workloads.heaviest_first(weights, wl.t))
    reqs, expected = make_requests(api, workloads, serving, [(wl, com.accepted)], [(0, 100), (0, 70)])
    srv = serving.BtxServer([com])
    for r in reqs:
        srv.submit(r)
    assert [r.padded for r in reqs] == [100, 100]
    srv.drain()
    for r, want in zip(reqs, expected):
        assert api.to_host(r.messages) == want
        assert api.to_host(srv.decrypt_alone(r)) == want
    with pytest.raises(ValueError):
        srv.submit(serving.Request(99, 0, 101, reqs[0].ct1, reqs[0].ct2, reqs[0].shares))


def test_serving_pfe_and_btx_requests_alone_equal_mixed(dev):
    """weighted_pfe requests (completed to the committee's B = 16 slots) and weighted_btx
    requests in one server: every request's plaintexts are the encrypted messages, alone (plain
    per-call path) and in the fused mix, which runs all pfe requests of a committee in one fused
    precompute + open."""
    import torch

    from weighted_bte_amd import api, serving, workloads

    B = 16
    pw = [2, 1, 3, 2]
    pfe = workloads.PfeWorkload(pw, threshold=4, batch_size=B, seed=23)
    pcom_ids = workloads.heaviest_first(pw, pfe.t)
    pcom = serving.PfeCommittee(pfe.material, pfe.W, B, pfe.offsets, pcom_ids, pfe.alpha, pfe.global_key)
    bw = [1, 2, 2, 1]
    btx = workloads.BtxWorkload(bw, threshold=3, max_batch=16, seed=24)
    bcom = serving.Committee(btx.material, btx.W, btx.L, btx.offsets, workloads.heaviest_first(bw, btx.t))
    reqs, expected = [], []
    for rid, b in enumerate((16, 5, 11, 1)):
        ct1, ct2, msgs = pfe.encrypt()
        c1, c2 = serving.pad_pfe_batch(ct1[:96 * b], ct2[:576 * b], B)
        sigma = api.pfe_partial_decrypt(c1, pfe.g)
        shares = torch.cat([sigma[144 * j:144 * (j + 1)] for j in pcom_ids])
        reqs.append(serving.Request(rid, 0, b, c1, c2, shares))
        expected.append(api.to_host(msgs)[:576 * b])
    breqs, bexp = make_requests(api, workloads, serving, [(btx, bcom.accepted)], [(0, 7), (0, 16)])
    for r in breqs:
        r.committee = 1
        r.rid += 10
    srv = serving.BtxServer([pcom, bcom], max_fuse=8)
    for r in reqs + breqs:
        srv.submit(r)
    srv.drain()
    for r, want in zip(reqs + breqs, expected + bexp):
        got = api.to_host(r.messages)
        assert got == want, "request %d (mixed)" % r.rid
        assert api.to_host(srv.decrypt_alone(r)) == want, "request %d (alone)" % r.rid
