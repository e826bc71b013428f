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
    from weighted_bte_amd import api, serving, workloads

    wls = []
    for seed, weights in ((3, [2, 1, 3, 1]), (4, [1, 1, 1, 2, 4])):
        wl = workloads.BtxWorkload(weights, threshold=sum(weights) // 2, max_batch=16, seed=seed)
        wls.append((wl, workloads.heaviest_first(weights, wl.t)))
    coms = [serving.Committee(wl.material, wl.W, wl.L, wl.offsets, c) for wl, c in wls]
    sizes = [(0, 3), (1, 5), (0, 8), (0, 2), (1, 16), (1, 7), (0, 3), (1, 1)]
    reqs, expected = make_requests(api, workloads, serving, wls, sizes)
    small = max(r.footprint() for r in reqs) * 3  # forces several admission rounds
    srv = serving.BtxServer(coms, hbm_budget=small, max_fuse=4)
    for r in reqs:
        srv.submit(r)
    rounds = 0
    while srv.waiting:
        srv.step()
        rounds += 1
    assert rounds >= 2
    for r, want in zip(reqs, expected):
        mixed = api.to_host(r.messages)
        assert mixed == want, "request %d (mixed)" % r.rid
        assert api.to_host(srv.decrypt_alone(r)) == mixed, "request %d (alone)" % r.rid
