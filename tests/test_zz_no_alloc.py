"""The allocation-free multi-Miller entry point (bte_multi_miller_loop_noalloc, scratch in a
file-scope device array) against bte_multi_miller_loop through the C ABI: bit-identical Miller
values on two calls in a row, the second smaller (the array keeps the first call's contents), a
full pairing check on one segment, and the size limit refused before any device work."""
import ctypes
import random

import pytest

from helpers import O, g1_mul_gen, g2_mul_gen

pytestmark = pytest.mark.gpu

MAX_PAIRS = (32 << 20) // 576


def _run(name, P, Q, npairs, off, nseg):
    import torch

    from weighted_bte_amd import call

    out = torch.empty(576 * nseg, dtype=torch.uint8, device="cuda")
    call(name, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(Q.data_ptr()), npairs, ctypes.c_void_p(off.data_ptr()),
         nseg, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return out


def test_noalloc_multi_miller_matches_allocating_entry_point(dev):
    import torch

    from weighted_bte_amd import api, inputs

    st = inputs.Stream(71, 4500)
    for npairs, arities in ((3000, None), (700, [0, 1, 5, 3, 200, 491])):
        rnd = random.Random(npairs)
        if arities is None:
            arities = []
            left = npairs
            while left:
                a = min(left, rnd.randrange(1, 60))
                arities.append(a)
                left -= a
        off = [0]
        for a in arities:
            off.append(off[-1] + a)
        assert off[-1] == npairs
        a_s, b_s = st.scalars(npairs), st.scalars(npairs)
        P = api.g1_generator_mul(api.to_device(inputs.pack_scalars(a_s)))
        Q = api.g2_generator_mul(api.to_device(inputs.pack_scalars(b_s)))
        offs = torch.tensor(off, dtype=torch.int64, device="cuda")
        want = _run("bte_multi_miller_loop", P, Q, npairs, offs, len(arities))
        got = _run("bte_multi_miller_loop_noalloc", P, Q, npairs, offs, len(arities))
        torch.cuda.synchronize()
        assert torch.equal(got, want), "npairs %d" % npairs
        # one segment through the final exponentiation against the oracle's closed form
        k = max(range(len(arities)), key=lambda i: arities[i])
        e = api.full_final_exponentiation(got[576 * k:576 * (k + 1)])
        exp = sum(a * b for a, b in zip(a_s[off[k]:off[k + 1]], b_s[off[k]:off[k + 1]])) % O.R
        assert O.f12_from_bytes(api.to_host(e)) == O.f12_pow(O.gt_generator(), exp)


def test_noalloc_multi_miller_refuses_more_pairs_than_its_scratch(dev):
    import torch

    from weighted_bte_amd import BteError, call

    dummy = torch.zeros(1024, dtype=torch.uint8, device="cuda")
    p = ctypes.c_void_p(dummy.data_ptr())
    with pytest.raises(BteError):
        call("bte_multi_miller_loop_noalloc", p, p, MAX_PAIRS + 1, p, 1, p,
             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
