"""ctypes loader for the oracle's C restatement (oracle/csrc/bls_cpu.c) and the CPU-baseline
timing used by bench.py.  TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py cpu_baseline)."""
from __future__ import annotations

import ctypes
import os
import subprocess
import time

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libbls_cpu.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    if not os.path.exists(LIB):
        try:
            build()
        except Exception:
            return None
    lib = ctypes.CDLL(LIB)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.oracle_g1_msm.argtypes = [vp, vp, sz, ctypes.c_int, vp]
    lib.oracle_g2_msm.argtypes = [vp, vp, sz, ctypes.c_int, vp]
    lib.oracle_fp_mul.argtypes = [vp, vp, vp]
    lib.oracle_g1_consecutive_multiples.argtypes = [vp, sz, vp]
    lib.oracle_max_threads.restype = ctypes.c_int
    for name, args in (("oracle_miller_pairs", [vp, vp, sz, vp]), ("oracle_multi_miller", [vp, vp, sz, sz, vp]),
                       ("oracle_final_exp", [vp, sz, ctypes.c_int, ctypes.c_int, vp]),
                       ("oracle_cyc_pow_plans", [vp, vp, sz]), ("oracle_cyc_stage", [vp, sz, sz, vp]),
                       ("oracle_g1_stage", [vp, sz, sz, vp]), ("oracle_set_frobenius", [vp]),
                       ("oracle_prepare_lines", [vp, sz, vp]), ("oracle_miller_prepared", [vp, vp, sz, vp])):
        getattr(lib, name).argtypes = args
    _init_frobenius(lib)
    return lib


def _init_frobenius(lib):
    from oracle import bls12_381 as O

    words = []
    for k in (1, 2, 3):
        for i in range(6):
            e = i * (O.P**k - 1) // 6
            words += [(e >> (64 * j)) & (2**64 - 1) for j in range(24)]
    arr = (ctypes.c_uint64 * len(words))(*words)
    lib.oracle_set_frobenius(arr)


def plan_bytes(s):
    """fft.rs:33-50 CyclotomicScalarPlan packed as 4 x 65 int8 digits + count (+3 pad)."""
    from oracle import bls12_381 as O

    nafs, count = O.cyclotomic_scalar_plan(s)
    out = bytearray()
    for d in nafs:
        out += bytes((x & 0xFF) for x in d)
    out += bytes([count, 0, 0, 0])
    return bytes(out)


def _buf(b):
    return ctypes.create_string_buffer(bytes(b), len(b))


def g1_msm(lib, points: bytes, scalars: bytes, c: int = 0) -> bytes:
    n = len(points) // 96
    out = ctypes.create_string_buffer(144)
    lib.oracle_g1_msm(_buf(points), _buf(scalars), n, c, out)
    return out.raw


def g2_msm(lib, points: bytes, scalars: bytes, c: int = 0) -> bytes:
    n = len(points) // 192
    out = ctypes.create_string_buffer(288)
    lib.oracle_g2_msm(_buf(points), _buf(scalars), n, c, out)
    return out.raw


def g1_consecutive_multiples(lib, gen_affine_bytes: bytes, n: int) -> bytes:
    out = ctypes.create_string_buffer(96 * n)
    lib.oracle_g1_consecutive_multiples(_buf(gen_affine_bytes), n, out)
    return out.raw


def time_g1_msm_sample(lib, logn: int, seed: int = 2):
    """Times one G1 MSM of 2^logn points on all host cores (bases (i+1)G, the bench's scalar
    distribution).  Returns the cpu_baseline record in the metric's unit (ms per MSM)."""
    from oracle import bls12_381 as O

    n = 1 << logn
    gen = O.g1_affine_to_bytes(O.G1_GEN)
    pts = g1_consecutive_multiples(lib, gen, n)
    import numpy as np

    rng = np.random.default_rng(seed)
    raw = rng.integers(0, 2**32, size=(n, 8), dtype=np.uint64).astype("<u4")
    raw[:, 7] &= 0x3FFFFFFF  # below r
    n_sparse = n // 10
    sparse = np.zeros((n_sparse, 8), dtype="<u4")
    for i in range(n_sparse):
        for b in rng.choice(255, size=1 + int(rng.integers(16)), replace=False):
            sparse[i, b // 32] |= np.uint32(1 << (b % 32))
    raw[n - n_sparse:] = sparse
    sc = raw.tobytes()
    t0 = time.perf_counter()
    g1_msm(lib, pts, sc)
    dt = time.perf_counter() - t0
    return {
        "value": round(dt * 1e3, 2),
        "unit": "ms",
        "cores": int(lib.oracle_max_threads()),
        "kind": "port",
        "sample": "one full G1 MSM of 2^%d points (bases (i+1)G, 10%% sparse scalars), C "
                  "Pippenger restatement, OpenMP over windows" % logn,
    }


def time_btx_decrypt_sample(lib, B=512, T=4):
    """Times one weighted_btx precompute + open of B ciphertexts with a committee of T parties
    on all host cores (the reference's per-batch work: G1 FFT of m = nextpow2(2B), m prepared
    Miller loops, m easy parts, the cyclotomic prefix inverse FFT, B hard parts; B multi-Miller
    loops of arity T and B final exponentiations).  Point values are valid curve points but
    not a real key (timing only).  Returns the record in ctxts/s."""
    import random

    from oracle import bls12_381 as O

    m = 1
    while m < 2 * B:
        m <<= 1
    rnd = random.Random(5)
    g1 = g1_consecutive_multiples(lib, O.g1_affine_to_bytes(O.G1_GEN), max(m, B * T))
    g2_small = [O.g2_affine_to_bytes(O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), rnd.randrange(1, O.R))))
                for _ in range(16)]
    kernel_q = b"".join(g2_small[i % 16] for i in range(m))
    open_q = b"".join(g2_small[(7 * i) % 16] for i in range(B * T))
    one = O.fp_to_mont_bytes(1)
    c_jac = b"".join(g1[96 * i:96 * i + 96] + one for i in range(m))
    dom = O.Radix2Domain(m)
    tw_fwd = O.stage_twiddles(m, dom.generator)
    tw_inv = O.stage_twiddles(m, dom.generator_inverse)
    lines = ctypes.create_string_buffer(m * 68 * 288)
    lib.oracle_prepare_lines(_buf(kernel_q), m, lines)  # committee-dependent (prepare), untimed
    inv_plans = [b"".join(plan_bytes(t) for t in tw) for tw in tw_inv]
    t0 = time.perf_counter()
    # precompute_batch (decryption.rs:555-603)
    buf = _buf(c_jac)
    for st, tw in enumerate(tw_fwd):
        h = len(tw)
        words = b"".join(t.to_bytes(32, "little") for t in tw)
        lib.oracle_g1_stage(buf, m, h, _buf(words))
    # normalisation is folded into the Miller inputs below (affine points of the same count)
    ml = ctypes.create_string_buffer(576 * m)
    lib.oracle_miller_prepared(lines, _buf(g1[:96 * m]), m, ml)
    conv = ctypes.create_string_buffer(576 * m)
    lib.oracle_final_exp(ml, m, 1, 0, conv)
    for st, pl in enumerate(inv_plans[:-1]):
        lib.oracle_cyc_stage(conv, m, len(tw_inv[st]), _buf(pl))
    last = ctypes.create_string_buffer(576 * B)
    ctypes.memmove(last, ctypes.addressof(conv) + 576 * (m // 2), 576 * B)
    lib.oracle_cyc_pow_plans(last, _buf(inv_plans[-1][:264 * B]), B)
    beta = ctypes.create_string_buffer(576 * B)
    lib.oracle_final_exp(last, B, 0, 1, beta)
    # open_batch (decryption.rs:607-695)
    mm = ctypes.create_string_buffer(576 * B)
    lib.oracle_multi_miller(_buf(g1[:96 * B * T]), _buf(open_q), B, T, mm)
    alpha = ctypes.create_string_buffer(576 * B)
    lib.oracle_final_exp(mm, B, 1, 1, alpha)
    dt = time.perf_counter() - t0
    return {
        "value": round(B / dt, 1),
        "unit": "ctxts/s",
        "cores": int(lib.oracle_max_threads()),
        "kind": "port",
        "sample": "one weighted_btx precompute + open at B=%d, committee %d (G1 FFT m=%d, %d prepared "
                  "Miller loops, cyclotomic prefix IFFT, %d hard parts, %d x %d multi-Miller, %d final "
                  "exps), C restatement, OpenMP; %.2f s" % (B, T, m, m, B, B, T, B, dt),
    }
