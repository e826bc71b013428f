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
    lib.oracle_set_threads.argtypes = [ctypes.c_int]
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


def median_time(fn, reps=5):
    """One untimed warm-up, then the median wall time of `reps` runs (the reference's example
    protocol, weighted_pfe/examples/paper_reproduction.rs:399-423)."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def with_threads(lib, n, fn):
    lib.oracle_set_threads(n)
    try:
        return fn()
    finally:
        lib.oracle_set_threads(0)


def time_g1_msm_sample(lib, logn: int, seed: int = 2, reps: int = 5):
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
    dt = median_time(lambda: g1_msm(lib, pts, sc), reps)
    # one thread on a 2^(logn-4) prefix of the same inputs (median of 3 after a warm-up)
    small = 1 << max(1, logn - 4)
    dt1 = with_threads(lib, 1, lambda: median_time(lambda: g1_msm(lib, pts[:96 * small], sc[:32 * small]), 3))
    return {
        "value": round(dt * 1e3, 2),
        "unit": "ms",
        "cores": int(lib.oracle_max_threads()),
        "kind": "port",
        "reps": reps,
        "sample": "one full G1 MSM of 2^%d points (bases (i+1)G, 10%% sparse scalars), C "
                  "Pippenger restatement, OpenMP over windows; median of %d after a warm-up" % (logn, reps),
        "one_thread": {"n": small, "ms": round(dt1 * 1e3, 2), "reps": 3},
    }


def time_fft_sample(lib, group, logn, reps=3):
    """Median time (s) of one forward radix-2 transform of 2^logn G1 points (group 1, Jacobian
    butterflies with a scalar multiplication by the twiddle) or GT elements (group 3, cyclotomic
    butterflies: a decomposed power by the twiddle, a product and a conjugate product) on all
    host cores, after a warm-up."""
    from oracle import bls12_381 as O

    n = 1 << logn
    dom = O.Radix2Domain(n)
    tws = O.stage_twiddles(n, dom.generator)
    if group == 1:
        g1 = g1_consecutive_multiples(lib, O.g1_affine_to_bytes(O.G1_GEN), n)
        one = O.fp_to_mont_bytes(1)
        data = b"".join(g1[96 * i:96 * i + 96] + one for i in range(n))
        words = [_buf(b"".join(t.to_bytes(32, "little") for t in tw)) for tw in tws]

        def run():
            buf = _buf(data)
            for tw, w in zip(tws, words):
                lib.oracle_g1_stage(buf, n, len(tw), w)
    else:
        data = O.f12_to_bytes(O.F12_ONE) * n
        plans = [_buf(b"".join(plan_bytes(t) for t in tw)) for tw in tws]

        def run():
            buf = _buf(data)
            for tw, pl in zip(tws, plans):
                lib.oracle_cyc_stage(buf, n, len(tw), pl)
    return median_time(run, reps)


def _btx_decrypt_work(lib, B, T):
    """Inputs and the timed body of one weighted_btx precompute + open of B ciphertexts with a
    committee of T parties: G1 FFT of m = nextpow2(2B), m prepared Miller loops, m easy parts,
    the cyclotomic prefix inverse FFT, B hard parts; B multi-Miller loops of arity T and B final
    exponentiations.  Point values are valid curve points but not a real key (timing only)."""
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
    tw_fwd = [b"".join(t.to_bytes(32, "little") for t in tw) for tw in O.stage_twiddles(m, dom.generator)]
    tw_inv = O.stage_twiddles(m, dom.generator_inverse)
    lines = ctypes.create_string_buffer(m * 68 * 288)
    lib.oracle_prepare_lines(_buf(kernel_q), m, lines)  # committee-dependent (prepare), untimed
    inv_plans = [b"".join(plan_bytes(t) for t in tw) for tw in tw_inv]
    g1m, g1bt, oq = _buf(g1[:96 * m]), _buf(g1[:96 * B * T]), _buf(open_q)

    def run():
        # precompute_batch (decryption.rs:555-603)
        buf = _buf(c_jac)
        for words in tw_fwd:
            lib.oracle_g1_stage(buf, m, len(words) // 32, _buf(words))
        # normalisation is folded into the Miller inputs below (affine points of the same count)
        ml = ctypes.create_string_buffer(576 * m)
        lib.oracle_miller_prepared(lines, g1m, m, ml)
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
        lib.oracle_multi_miller(g1bt, oq, B, T, mm)
        alpha = ctypes.create_string_buffer(576 * B)
        lib.oracle_final_exp(mm, B, 1, 1, alpha)
    return run, m


def time_btx_decrypt_sample(lib, B=512, T=4, reps=5, B1=128):
    """Times one weighted_btx precompute + open of B ciphertexts (committee of T) on all host
    cores, median of `reps` after a warm-up, plus a one-thread row at B1 ciphertexts (median of
    3).  Returns the record in ctxts/s."""
    run, m = _btx_decrypt_work(lib, B, T)
    dt = median_time(run, reps)
    run1, _ = _btx_decrypt_work(lib, B1, T)
    dt1 = with_threads(lib, 1, lambda: median_time(run1, 3))
    return {
        "value": round(B / dt, 1),
        "unit": "ctxts/s",
        "cores": int(lib.oracle_max_threads()),
        "kind": "port",
        "reps": reps,
        "sample": "one weighted_btx precompute + open at B=%d, committee %d (G1 FFT m=%d, %d prepared "
                  "Miller loops, cyclotomic prefix IFFT, %d hard parts, %d x %d multi-Miller, %d final "
                  "exps), C restatement, OpenMP; median of %d after a warm-up, %.3f s"
                  % (B, T, m, m, B, B, T, B, reps, dt),
        "one_thread": {"B": B1, "ctxts_per_s": round(B1 / dt1, 1), "reps": 3},
    }


def _pfe_decrypt_work(lib, B, T):
    """Inputs and the timed body of one weighted_pfe precompute + open of B ciphertexts
    (weighted_pfe/src/decryption.rs:540-644): the G1 Cauchy transform (size-B G1 FFT and
    unscaled inverse FFT), B prepared Miller loops against the per-slot total lines, B easy parts,
    the cyclotomic Cauchy transform (size-B forward and inverse cyclotomic FFT); then B
    multi-Miller loops of arity T, B prepared mask Miller loops, 2B easy parts, B cyclotomic
    powers by alpha(alpha + 1) and B hard parts.  The signed-small multiplications by B - 1 - 2k
    (a few dozen group operations per slot) are left out.  Timing only."""
    import random

    from oracle import bls12_381 as O

    rnd = random.Random(6)
    g1 = g1_consecutive_multiples(lib, O.g1_affine_to_bytes(O.G1_GEN), max(B, B * T))
    g2_small = [O.g2_affine_to_bytes(O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), rnd.randrange(1, O.R))))
                for _ in range(16)]
    one = O.fp_to_mont_bytes(1)
    c_jac = b"".join(g1[96 * i:96 * i + 96] + one for i in range(B))
    dom = O.Radix2Domain(B)
    tw_fwd = [b"".join(t.to_bytes(32, "little") for t in tw) for tw in O.stage_twiddles(B, dom.generator)]
    tw_inv = [b"".join(t.to_bytes(32, "little") for t in tw) for tw in O.stage_twiddles(B, dom.generator_inverse)]
    cyc_fwd = [(len(tw), b"".join(plan_bytes(t) for t in tw)) for tw in O.stage_twiddles(B, dom.generator)]
    cyc_inv = [(len(tw), b"".join(plan_bytes(t) for t in tw)) for tw in O.stage_twiddles(B, dom.generator_inverse)]
    lines = ctypes.create_string_buffer(2 * B * 68 * 288)
    lib.oracle_prepare_lines(_buf(b"".join(g2_small[i % 16] for i in range(2 * B))), 2 * B, lines)  # prepare
    mask_lines = ctypes.addressof(lines) + B * 68 * 288
    open_q = _buf(b"".join(g2_small[(5 * i) % 16] for i in range(B * T)))
    aa1 = _buf(b"".join(plan_bytes(rnd.randrange(O.R)) for _ in range(B)))
    g1b, g1bt = _buf(g1[:96 * B]), _buf(g1[:96 * B * T])

    def run():
        buf = _buf(c_jac)  # G1 Cauchy: forward then unscaled inverse transform
        for words in tw_fwd + tw_inv:
            lib.oracle_g1_stage(buf, B, len(words) // 32, _buf(words))
        ml = ctypes.create_string_buffer(576 * 2 * B)
        lib.oracle_miller_prepared(lines, g1b, B, ml)
        cyc = ctypes.create_string_buffer(576 * B)
        lib.oracle_final_exp(ml, B, 1, 0, cyc)
        for h, pl in cyc_fwd + cyc_inv:  # cyclotomic Cauchy: forward then inverse transform
            lib.oracle_cyc_stage(cyc, B, h, _buf(pl))
        mm = ctypes.create_string_buffer(576 * 2 * B)  # opening: multi-Miller and mask loops
        lib.oracle_multi_miller(g1bt, open_q, B, T, mm)
        lib.oracle_miller_prepared(ctypes.c_void_p(mask_lines), g1b, B, ctypes.addressof(mm) + 576 * B)
        easy = ctypes.create_string_buffer(576 * 2 * B)
        lib.oracle_final_exp(mm, 2 * B, 1, 0, easy)
        lib.oracle_cyc_pow_plans(easy, aa1, B)
        out = ctypes.create_string_buffer(576 * B)
        lib.oracle_final_exp(easy, B, 0, 1, out)
    return run


def time_pfe_decrypt_sample(lib, B=512, T=4, reps=5, B1=128):
    """weighted_pfe precompute + open of B ciphertexts on all host cores (median of `reps` after
    a warm-up) and on one thread at B1 (median of 3); ctxts/s."""
    dt = median_time(_pfe_decrypt_work(lib, B, T), reps)
    run1 = _pfe_decrypt_work(lib, B1, T)
    dt1 = with_threads(lib, 1, lambda: median_time(run1, 3))
    return {
        "value": round(B / dt, 1),
        "unit": "ctxts/s",
        "cores": int(lib.oracle_max_threads()),
        "kind": "port",
        "reps": reps,
        "sample": "one weighted_pfe precompute + open at B=%d, committee %d (G1 and cyclotomic Cauchy "
                  "transforms of size %d, %d prepared Miller loops, %d x %d multi-Miller, %d mask loops, "
                  "%d easy parts, %d alpha powers, %d hard parts), C restatement, OpenMP; median of %d "
                  "after a warm-up, %.3f s" % (B, T, B, B, B, T, B, 2 * B, B, B, reps, dt),
        "one_thread": {"B": B1, "ctxts_per_s": round(B1 / dt1, 1), "reps": 3},
    }
