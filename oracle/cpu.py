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
    return lib


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
