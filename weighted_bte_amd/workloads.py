"""Synthetic weighted_btx workloads on the device (bench and tests).

Restates the algebra of the reference's trusted-dealer setup (weighted_btx/src/setup.rs:98-233)
and encryption (encryption.rs:28-49) with seeded scalars; group elements come from this
library's fixed-base kernels and GT values from its pairing, so nothing here depends on a CPU
implementation of the curve.
"""
from __future__ import annotations

import numpy as np

from . import api, inputs

R = inputs.R
ROOT_OF_UNITY = 0x16A2A19EDFE81F20D09B681922C813B4B63683508C2280B93829971F439F0D2B  # 2^32-th root


def pareto_weights(n, total, alpha=1.16, seed=4):
    """Pareto(alpha) weights rounded to positive integers, scaled to sum to `total` with a
    largest-remainder adjustment."""
    rng = np.random.default_rng(seed)
    raw = rng.pareto(alpha, n) + 1.0
    scaled = raw / raw.sum() * (total - n)
    base = np.floor(scaled).astype(np.int64)
    rem = int(total - n - base.sum())
    order = np.argsort(-(scaled - base), kind="stable")
    base[order[:rem]] += 1
    return [int(w) + 1 for w in base]


def heaviest_first(weights, threshold):
    """Heaviest parties until the weight exceeds the threshold (paper_reproduction.rs:581-607),
    returned in ascending party order."""
    order = sorted(range(len(weights)), key=lambda j: (-weights[j], j))
    out, w = [], 0
    for j in order:
        out.append(j)
        w += weights[j]
        if w > threshold:
            break
    return sorted(out)


def gt_generator_bytes():
    """e(G1, G2) through the library (Miller loop + final exponentiation)."""
    g1 = api.g1_generator_mul(api.to_device((1).to_bytes(32, "little")))
    g2 = api.g2_generator_mul(api.to_device((1).to_bytes(32, "little")))
    return api.to_host(api.full_final_exponentiation(api.miller_loop_pairs(g1, g2)))


class BtxWorkload:
    """Key material for `weights`, threshold and max batch L, with device-resident G2 material
    laid out as the C ABI expects (negative blocks then positive blocks)."""

    def __init__(self, weights, threshold, max_batch, seed=1):
        st = inputs.Stream(seed, 4000)
        self.weights = list(weights)
        self.N = len(weights)
        self.W = sum(weights)
        self.t = threshold
        self.L = max_batch
        self.offsets = [0]
        for w in weights:
            self.offsets.append(self.offsets[-1] + w)
        owner = np.repeat(np.arange(self.N), weights)
        dom = 1
        while dom < self.W:
            dom <<= 1
        g = pow(ROOT_OF_UNITY, (1 << 32) // dom, R)
        xs = [pow(g, v, R) for v in range(self.W)]
        poly = st.scalars(threshold + 1)
        self.z = poly[0]
        zv = []
        for x in xs:  # Horner
            acc = 0
            for c in reversed(poly):
                acc = (acc * x + c) % R
            zv.append(acc)
        self.q = [q or 1 for q in st.scalars(self.N)]
        qinv = [pow(q, -1, R) for q in self.q]
        neg = bytearray()
        pos = bytearray()
        qp = list(qinv)  # q^-i, starting at i = 1
        for i in range(1, self.L + 1):
            neg += b"".join((zv[v] * qp[owner[v]] % R).to_bytes(32, "little") for v in range(self.W))
            qp = [a * b % R for a, b in zip(qp, qinv)]
        qp = list(self.q)
        for i in range(1, self.L):
            pos += b"".join((zv[v] * qp[owner[v]] % R).to_bytes(32, "little") for v in range(self.W))
            qp = [a * b % R for a, b in zip(qp, self.q)]
        self.material = api.g2_generator_mul(api.to_device(bytes(neg + pos)))
        self.stream = st
        self.gt = gt_generator_bytes()

    def encrypt(self, B, count=1):
        """count batches of B ciphertexts: messages gT^s, ct1 = r G1, ct2 = gT^(s + z r)."""
        n = B * count
        s = self.stream.scalars(n)
        r = self.stream.scalars(n)
        ct1 = api.g1_generator_mul(api.to_device(inputs.pack_scalars(r)))
        gts = api.to_device(self.gt * n)
        ct2 = api.cyclotomic_mul_scalar_decomposed(gts, [(a + self.z * b) % R for a, b in zip(s, r)])
        msgs = api.cyclotomic_mul_scalar_decomposed(gts, s)
        return ct1, ct2, msgs
