"""Synthetic weighted_btx workloads on the device (bench and tests).

The trusted-dealer setup (weighted_btx/src/setup.rs:98-233) runs on the device
(api.btx_keygen) from seeded dealer choices; encryption (encryption.rs:28-49) uses seeded
scalars, the library's fixed-base kernels and its pairing, so nothing here depends on a CPU
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

    def __init__(self, weights, threshold, max_batch, seed=1, gpu_keygen=True):
        st = inputs.Stream(seed, 4000)
        self.weights = list(weights)
        self.N = len(weights)
        self.W = sum(weights)
        self.t = threshold
        self.L = max_batch
        self.offsets = [0]
        for w in weights:
            self.offsets.append(self.offsets[-1] + w)
        poly = st.scalars(threshold + 1)
        self.z = poly[0]
        self.q = [q or 1 for q in st.scalars(self.N)]
        self.stream = st
        self.gt = gt_generator_bytes()
        if gpu_keygen:
            self.material = api.btx_keygen(poly, self.q, self.offsets, self.L)
        else:
            self.material = self.host_material(poly)

    def host_material(self, poly):
        """The same material with the scalar tables computed on the host (the check of the
        device key generation)."""
        owner = np.repeat(np.arange(self.N), self.weights)
        dom = 1
        while dom < self.W:
            dom <<= 1
        g = pow(ROOT_OF_UNITY, (1 << 32) // dom, R)
        xs = [pow(g, v, R) for v in range(self.W)]
        zv = []
        for x in xs:  # Horner
            acc = 0
            for c in reversed(poly):
                acc = (acc * x + c) % R
            zv.append(acc)
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
        return api.g2_generator_mul(api.to_device(bytes(neg + pos)))

    def encrypt(self, B, count=1):
        """count batches of B ciphertexts: messages gT^s, ct1 = r G1, ct2 = gT^(s + z r)."""
        n = B * count
        s = self.stream.scalars(n)
        r = self.stream.scalars(n)
        self.s, self.r = s, r  # the messages' exponents, for checks against gT^s
        ct1 = api.g1_generator_mul(api.to_device(inputs.pack_scalars(r)))
        gts = api.to_device(self.gt * n)
        ct2 = api.cyclotomic_mul_scalar_decomposed(gts, [(a + self.z * b) % R for a, b in zip(s, r)])
        msgs = api.cyclotomic_mul_scalar_decomposed(gts, s)
        return ct1, ct2, msgs


class PfeWorkload:
    """weighted_pfe key material (weighted_pfe/src/setup.rs:149-354) for `weights`, threshold and
    power-of-two batch size B: device G2 array d1 (B blocks of W) then d2 (B blocks of W), the
    global key V, alpha points and party fractions g_{j,i} = 1/(rho_j + alpha_i)."""

    def __init__(self, weights, threshold, batch_size, seed=1, gpu_keygen=True):
        st = inputs.Stream(seed, 5000)
        B = batch_size
        self.weights = list(weights)
        self.N = len(weights)
        self.W = sum(weights)
        self.t = threshold
        self.B = B
        self.offsets = [0]
        for w in weights:
            self.offsets.append(self.offsets[-1] + w)
        owner = [j for j in range(self.N) for _ in range(weights[j])]
        dom = 1
        while dom < self.W:
            dom <<= 1
        self._omega = pow(ROOT_OF_UNITY, (1 << 32) // dom, R)
        poly = st.scalars(threshold + 1)
        poly[0] = poly[0] or 1
        self.z = poly[0]
        self._zv = None
        # alpha_i = omega_B^i + c with the smallest c >= 2 (setup.rs:605-618)
        gb = pow(ROOT_OF_UNITY, (1 << 32) // B, R)
        roots = [pow(gb, i, R) for i in range(B)]
        c = 2
        while any((r + c) % R == 0 or (r + c + 1) % R == 0 for r in roots):
            c += 1
        self.alpha = [(r + c) % R for r in roots]
        bm1, b0 = st.scalars(2)
        self.beta_m1, self.beta_0 = bm1 or 1, b0 or 1
        self.rho = []
        labels = set()
        while len(self.rho) < self.N:
            rho = st.scalars(1)[0]
            lab = [(rho + a) % R for a in self.alpha]
            if any(x == 0 or x in labels for x in lab):
                continue
            labels.update(lab)
            self.rho.append(rho)
        self.poly = poly
        if gpu_keygen:
            self.material, self.global_key, self.g, _ = api.pfe_keygen(
                poly, self.rho, self.alpha, (self.beta_m1, self.beta_0), self.offsets, B)
        else:
            self.material, self.global_key, self.g = self.host_material(self.zv, owner)
        self.owner = owner
        self.stream = st
        self.gt = gt_generator_bytes()

    @property
    def zv(self):
        """Share values z_v = poly(omega^v) at the W virtual points (host, computed on first use:
        the device key generation evaluates them itself)."""
        if self._zv is None:
            zv = []
            for v in range(self.W):
                x = pow(self._omega, v, R)
                acc = 0
                for c in reversed(self.poly):
                    acc = (acc * x + c) % R
                zv.append(acc)
            self._zv = zv
        return self._zv

    def host_material(self, zv, owner):
        """The key material with the scalar tables computed on the host (the check of the device
        key generation): (d1 || d2, global key, fractions g[j][i])."""
        B = self.B
        g = [[pow((rho + a) % R, -1, R) for a in self.alpha] for rho in self.rho]
        ainv = [pow(a, -1, R) for a in self.alpha]
        a1inv = [pow(a + 1, -1, R) for a in self.alpha]
        u = [[(gj[i] * gj[i] + self.beta_m1 * a1inv[i] + self.beta_0 * ainv[i]) % R for i in range(B)]
             for gj in g]
        d1 = b"".join((zv[v] * g[owner[v]][i] % R).to_bytes(32, "little")
                      for i in range(B) for v in range(self.W))
        d2 = b"".join((zv[v] * u[owner[v]][i] % R).to_bytes(32, "little")
                      for i in range(B) for v in range(self.W))
        material = api.g2_generator_mul(api.to_device(d1 + d2))
        vkey = self.z * (self.beta_m1 + self.beta_0) % R
        global_key = api.g2_generator_mul(api.to_device(vkey.to_bytes(32, "little")))
        return material, global_key, g

    def encrypt(self, count=1):
        """count batches of B messages gT^s with ct1 = r G1 and ct2 = gT^(s + z beta_0 r)
        (encryption key gT^(z beta_0))."""
        B = self.B * count
        self.s = self.stream.scalars(B)
        self.r = self.stream.scalars(B)
        ct1 = api.g1_generator_mul(api.to_device(inputs.pack_scalars(self.r)))
        gts = api.to_device(self.gt * B)
        e = self.z * self.beta_0 % R
        ct2 = api.cyclotomic_mul_scalar_decomposed(gts, [(a + e * b) % R for a, b in zip(self.s, self.r)])
        msgs = api.cyclotomic_mul_scalar_decomposed(gts, self.s)
        return ct1, ct2, msgs
