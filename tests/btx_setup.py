"""Seeded weighted_btx key material and ciphertexts for tests (restating the math of
weighted_btx/src/setup.rs:98-233 and encryption.rs:28-49; scalars on the host with the oracle,
group elements through the library's fixed-base kernels)."""
from helpers import O


def heaviest_first_committee(weights, threshold):
    """weighted_btx/examples/paper_reproduction.rs:581-607: heaviest parties until the weight
    exceeds the threshold; returned in ascending party order."""
    order = sorted(range(len(weights)), key=lambda j: (-weights[j], j))
    chosen, w = [], 0
    for j in order:
        chosen.append(j)
        w += weights[j]
        if w > threshold:
            break
    return sorted(chosen)


class BtxSetup:
    def __init__(self, api, weights, threshold, max_batch, seed=1):
        rng = O.SeededRng(seed, 77)
        self.api = api
        self.weights = list(weights)
        self.N = len(weights)
        self.W = sum(weights)
        self.t = threshold
        self.L = max_batch
        self.offsets = [0]
        for w in weights:
            self.offsets.append(self.offsets[-1] + w)
        self.owner = [j for j in range(self.N) for _ in range(weights[j])]
        dom = O.Radix2Domain(1 << (self.W - 1).bit_length() if self.W > 1 else 1)
        self.x = [pow(dom.generator, v, O.R) for v in range(self.W)]
        poly = [rng.scalar() for _ in range(threshold + 1)]
        self.z = poly[0]
        self.zv = [sum(c * pow(x, k, O.R) for k, c in enumerate(poly)) % O.R for x in self.x]
        self.q = [rng.scalar() or 1 for _ in range(self.N)]
        neg, pos = [], []
        for i in range(1, self.L + 1):
            for v in range(self.W):
                neg.append(self.zv[v] * pow(self.q[self.owner[v]], -i, O.R) % O.R)
        for i in range(1, self.L):
            for v in range(self.W):
                pos.append(self.zv[v] * pow(self.q[self.owner[v]], i, O.R) % O.R)
        scal = b"".join(O.scalar_to_le32(k) for k in neg + pos)
        self.material = api.g2_generator_mul(api.to_device(scal))
        self.rng = rng

    def encrypt_batch(self, B):
        """messages gT^s_i and ciphertexts (r_i G1, gT^(s_i + z r_i))."""
        api = self.api
        self.s = [self.rng.scalar() for _ in range(B)]
        self.r = [self.rng.scalar() for _ in range(B)]
        ct1 = api.g1_generator_mul(api.to_device(b"".join(O.scalar_to_le32(k) for k in self.r)))
        gt = api.to_device(O.f12_to_bytes(O.gt_generator()) * B)
        ct2 = api.cyclotomic_mul_scalar_decomposed(
            gt, [(s + self.z * r) % O.R for s, r in zip(self.s, self.r)])
        return ct1, ct2

    def kappa(self, selected, lam, d):
        return sum(l * self.zv[v] * pow(self.q[self.owner[v]], d, O.R)
                   for v, l in zip(selected, lam)) % O.R
