"""GPU Schnorr proof checks of weighted_btx ciphertexts (weighted_btx/src/proof.rs:56-70,
SchnorrProof::verify): s_i G == A_i + e_i C_i for each ciphertext of a batch.  Inputs are
generator multiples with known discrete logs, so the expected verdict is the scalar identity
s_i == k_i + e_i r_i (mod r); the tampered cases mirror the reference test (a changed ciphertext,
a changed response) plus identity commitments and ciphertexts."""
import random

import pytest

from helpers import O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(dev):
    from weighted_bte_amd import api as A
    return A


def _pack(api, vals):
    return api.to_device(b"".join((v % O.R).to_bytes(32, "little") for v in vals))


@pytest.mark.parametrize("B", [1, 37, 512])
def test_btx_schnorr_batch(api, B):
    rnd = random.Random(70 + B)
    r = [rnd.randrange(1, O.R) for _ in range(B)]      # ciphertext randomness: C_i = r_i G
    k = [rnd.randrange(1, O.R) for _ in range(B)]      # nonces: A_i = k_i G
    e = [rnd.randrange(O.R) for _ in range(B)]         # hashed challenges
    s = [(ki + ei * ri) % O.R for ki, ei, ri in zip(k, e, r)]
    if B > 8:
        s[3] = (s[3] + 1) % O.R                       # wrong response
        r[5] = (r[5] + 1) % O.R                       # ciphertext changed after proving
        k[6] = 0                                      # identity commitment, response k + e r
        s[6] = (e[6] * r[6]) % O.R
        r[7] = 0                                      # identity ciphertext: s = k passes
        s[7] = k[7]
        e[8] = 0                                      # zero challenge: s = k passes
        s[8] = k[8]
    ok = api.btx_verify_proofs(api.g1_generator_mul(_pack(api, k)), api.g1_generator_mul(_pack(api, r)), s, e)
    want = [si == (ki + ei * ri) % O.R for si, ki, ei, ri in zip(s, k, e, r)]
    assert ok == want
    if B > 8:
        assert not ok[3] and not ok[5] and ok[6] and ok[7] and ok[8]


def test_btx_schnorr_rejects_noncanonical(api):
    from weighted_bte_amd import BteError
    one = _pack(api, [1])
    with pytest.raises(BteError):
        api.btx_verify_proofs(api.g1_generator_mul(one), api.g1_generator_mul(one), [O.R], [0])
