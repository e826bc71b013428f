"""The package's vectorised seeded input generator equals the oracle's scalar restatement."""
from helpers import O


def test_stream_matches_oracle():
    from weighted_bte_amd import inputs

    a = inputs.Stream(5, 7)
    b = O.SeededRng(5, 7)
    assert a.scalars(50) == [b.scalar() for _ in range(50)]
    assert inputs.msm_scalars(a, 40) == O.msm_config_scalars(b, 40)
    assert a.bytes(200) == b.bytes(200)


def test_chacha_block_vector():
    from weighted_bte_amd import inputs

    key = bytes(range(32))
    # RFC 8439 2.3.2 uses nonce 000000090000004a00000000; the package stream uses a zero nonce,
    # so compare the vectorised block function against the oracle's for several counters.
    for ctr in (0, 1, 77):
        assert inputs.chacha20_blocks(key, ctr, 1) == O.chacha20_block(key, ctr, bytes(12))
