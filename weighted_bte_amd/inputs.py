"""Seeded synthetic inputs shared by the bench and the tests (vectorised with numpy).

The reference samples from OsRng; this build defines its own seeded streams: a ChaCha20
keystream (RFC 8439 block function, zero nonce, counter from 0) keyed by
SHA-256(b"weighted-bte-amd/<seed>/<stream>").  A scalar is 64 keystream bytes read little-endian
and reduced mod r.  Sparse scalars (the 10% tail of the MSM workload) take 1 + (u64 % 16)
distinct bit positions below 255, each position u64 % 255, redrawn on repeats.
"""
from __future__ import annotations

import hashlib

import numpy as np

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
_CONST = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574], dtype=np.uint32)


def _rotl(v, c):
    return (v << np.uint32(c)) | (v >> np.uint32(32 - c))


def chacha20_blocks(key32: bytes, first_counter: int, count: int) -> bytes:
    """`count` consecutive ChaCha20 blocks (64 bytes each) for a zero nonce."""
    k = np.frombuffer(key32, dtype="<u4").astype(np.uint32)
    st = np.zeros((count, 16), dtype=np.uint32)
    st[:, 0:4] = _CONST
    st[:, 4:12] = k
    st[:, 12] = (np.arange(count, dtype=np.uint64) + first_counter).astype(np.uint32)
    w = st.copy()
    x = [w[:, i].copy() for i in range(16)]

    def qr(a, b, c, d):
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 16)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 12)
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 8)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 7)

    with np.errstate(over="ignore"):
        for _ in range(10):
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
        out = np.stack(x, axis=1) + st
    return out.astype("<u4").tobytes()


class Stream:
    def __init__(self, seed: int, stream: int = 0):
        self.key = hashlib.sha256(b"weighted-bte-amd/%d/%d" % (seed, stream)).digest()
        self.counter = 0
        self.buf = b""
        self.pos = 0

    def bytes(self, n: int) -> bytes:
        if len(self.buf) - self.pos < n:
            rest = self.buf[self.pos:]
            need = n - len(rest)
            # refill ahead in bulk: the byte sequence is the same, small reads stay cheap
            blocks = max((need + 63) // 64, 1024)
            self.buf = rest + chacha20_blocks(self.key, self.counter, blocks)
            self.pos = 0
            self.counter += blocks
        out = self.buf[self.pos:self.pos + n]
        self.pos += n
        return out

    def scalars(self, n: int):
        raw = self.bytes(64 * n)
        return [int.from_bytes(raw[64 * i: 64 * i + 64], "little") % R for i in range(n)]

    def u64(self) -> int:
        return int.from_bytes(self.bytes(8), "little")

    def sparse_scalar(self, max_bits: int = 16) -> int:
        bits = set()
        k = 1 + self.u64() % max_bits
        while len(bits) < k:
            bits.add(self.u64() % 255)
        return sum(1 << b for b in bits) % R


def msm_scalars(stream: Stream, n: int, sparse_fraction: float = 0.1):
    """90% uniform scalars followed by a 10% tail of sparse ones (<= 16 set bits)."""
    n_sparse = int(n * sparse_fraction)
    return stream.scalars(n - n_sparse) + [stream.sparse_scalar() for _ in range(n_sparse)]


def pack_scalars(scalars) -> bytes:
    return b"".join(s.to_bytes(32, "little") for s in scalars)
