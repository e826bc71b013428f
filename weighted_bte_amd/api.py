"""Python mirror of the reference's crate-private L1 interface, over the C ABI.

Names, argument meaning and error behaviour follow weighted_btx/src/{blst_utils,fft,
final_exponentiation}.rs; values are torch uint8 tensors in device memory holding blst layouts
(Montgomery Fp, 32-byte little-endian scalars).  Length mismatches raise ValueError where the
reference panics (assert_eq!).
"""
from __future__ import annotations

import ctypes

import torch

from . import call

G1_AFFINE = 96
G1_JAC = 144
G2_AFFINE = 192
G2_JAC = 288
FP12 = 576
SCALAR = 32


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t):
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if t.dtype != torch.uint8 and t.dtype != torch.int32:
        raise ValueError("expected uint8 or int32 tensor")
    return t.contiguous()


def to_device(data: bytes, device="cuda"):
    """Host bytes -> device uint8 tensor."""
    import numpy as np

    arr = np.frombuffer(bytes(data), dtype=np.uint8).copy()
    return torch.from_numpy(arr).to(device)


def to_host(t) -> bytes:
    return bytes(t.detach().cpu().numpy().tobytes())


def _empty(nbytes, like):
    return torch.empty(nbytes, dtype=torch.uint8, device=like.device)


# --- a1 / a2: MSM (weighted_btx/src/blst_utils.rs:43-152) -----------------------------------
def g1_multi_exp_affine_bytes(points, scalar_bytes):
    points, scalar_bytes = _dev(points), _dev(scalar_bytes)
    n = points.numel() // G1_AFFINE
    if scalar_bytes.numel() != n * SCALAR:
        raise ValueError("scalar_bytes.len() != points.len() * 32")
    out = _empty(G1_JAC, points)
    call("bte_g1_msm", _ptr(points), _ptr(scalar_bytes), n, _ptr(out), _stream())
    return out


def g2_multi_exp_affine_bytes(points, scalar_bytes):
    points, scalar_bytes = _dev(points), _dev(scalar_bytes)
    n = points.numel() // G2_AFFINE
    if scalar_bytes.numel() != n * SCALAR:
        raise ValueError("scalar_bytes.len() != points.len() * 32")
    out = _empty(G2_JAC, points)
    call("bte_g2_msm", _ptr(points), _ptr(scalar_bytes), n, _ptr(out), _stream())
    return out


def g2_multi_exp_affine_indexed(points, indices, scalar_bytes):
    points, indices, scalar_bytes = _dev(points), _dev(indices), _dev(scalar_bytes)
    n = indices.numel()
    if scalar_bytes.numel() != n * SCALAR:
        raise ValueError("scalar_bytes.len() != indices.len() * 32")
    out = _empty(G2_JAC, points)
    call("bte_g2_msm_indexed", _ptr(points), points.numel() // G2_AFFINE, _ptr(indices),
         _ptr(scalar_bytes), n, _ptr(out), _stream())
    return out


# --- a3: batch normalisation (weighted_btx/src/blst_utils.rs:11-36) ------------------------
def batch_normalize_g1(projective):
    projective = _dev(projective)
    n = projective.numel() // G1_JAC
    out = _empty(n * G1_AFFINE, projective)
    call("bte_g1_batch_normalize", _ptr(projective), n, _ptr(out), _stream())
    return out


def batch_normalize_g2(projective):
    projective = _dev(projective)
    n = projective.numel() // G2_JAC
    out = _empty(n * G2_AFFINE, projective)
    call("bte_g2_batch_normalize", _ptr(projective), n, _ptr(out), _stream())
    return out


# --- fixed-base generator multiples ----------------------------------------------------------
def g1_generator_mul(scalar_bytes):
    scalar_bytes = _dev(scalar_bytes)
    n = scalar_bytes.numel() // SCALAR
    out = _empty(n * G1_AFFINE, scalar_bytes)
    call("bte_g1_generator_mul", _ptr(scalar_bytes), n, _ptr(out), _stream())
    return out


def g2_generator_mul(scalar_bytes):
    scalar_bytes = _dev(scalar_bytes)
    n = scalar_bytes.numel() // SCALAR
    out = _empty(n * G2_AFFINE, scalar_bytes)
    call("bte_g2_generator_mul", _ptr(scalar_bytes), n, _ptr(out), _stream())
    return out


def set_msm_window(group, n, bits):
    call("bte_msm_set_window", group, n, bits)
