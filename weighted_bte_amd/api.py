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
    """Checks a device input and returns it unchanged.  Non-contiguous tensors are refused rather
    than copied: a temporary copy would be released (and its memory handed out again by the
    caching allocator) while the asynchronous kernels still read it."""
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if t.dtype != torch.uint8 and t.dtype != torch.int32:
        raise ValueError("expected uint8 or int32 tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t


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


# --- canonical compressed encodings (blstrs to_compressed / from_compressed) -----------------
def g1_compress(points_affine):
    p = _dev(points_affine)
    n = p.numel() // G1_AFFINE
    out = _empty(48 * n, p)
    call("bte_g1_compress", _ptr(p), n, _ptr(out), _stream())
    return out


def g2_compress(points_affine):
    p = _dev(points_affine)
    n = p.numel() // G2_AFFINE
    out = _empty(96 * n, p)
    call("bte_g2_compress", _ptr(p), n, _ptr(out), _stream())
    return out


def g1_decompress(encodings):
    """-> (affine points, validity flags as a uint8 tensor)."""
    e = _dev(encodings)
    n = e.numel() // 48
    pts, ok = _empty(n * G1_AFFINE, e), _empty(n, e)
    call("bte_g1_decompress", _ptr(e), n, _ptr(pts), _ptr(ok), _stream())
    return pts, ok


def g2_decompress(encodings):
    e = _dev(encodings)
    n = e.numel() // 96
    pts, ok = _empty(n * G2_AFFINE, e), _empty(n, e)
    call("bte_g2_decompress", _ptr(e), n, _ptr(pts), _ptr(ok), _stream())
    return pts, ok


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


def _msm_batched(fn, width, jwidth, points, scalar_bytes, seg_offsets, indices):
    points, scalar_bytes = _dev(points), _dev(scalar_bytes)
    nterms = scalar_bytes.numel() // SCALAR
    if indices is not None:
        indices = _dev(indices)
        if indices.numel() != nterms:
            raise ValueError("indices.len() != scalars.len()")
    elif points.numel() != nterms * width:
        raise ValueError("points.len() != scalars.len()")
    if seg_offsets[0] != 0 or seg_offsets[-1] != nterms or any(
            b < a for a, b in zip(seg_offsets, seg_offsets[1:])):
        raise ValueError("bad segment offsets")
    offs = torch.tensor(seg_offsets, dtype=torch.int64, device=points.device)
    nseg = len(seg_offsets) - 1
    out = _empty(nseg * jwidth, points)
    status = torch.zeros(1, dtype=torch.int32, device=points.device)
    call(fn, _ptr(points), points.numel() // width, _ptr(indices), _ptr(scalar_bytes), nterms,
         _ptr(offs), nseg, _ptr(out), _ptr(status), _stream())
    if indices is not None and int(status.item()) != 0:  # waits for the stream
        raise ValueError("point index out of range")
    return out


def g1_multi_exp_batched(points, scalar_bytes, seg_offsets, indices=None):
    """One G1 MSM per segment of terms (rayon fan-out of g1_multi_exp_affine_bytes)."""
    return _msm_batched("bte_g1_msm_batched", G1_AFFINE, G1_JAC, points, scalar_bytes, seg_offsets,
                        indices)


def g2_multi_exp_batched(points, scalar_bytes, seg_offsets, indices=None):
    """One G2 MSM per segment of terms (rayon fan-out of g2_multi_exp_affine_bytes/_indexed)."""
    return _msm_batched("bte_g2_msm_batched", G2_AFFINE, G2_JAC, points, scalar_bytes, seg_offsets,
                        indices)


def set_msm_window(group, n, bits):
    call("bte_msm_set_window", group, n, bits)


# Launch-shape knobs of bte_tune_set (include/bte_hip.h); results never depend on them.
TUNE_MSM_ACC_ROUNDS, TUNE_CYC_THREE_LANE_MAX, TUNE_MM_CHUNK_MAX, TUNE_MM_MIN_THREADS = 1, 2, 3, 4
TUNE_DEFAULTS = {TUNE_MSM_ACC_ROUNDS: 3, TUNE_CYC_THREE_LANE_MAX: 32768, TUNE_MM_CHUNK_MAX: 8,
                 TUNE_MM_MIN_THREADS: 65536}


def tune_set(knob, value):
    call("bte_tune_set", knob, int(value))


def tune_get(knob):
    v = ctypes.c_int64()
    call("bte_tune_get", knob, ctypes.byref(v))
    return v.value


# --- a4..a9: Radix2Domain transforms (weighted_btx/src/fft.rs, weighted_pfe/src/fft.rs) -----
class Radix2Domain:
    """Radix2Domain::new (fft.rs:164-191): the size must be a nonzero power of two <= 2^32."""

    def __init__(self, size):
        if size <= 0 or size & (size - 1) or size > (1 << 32):
            raise ValueError("InvalidBatchSize(%d)" % size)
        self.size = size

    def _check(self, values, width):
        values = _dev(values)
        if values.numel() != self.size * width:
            raise ValueError("values.len() != domain size")
        return values

    def fft_g1(self, values):
        v = self._check(values, G1_JAC).clone()
        call("bte_fft_g1", _ptr(v), self.size, 0, 0, _stream())
        return v

    def ifft_g1(self, values, scaled=True):
        v = self._check(values, G1_JAC).clone()
        call("bte_fft_g1", _ptr(v), self.size, 1, 1 if scaled else 0, _stream())
        return v

    def fft_g2(self, values):
        v = self._check(values, G2_JAC).clone()
        call("bte_fft_g2", _ptr(v), self.size, 0, 0, _stream())
        return v

    def ifft_g2(self, values, scaled=True):
        v = self._check(values, G2_JAC).clone()
        call("bte_fft_g2", _ptr(v), self.size, 1, 1 if scaled else 0, _stream())
        return v

    def fft_gt(self, values):
        v = self._check(values, FP12).clone()
        call("bte_fft_gt", _ptr(v), self.size, 0, 0, _stream())
        return v

    def ifft_gt(self, values, scaled=True):
        v = self._check(values, FP12).clone()
        call("bte_fft_gt", _ptr(v), self.size, 1, 1 if scaled else 0, _stream())
        return v

    def fft_cyclotomic(self, values):
        v = self._check(values, FP12).clone()
        call("bte_fft_cyclotomic", _ptr(v), self.size, 0, _stream())
        return v

    def ifft_cyclotomic_unscaled(self, values):
        v = self._check(values, FP12).clone()
        call("bte_fft_cyclotomic", _ptr(v), self.size, 1, _stream())
        return v

    def ifft_cyclotomic_prefix_unscaled(self, values, output_count):
        if output_count > self.size // 2:
            raise ValueError("output_count > size / 2")
        v = self._check(values, FP12).clone()
        call("bte_ifft_cyclotomic_prefix_unscaled", _ptr(v), self.size, output_count, _stream())
        return v[: output_count * FP12]


def fft_g1_segmented(values, log_sizes, inverse=False, scaled=True):
    """Many G1 transforms of uneven power-of-two sizes (Jacobian, back to back) in one launch
    per stage level."""
    v = _dev(values).clone()
    if v.numel() != sum(1 << l for l in log_sizes) * G1_JAC:
        raise ValueError("values.len() != sum of sizes")
    arr = (ctypes.c_int * len(log_sizes))(*log_sizes)
    call("bte_fft_g1_segmented", _ptr(v), arr, len(log_sizes), 1 if inverse else 0, 1 if scaled else 0,
         _stream())
    return v


def fft_fr(values, inverse=False, scaled=True, count=1):
    """Radix2Domain::{fft, ifft} over Scalar (fft.rs:197-234) on canonical 32-byte scalars
    (device), `count` transforms of equal power-of-two size back to back."""
    v = _dev(values).clone()
    n = v.numel() // SCALAR // count
    if n * count * SCALAR != v.numel():
        raise ValueError("values.len() != count * n")
    call("bte_fft_fr_batched", _ptr(v), n, count, 1 if inverse else 0, 1 if scaled else 0, _stream())
    return v


def cyclotomic_mul_scalar_decomposed(values, scalars):
    """fft.rs:65-71, element-wise; scalars are canonical ints (host)."""
    v = _dev(values).clone()
    n = v.numel() // FP12
    if len(scalars) != n:
        raise ValueError("scalars.len() != values.len()")
    hs = b"".join(int(s).to_bytes(32, "little") for s in scalars)
    buf = ctypes.create_string_buffer(hs, len(hs))
    call("bte_cyclotomic_pow_scalars", _ptr(v), buf, n, _stream())
    torch.cuda.current_stream().synchronize()
    return v


def mul_cyclotomic_signed_small(values, coefficients):
    """weighted_pfe/src/fft.rs:671-697, element-wise."""
    v = _dev(values).clone()
    n = v.numel() // FP12
    c = torch.tensor(list(coefficients), dtype=torch.int64, device=v.device)
    if c.numel() != n:
        raise ValueError("coefficients.len() != values.len()")
    call("bte_cyclotomic_pow_small", _ptr(v), _ptr(c), n, _stream())
    return v


# --- a10..a14: pairing (weighted_btx/src/final_exponentiation.rs) -----------------------------
MILLER_LINE_COUNT = 68
LINES_BYTES = MILLER_LINE_COUNT * 288


def prepare_g2_lines(q_affine):
    """PreparedG2Lines::from_affine for every point (final_exponentiation.rs:186-195)."""
    q = _dev(q_affine)
    n = q.numel() // G2_AFFINE
    out = _empty(n * LINES_BYTES, q)
    call("bte_g2_prepare_lines", _ptr(q), n, _ptr(out), _stream())
    return out


def multi_miller_loop_prepared(lines, p_affine, seg_offsets):
    """unprepared_multi_miller_loop with the G2 side prepared: pair g uses table g of `lines`
    (prepare_g2_lines of the G2 points) and point g; one Fp12 per segment of seg_offsets."""
    lines, p = _dev(lines), _dev(p_affine)
    n = p.numel() // G1_AFFINE
    if lines.numel() != n * LINES_BYTES or seg_offsets[0] != 0 or seg_offsets[-1] != n:
        raise ValueError("lines / offsets do not match the points")
    tabs = torch.arange(n, dtype=torch.int64, device=p.device) * LINES_BYTES + lines.data_ptr()
    nseg = len(seg_offsets) - 1
    offs = (ctypes.c_uint64 * len(seg_offsets))(*seg_offsets)
    out = _empty(nseg * FP12, p)
    call("bte_multi_miller_loop_prepared", _ptr(tabs), _ptr(p), n, offs, nseg, _ptr(out), _stream())
    return out


def multi_pairing_prepared_fused(lines, p_affine, seg_offsets):
    """Per segment, the final exponentiation of multi_miller_loop_prepared (the pairing product
    final_exponentiation.rs:229-273 then :348-354), in one kernel launch."""
    lines, p = _dev(lines), _dev(p_affine)
    n = p.numel() // G1_AFFINE
    if lines.numel() != n * LINES_BYTES or seg_offsets[0] != 0 or seg_offsets[-1] != n:
        raise ValueError("lines / offsets do not match the points")
    tabs = torch.arange(n, dtype=torch.int64, device=p.device) * LINES_BYTES + lines.data_ptr()
    nseg = len(seg_offsets) - 1
    offs = (ctypes.c_uint64 * len(seg_offsets))(*seg_offsets)
    out = _empty(nseg * FP12, p)
    call("bte_multi_pairing_prepared_fused", _ptr(tabs), _ptr(p), n, offs, nseg, _ptr(out), _stream())
    return out


def miller_loop_prepared(lines, p_affine):
    lines, p = _dev(lines), _dev(p_affine)
    n = p.numel() // G1_AFFINE
    if lines.numel() != n * LINES_BYTES:
        raise ValueError("lines.len() != points.len()")
    out = _empty(n * FP12, p)
    call("bte_miller_loop_prepared", _ptr(lines), _ptr(p), n, _ptr(out), _stream())
    return out


def unprepared_multi_miller_loop(left, right, seg_offsets=None):
    """final_exponentiation.rs:229-273; with seg_offsets (a list, len nseg+1) it evaluates one
    multi-Miller loop per segment of pairs (uneven arities)."""
    left, right = _dev(left), _dev(right)
    npairs = left.numel() // G1_AFFINE
    if right.numel() != npairs * G2_AFFINE:
        raise ValueError("left.len() != right.len()")
    if seg_offsets is None:
        seg_offsets = [0, npairs]
    if seg_offsets[0] != 0 or seg_offsets[-1] != npairs or any(
            b < a for a, b in zip(seg_offsets, seg_offsets[1:])):
        raise ValueError("bad segment offsets")
    offs = torch.tensor(seg_offsets, dtype=torch.int64, device=left.device)
    nseg = len(seg_offsets) - 1
    out = _empty(nseg * FP12, left)
    call("bte_multi_miller_loop", _ptr(left), _ptr(right), npairs, _ptr(offs), nseg, _ptr(out),
         _stream())
    return out


def miller_loop_pairs(left, right):
    left, right = _dev(left), _dev(right)
    n = left.numel() // G1_AFFINE
    out = _empty(n * FP12, left)
    call("bte_miller_loop_pairs", _ptr(left), _ptr(right), n, _ptr(out), _stream())
    return out


def easy_final_exponentiation(values):
    v = _dev(values)
    out = torch.empty_like(v)
    call("bte_final_exp_easy", _ptr(v), v.numel() // FP12, _ptr(out), _stream())
    return out


batch_easy_final_exponentiation = easy_final_exponentiation


def hard_final_exponentiation(values):
    v = _dev(values)
    out = torch.empty_like(v)
    call("bte_final_exp_hard", _ptr(v), v.numel() // FP12, _ptr(out), _stream())
    return out


def full_final_exponentiation(values):
    v = _dev(values)
    out = torch.empty_like(v)
    call("bte_final_exp", _ptr(v), v.numel() // FP12, _ptr(out), _stream())
    return out


# --- a16: weighted_btx phases (weighted_btx/src/decryption.rs) ------------------------------
def lagrange_at_zero(total_weight, selected):
    """interpolation.rs:31-52 (direct formula up to 64 points, device product tree above);
    returns canonical ints."""
    sel = (ctypes.c_uint64 * len(selected))(*selected)
    out = ctypes.create_string_buffer(32 * len(selected))
    call("bte_lagrange_at_zero", total_weight, sel, len(selected), out)
    raw = out.raw
    return [int.from_bytes(raw[32 * i: 32 * i + 32], "little") for i in range(len(selected))]


def btx_prepare(material, total_weight, max_batch, party_offsets, accepted, batch_size):
    """prepare_decryption (decryption.rs:427-551) -> (kernel_lines, opening_keys)."""
    material = _dev(material)
    m = 1
    while m < 2 * batch_size:
        m <<= 1
    offs = (ctypes.c_uint64 * len(party_offsets))(*party_offsets)
    acc = (ctypes.c_uint32 * len(accepted))(*accepted)
    lines = _empty(m * LINES_BYTES, material)
    keys = _empty(batch_size * len(accepted) * G2_AFFINE, material)
    call("bte_btx_prepare", _ptr(material), total_weight, max_batch, offs, len(party_offsets) - 1, acc,
         len(accepted), batch_size, _ptr(lines), _ptr(keys), _stream())
    return lines, keys


def btx_precompute(ct1_affine, valid, kernel_lines):
    """precompute_batch (decryption.rs:555-603) -> beta (B x Fp12)."""
    ct1 = _dev(ct1_affine)
    B = ct1.numel() // G1_AFFINE
    v = (ctypes.c_uint8 * B)(*[1 if x else 0 for x in valid])
    beta = _empty(B * FP12, ct1)
    call("bte_btx_precompute", _ptr(ct1), v, B, _ptr(_dev(kernel_lines)), _ptr(beta), _stream())
    return beta


def btx_open(shares_jac, opening_keys, ct2, beta, valid):
    """open_batch (decryption.rs:607-695) -> messages (B x Fp12; one where invalid)."""
    shares = _dev(shares_jac)
    T = shares.numel() // G1_JAC
    ct2 = _dev(ct2)
    B = ct2.numel() // FP12
    v = (ctypes.c_uint8 * B)(*[1 if x else 0 for x in valid])
    out = _empty(B * FP12, ct2)
    call("bte_btx_open", _ptr(shares), T, _ptr(_dev(opening_keys)), _ptr(ct2), _ptr(_dev(beta)), v, B,
         _ptr(out), _stream())
    return out


def btx_precompute_batched(ct1_affine, valid, kernel_lines, batch_size):
    """precompute_batch for count = len(valid) / B batches fused in each launch."""
    ct1 = _dev(ct1_affine)
    n = ct1.numel() // G1_AFFINE
    count = n // batch_size
    v = (ctypes.c_uint8 * n)(*[1 if x else 0 for x in valid])
    beta = _empty(n * FP12, ct1)
    call("bte_btx_precompute_batched", _ptr(ct1), v, batch_size, count, _ptr(_dev(kernel_lines)),
         _ptr(beta), _stream())
    return beta


def btx_open_batched(shares_jac, n_accepted, opening_keys, ct2, beta, valid, batch_size):
    ct2 = _dev(ct2)
    n = ct2.numel() // FP12
    count = n // batch_size
    v = (ctypes.c_uint8 * n)(*[1 if x else 0 for x in valid])
    out = _empty(n * FP12, ct2)
    call("bte_btx_open_batched", _ptr(_dev(shares_jac)), n_accepted, _ptr(_dev(opening_keys)), _ptr(ct2),
         _ptr(_dev(beta)), v, batch_size, count, _ptr(out), _stream())
    return out


class _BtxRequest(ctypes.Structure):
    """Mirror of bte_btx_request (include/bte_hip.h)."""
    _fields_ = [("ct1", ctypes.c_void_p), ("ct2", ctypes.c_void_p), ("valid", ctypes.c_void_p),
                ("count", ctypes.c_size_t), ("batch_size", ctypes.c_size_t),
                ("kernel_lines", ctypes.c_void_p), ("opening_keys", ctypes.c_void_p),
                ("opening_lines", ctypes.c_void_p), ("shares", ctypes.c_void_p), ("n_accepted", ctypes.c_size_t),
                ("messages_out", ctypes.c_void_p)]


def btx_decrypt_requests(requests):
    """precompute_batch + open_batch (decryption.rs:555-695) for requests of different committees
    and batch sizes in one fused pass.  Each request is a dict with ct1 (count G1 affine), ct2
    (count Fp12), valid (count bools), batch_size (the size its lines/keys were prepared for),
    lines, keys, optionally key_lines (the keys' line tables), shares (T G1 Jacobian) and
    n_accepted.  Returns one message tensor per request."""
    arr = (_BtxRequest * len(requests))()
    keep, outs = [], []
    for k, r in enumerate(requests):
        ct2 = _dev(r["ct2"])
        n = ct2.numel() // FP12
        out = _empty(n * FP12, ct2)
        v = (ctypes.c_uint8 * max(n, 1))(*[1 if x else 0 for x in r["valid"]])
        keep.append(v)
        outs.append(out)
        kl = r.get("key_lines")
        arr[k] = _BtxRequest(_ptr(_dev(r["ct1"])), _ptr(ct2), ctypes.cast(v, ctypes.c_void_p), n,
                             r["batch_size"], _ptr(_dev(r["lines"])), _ptr(_dev(r["keys"])),
                             _ptr(_dev(kl)) if kl is not None else ctypes.c_void_p(0),
                             _ptr(_dev(r["shares"])), r["n_accepted"], _ptr(out))
    call("bte_btx_decrypt_requests", ctypes.cast(arr, ctypes.c_void_p), len(requests), _stream())
    return outs


def ifft_cyclotomic_prefix_unscaled_segmented(values, log_sizes, output_counts):
    """Prefix inverse cyclotomic transforms of uneven sizes (back to back); returns the
    concatenated prefixes."""
    v = _dev(values).clone()
    if v.numel() != sum(1 << l for l in log_sizes) * FP12:
        raise ValueError("values.len() != sum of sizes")
    la = (ctypes.c_int * len(log_sizes))(*log_sizes)
    oa = (ctypes.c_uint32 * len(log_sizes))(*output_counts)
    call("bte_ifft_cyclotomic_prefix_unscaled_segmented", _ptr(v), la, oa, len(log_sizes), _stream())
    parts, off = [], 0
    for l, c in zip(log_sizes, output_counts):
        parts.append(v[off * FP12:(off + c) * FP12])
        off += 1 << l
    return torch.cat(parts) if parts else v[:0]


def btx_open_prepared_batched(shares_jac, n_accepted, key_lines, ct2, beta, valid, batch_size):
    """open_batch against the opening keys' line tables (prepare_g2_lines(keys))."""
    ct2 = _dev(ct2)
    n = ct2.numel() // FP12
    count = n // batch_size
    v = (ctypes.c_uint8 * n)(*[1 if x else 0 for x in valid])
    out = _empty(n * FP12, ct2)
    call("bte_btx_open_prepared_batched", _ptr(_dev(shares_jac)), n_accepted, _ptr(_dev(key_lines)), _ptr(ct2),
         _ptr(_dev(beta)), v, batch_size, count, _ptr(out), _stream())
    return out


def btx_keygen(poly, q, party_offsets, max_batch, verification_keys=False):
    """setup.rs:98-233 key material from the dealer's choices (polynomial coefficients and party
    q's as canonical ints): returns the device G2 material (negative then positive blocks) and,
    if asked, the L x N verification keys."""
    N, L = len(q), max_batch
    W = party_offsets[-1]
    mat = torch.empty((2 * L - 1) * W * G2_AFFINE, dtype=torch.uint8, device="cuda")
    vk = torch.empty(L * N * G2_AFFINE, dtype=torch.uint8, device="cuda") if verification_keys else None
    offs = (ctypes.c_uint64 * len(party_offsets))(*party_offsets)
    call("bte_btx_keygen", _host_scalars(poly), len(poly), _host_scalars(q), offs, N, L, _ptr(mat), _ptr(vk),
         _stream())
    return (mat, vk) if verification_keys else mat


def pfe_keygen(poly, rho, alpha, beta, party_offsets, batch_size, verification_keys=False):
    """weighted_pfe setup.rs:149-354 key material from the dealer's choices (canonical ints):
    returns (material d1 || d2 on the device, global key on the device, fractions g[j][i] as
    ints, verification keys on the device or None)."""
    N, B = len(rho), batch_size
    W = party_offsets[-1]
    mat = torch.empty(2 * B * W * G2_AFFINE, dtype=torch.uint8, device="cuda")
    gk = torch.empty(G2_AFFINE, dtype=torch.uint8, device="cuda")
    vk = torch.empty(B * N * G2_AFFINE, dtype=torch.uint8, device="cuda") if verification_keys else None
    frac = ctypes.create_string_buffer(32 * N * B)
    offs = (ctypes.c_uint64 * len(party_offsets))(*party_offsets)
    call("bte_pfe_keygen", _host_scalars(poly), len(poly), _host_scalars(rho), _host_scalars(alpha),
         _host_scalars(beta), offs, N, B, _ptr(mat), _ptr(vk), _ptr(gk), frac, _stream())
    raw = frac.raw
    g = [[int.from_bytes(raw[32 * (j * B + i): 32 * (j * B + i) + 32], "little") for i in range(B)]
         for j in range(N)]
    return mat, gk, g, vk


def btx_verify_shares(shares_jac, party_index, verification_keys, n_parties, max_batch, ct1_affine, valid):
    """verify_decryption_share (decryption.rs:236-272) for every share: a list of bools."""
    T = len(party_index)
    B = len(valid)
    pidx = (ctypes.c_uint32 * max(T, 1))(*party_index)
    v = (ctypes.c_uint8 * B)(*[1 if x else 0 for x in valid])
    ok = (ctypes.c_uint8 * max(T, 1))()
    call("bte_btx_verify_shares", _ptr(_dev(shares_jac)), pidx, T, _ptr(_dev(verification_keys)), n_parties,
         max_batch, _ptr(_dev(ct1_affine)), v, B, ok, _stream())
    return [bool(ok[j]) for j in range(T)]


def btx_verify_shares_batched(shares_jac, party_index, challenges, verification_keys, n_parties, max_batch,
                              ct1_affine, valid):
    """verify_decryption_shares_batched (decryption.rs:356-422) with the given challenges."""
    T = len(party_index)
    B = len(valid)
    pidx = (ctypes.c_uint32 * max(T, 1))(*party_index)
    v = (ctypes.c_uint8 * B)(*[1 if x else 0 for x in valid])
    res = ctypes.c_int(0)
    call("bte_btx_verify_shares_batched", _ptr(_dev(shares_jac)), pidx, _host_scalars(challenges), T,
         _ptr(_dev(verification_keys)), n_parties, max_batch, _ptr(_dev(ct1_affine)), v, B, ctypes.byref(res),
         _stream())
    return bool(res.value)


class BtxPlan:
    """Launch-only plan of the fused decrypt step for `count` batches of B ciphertexts of one
    committee (kernel_lines / key_lines from btx_prepare and prepare_g2_lines(keys)); run()
    enqueues kernels only, launch_graph() replays a HIP graph of one run (bound to the buffers of
    its first call; the current stream must not be the null stream)."""

    def __init__(self, batch_size, count, n_accepted, kernel_lines, key_lines, valid):
        self._keep = (_dev(kernel_lines), _dev(key_lines))
        self.B, self.K, self.T = batch_size, count, n_accepted
        v = (ctypes.c_uint8 * (batch_size * count))(*[1 if x else 0 for x in valid])
        h = ctypes.c_void_p(0)
        call("bte_btx_plan_create", batch_size, count, n_accepted, _ptr(self._keep[0]), _ptr(self._keep[1]), v,
             ctypes.byref(h), _stream())
        self.handle = h

    def _args(self, ct1, shares, ct2, out):
        return (self.handle, _ptr(_dev(ct1)), _ptr(_dev(shares)), _ptr(_dev(ct2)), _ptr(out), _stream())

    def run(self, ct1, shares, ct2, out=None):
        out = out if out is not None else _empty(self.B * self.K * FP12, ct2)
        call("bte_btx_plan_run", *self._args(ct1, shares, ct2, out))
        return out

    def launch_graph(self, ct1, shares, ct2, out):
        call("bte_btx_plan_launch_graph", *self._args(ct1, shares, ct2, out))
        return out

    def close(self):
        if self.handle:
            call("bte_btx_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def btx_partial_decrypt(ct1_affine, q_scalars):
    """partial_decrypt (decryption.rs:205-233) for every party key q_j -> sigma (N x G1 Jacobian)."""
    ct1 = _dev(ct1_affine)
    B = ct1.numel() // G1_AFFINE
    N = len(q_scalars)
    qb = b"".join(int(q).to_bytes(32, "little") for q in q_scalars)
    buf = ctypes.create_string_buffer(qb, len(qb))
    out = _empty(N * G1_JAC, ct1)
    call("bte_btx_partial_decrypt", _ptr(ct1), B, buf, N, _ptr(out), _stream())
    return out


# --- a16: weighted_pfe phases (weighted_pfe/src/decryption.rs) ------------------------------
def _host_scalars(vals):
    b = b"".join(int(v).to_bytes(32, "little") for v in vals)
    return ctypes.create_string_buffer(b, len(b))


def pfe_prepare(material, total_weight, batch_size, party_offsets, accepted, alpha, global_key):
    """prepare_decryption (decryption.rs:421-537) -> (d1_by_party, total_lines, mask_lines)."""
    material = _dev(material)
    B, T = batch_size, len(accepted)
    offs = (ctypes.c_uint64 * len(party_offsets))(*party_offsets)
    acc = (ctypes.c_uint32 * T)(*accepted)
    d1p = _empty(B * T * G2_AFFINE, material)
    tl = _empty(B * LINES_BYTES, material)
    ml = _empty(B * LINES_BYTES, material)
    call("bte_pfe_prepare", _ptr(material), total_weight, B, offs, len(party_offsets) - 1, acc, T,
         _host_scalars(alpha), _ptr(_dev(global_key)), _ptr(d1p), _ptr(tl), _ptr(ml), _stream())
    return d1p, tl, ml


def btx_verify_proofs(commitments, ct1_affine, responses, challenges):
    """SchnorrProof::verify (weighted_btx/src/proof.rs:56-70) for every ciphertext of a batch,
    given the hashed challenges (canonical ints); returns one bool per ciphertext."""
    B = len(responses)
    if len(challenges) != B:
        raise ValueError("btx_verify_proofs: %d responses for %d challenges" % (B, len(challenges)))
    if B == 0:
        return []
    ok = (ctypes.c_uint8 * B)()
    call("bte_btx_verify_proofs", _ptr(_dev(commitments)), _ptr(_dev(ct1_affine)), _host_scalars(responses),
         _host_scalars(challenges), B, ok, _stream())
    return [bool(ok[i]) for i in range(B)]


def pfe_verify_proofs_batched(commitments, ct1_affine, responses, challenges, coefficients):
    """verify_batch (weighted_pfe/src/proof.rs:74-122) given the hashed challenges and the
    verifier's nonzero coefficients (canonical ints)."""
    B = len(responses)
    res = ctypes.c_int(0)
    call("bte_pfe_verify_proofs_batched", _ptr(_dev(commitments)), _ptr(_dev(ct1_affine)), _host_scalars(responses),
         _host_scalars(challenges), _host_scalars(coefficients), B, ctypes.byref(res), _stream())
    return bool(res.value)


def pfe_precompute_batched(ct1_affine, total_lines, batch_size):
    """precompute_batch for count = len / B batches of one committee in one fused pass."""
    ct1 = _dev(ct1_affine)
    n = ct1.numel() // G1_AFFINE
    count = n // batch_size
    fc = _empty(n * G1_JAC, ct1)
    pc = _empty(n * FP12, ct1)
    call("bte_pfe_precompute_batched", _ptr(ct1), batch_size, count, _ptr(_dev(total_lines)), _ptr(fc), _ptr(pc),
         _stream())
    return fc, pc


def pfe_open_prepared_batched(shares_jac, n_accepted, d1_lines, mask_lines, ct1_affine, ct2, first_cauchy,
                              pairing_cauchy, alpha, batch_size):
    """open_batch for count batches against d1_lines = prepare_g2_lines(d1_by_party)."""
    ct2 = _dev(ct2)
    n = ct2.numel() // FP12
    out = _empty(n * FP12, ct2)
    call("bte_pfe_open_prepared_batched", _ptr(_dev(shares_jac)), n_accepted, _ptr(_dev(d1_lines)),
         _ptr(_dev(mask_lines)), _ptr(_dev(ct1_affine)), _ptr(ct2), _ptr(_dev(first_cauchy)),
         _ptr(_dev(pairing_cauchy)), _host_scalars(alpha), batch_size, n // batch_size, _ptr(out), _stream())
    return out


def pfe_precompute(ct1_affine, total_lines):
    """precompute_batch (decryption.rs:540-573) -> (first_cauchy, pairing_cauchy)."""
    ct1 = _dev(ct1_affine)
    B = ct1.numel() // G1_AFFINE
    fc = _empty(B * G1_JAC, ct1)
    pc = _empty(B * FP12, ct1)
    call("bte_pfe_precompute", _ptr(ct1), B, _ptr(_dev(total_lines)), _ptr(fc), _ptr(pc), _stream())
    return fc, pc


def pfe_open(shares_jac, d1_by_party, mask_lines, ct1_affine, ct2, first_cauchy, pairing_cauchy, alpha):
    """open_batch (decryption.rs:576-644) -> messages (B x Fp12)."""
    shares = _dev(shares_jac)
    ct2 = _dev(ct2)
    B = ct2.numel() // FP12
    out = _empty(B * FP12, ct2)
    call("bte_pfe_open", _ptr(shares), shares.numel() // G1_JAC, _ptr(_dev(d1_by_party)),
         _ptr(_dev(mask_lines)), _ptr(_dev(ct1_affine)), _ptr(ct2), _ptr(_dev(first_cauchy)),
         _ptr(_dev(pairing_cauchy)), _host_scalars(alpha), B, _ptr(out), _stream())
    return out


def pfe_partial_decrypt(ct1_affine, fractions):
    """partial_decrypt (decryption.rs:230-251); fractions[j] = [g_{j,0}, ..., g_{j,B-1}]."""
    ct1 = _dev(ct1_affine)
    B = ct1.numel() // G1_AFFINE
    N = len(fractions)
    fr = to_device(b"".join(int(g).to_bytes(32, "little") for row in fractions for g in row), ct1.device)
    out = _empty(N * G1_JAC, ct1)
    call("bte_pfe_partial_decrypt", _ptr(ct1), B, _ptr(fr), N, _ptr(out), _stream())
    return out


class PfePlan:
    """Plan of the fused weighted_pfe step for `count` batches of one committee (total_lines /
    mask_lines from pfe_prepare, d1_lines = prepare_g2_lines(d1_by_party), alpha the B alpha
    points): run() is pfe_precompute_batched + pfe_open_prepared_batched, launch_graph() replays a
    HIP graph of one run (bound to the buffers of its first call; not on the null stream)."""

    def __init__(self, batch_size, count, n_accepted, total_lines, d1_lines, mask_lines, alpha):
        self._keep = (_dev(total_lines), _dev(d1_lines), _dev(mask_lines))
        self.B, self.K, self.T = batch_size, count, n_accepted
        h = ctypes.c_void_p(0)
        call("bte_pfe_plan_create", batch_size, count, n_accepted, _ptr(self._keep[0]), _ptr(self._keep[1]),
             _ptr(self._keep[2]), _host_scalars(alpha), ctypes.byref(h), _stream())
        self.handle = h

    def _args(self, ct1, shares, ct2, out):
        return (self.handle, _ptr(_dev(ct1)), _ptr(_dev(shares)), _ptr(_dev(ct2)), _ptr(out), _stream())

    def run(self, ct1, shares, ct2, out=None):
        out = out if out is not None else _empty(self.B * self.K * FP12, ct2)
        call("bte_pfe_plan_run", *self._args(ct1, shares, ct2, out))
        return out

    def launch_graph(self, ct1, shares, ct2, out):
        call("bte_pfe_plan_launch_graph", *self._args(ct1, shares, ct2, out))
        return out

    def close(self):
        if self.handle:
            call("bte_pfe_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
