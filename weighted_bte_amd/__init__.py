"""weighted_bte_amd: MI355X (gfx950) BLS12-381 hot path of weighted_btx / weighted_pfe batch
decryption behind a C ABI (include/bte_hip.h).

This package is a thin ctypes binding over lib/libbte_hip.so plus a Python mirror of the
reference's crate-private L1 functions (weighted_bte_amd.api).  There is no CPU fallback: if the
library is not built, loading fails with an error that says how to build it.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbte_hip.so")

_lib = None


class BteError(RuntimeError):
    pass


_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGNATURES = {
    "bte_last_error": [],
    "bte_version": [],
    "bte_g1_msm": [_vp, _vp, _sz, _vp, _vp],
    "bte_g2_msm": [_vp, _vp, _sz, _vp, _vp],
    "bte_g2_msm_indexed": [_vp, _sz, _vp, _vp, _sz, _vp, _vp],
    "bte_msm_set_window": [_i, _sz, _i],
    "bte_g1_msm_batched": [_vp, _sz, _vp, _vp, _sz, _vp, _sz, _vp, _vp, _vp],
    "bte_g2_msm_batched": [_vp, _sz, _vp, _vp, _sz, _vp, _sz, _vp, _vp, _vp],
    "bte_g1_batch_normalize": [_vp, _sz, _vp, _vp],
    "bte_g2_batch_normalize": [_vp, _sz, _vp, _vp],
    "bte_g1_generator_mul": [_vp, _sz, _vp, _vp],
    "bte_g2_generator_mul": [_vp, _sz, _vp, _vp],
    "bte_fft_g1": [_vp, _sz, _i, _i, _vp],
    "bte_fft_set_dft_max_log": [_i],
    "bte_tune_set": [_i, ctypes.c_int64],
    "bte_tune_get": [_i, _vp],
    "bte_fft_g2": [_vp, _sz, _i, _i, _vp],
    "bte_fft_gt": [_vp, _sz, _i, _i, _vp],
    "bte_fft_cyclotomic": [_vp, _sz, _i, _vp],
    "bte_ifft_cyclotomic_prefix_unscaled": [_vp, _sz, _sz, _vp],
    "bte_cyclotomic_pow_scalars": [_vp, _vp, _sz, _vp],
    "bte_cyclotomic_pow_small": [_vp, _vp, _sz, _vp],
    "bte_g2_prepare_lines": [_vp, _sz, _vp, _vp],
    "bte_miller_loop_prepared": [_vp, _vp, _sz, _vp, _vp],
    "bte_miller_loop_pairs": [_vp, _vp, _sz, _vp, _vp],
    "bte_multi_miller_loop": [_vp, _vp, _sz, _vp, _sz, _vp, _vp],
    "bte_multi_miller_loop_noalloc": [_vp, _vp, _sz, _vp, _sz, _vp, _vp],
    "bte_final_exp_easy": [_vp, _sz, _vp, _vp],
    "bte_final_exp_hard": [_vp, _sz, _vp, _vp],
    "bte_final_exp": [_vp, _sz, _vp, _vp],
    "bte_lagrange_at_zero": [_sz, _vp, _sz, _vp],
    "bte_lagrange_at_zero_device": [_sz, _vp, _sz, _vp, _vp],
    "bte_fft_fr": [_vp, _sz, _i, _i, _vp],
    "bte_fft_fr_batched": [_vp, _sz, _sz, _i, _i, _vp],
    "bte_btx_prepare": [_vp, _sz, _sz, _vp, _sz, _vp, _sz, _sz, _vp, _vp, _vp],
    "bte_btx_precompute": [_vp, _vp, _sz, _vp, _vp, _vp],
    "bte_btx_open": [_vp, _sz, _vp, _vp, _vp, _vp, _sz, _vp, _vp],
    "bte_btx_partial_decrypt": [_vp, _sz, _vp, _sz, _vp, _vp],
    "bte_btx_precompute_batched": [_vp, _vp, _sz, _sz, _vp, _vp, _vp],
    "bte_btx_open_batched": [_vp, _sz, _vp, _vp, _vp, _vp, _sz, _sz, _vp, _vp],
    "bte_miller_loop_prepared_tiled": [_vp, _sz, _vp, _sz, _vp, _vp],
    "bte_fft_g1_batched": [_vp, _sz, _sz, _i, _i, _vp],
    "bte_fft_g2_batched": [_vp, _sz, _sz, _i, _i, _vp],
    "bte_fft_cyclotomic_batched": [_vp, _sz, _sz, _i, _vp],
    "bte_ifft_cyclotomic_prefix_unscaled_batched": [_vp, _sz, _sz, _sz, _vp],
    "bte_pfe_prepare": [_vp, _sz, _sz, _vp, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp],
    "bte_pfe_precompute": [_vp, _sz, _vp, _vp, _vp, _vp],
    "bte_pfe_open": [_vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp],
    "bte_pfe_partial_decrypt": [_vp, _sz, _vp, _sz, _vp, _vp],
    "bte_fft_g1_segmented": [_vp, _vp, _sz, _i, _i, _vp],
    "bte_fft_g2_segmented": [_vp, _vp, _sz, _i, _i, _vp],
    "bte_ifft_cyclotomic_prefix_unscaled_segmented": [_vp, _vp, _vp, _sz, _vp],
    "bte_miller_loop_prepared_segmented": [_vp, _vp, _sz, _vp, _sz, _vp, _vp],
    "bte_btx_decrypt_requests": [_vp, _sz, _vp],
    "bte_btx_open_prepared_batched": [_vp, _sz, _vp, _vp, _vp, _vp, _sz, _sz, _vp, _vp],
    "bte_pfe_precompute_batched": [_vp, _sz, _sz, _vp, _vp, _vp, _vp],
    "bte_btx_plan_create": [_sz, _sz, _sz, _vp, _vp, _vp, ctypes.POINTER(_vp), _vp],
    "bte_btx_plan_run": [_vp, _vp, _vp, _vp, _vp, _vp],
    "bte_btx_plan_launch_graph": [_vp, _vp, _vp, _vp, _vp, _vp],
    "bte_btx_plan_destroy": [_vp],
    "bte_pfe_plan_create": [_sz, _sz, _sz, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp), _vp],
    "bte_pfe_plan_run": [_vp, _vp, _vp, _vp, _vp, _vp],
    "bte_pfe_plan_launch_graph": [_vp, _vp, _vp, _vp, _vp, _vp],
    "bte_pfe_plan_destroy": [_vp],
    "bte_btx_keygen": [_vp, _sz, _vp, _vp, _sz, _sz, _vp, _vp, _vp],
    "bte_btx_verify_shares": [_vp, _vp, _sz, _vp, _sz, _sz, _vp, _vp, _sz, _vp, _vp],
    "bte_pfe_verify_proofs_batched": [_vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp],
    "bte_btx_verify_proofs": [_vp, _vp, _vp, _vp, _sz, _vp, _vp],
    "bte_pfe_keygen": [_vp, _sz, _vp, _vp, _vp, _vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp],
    "bte_g1_compress": [_vp, _sz, _vp, _vp],
    "bte_g2_compress": [_vp, _sz, _vp, _vp],
    "bte_g1_decompress": [_vp, _sz, _vp, _vp, _vp],
    "bte_g2_decompress": [_vp, _sz, _vp, _vp, _vp],
    "bte_btx_verify_shares_batched": [_vp, _vp, _vp, _sz, _vp, _sz, _sz, _vp, _vp, _sz, _vp, _vp],
    "bte_pfe_open_prepared_batched": [_vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _sz, _vp, _vp],
    "bte_multi_miller_loop_prepared": [_vp, _vp, _sz, _vp, _sz, _vp, _vp],
    "bte_multi_pairing_prepared_fused": [_vp, _vp, _sz, _vp, _sz, _vp, _vp],
    "bte_profile_enable": [_i],
    "bte_cyclotomic_prefix_work": [_sz, _sz, _vp],
    "bte_profile_read": [ctypes.c_char_p, _sz],
    "bte_probe_mac_rate": [ctypes.POINTER(ctypes.c_double), _vp],
    "bte_device_alloc": [ctypes.POINTER(_vp), _sz],
    "bte_device_free": [_vp],
    "bte_memcpy_h2d": [_vp, _vp, _sz, _vp],
    "bte_memcpy_d2h": [_vp, _vp, _sz, _vp],
    "bte_stream_synchronize": [_vp],
    "bte_stream_reserve": [_vp, _sz, _sz],
    "bte_test_fail_alloc": [_i],
}
_RESTYPES = {"bte_last_error": ctypes.c_char_p, "bte_version": ctypes.c_char_p}


def lib():
    """Load (once) and return the ctypes handle of libbte_hip.so."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BteError(
                "%s is not built: run `python -m weighted_bte_amd.build` (hipcc, gfx950)" % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = handle
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().bte_last_error().decode(errors="replace")
        raise BteError("bte status %d: %s" % (rc, msg))
    return rc


def call(name, *args):
    """Call a C ABI entry point and raise BteError on a nonzero status."""
    return check(getattr(lib(), name)(*args))


def profile_enable(on=True):
    call("bte_profile_enable", 1 if on else 0)


def profile_read():
    """Per-phase milliseconds of the most recent profiled call, as a dict."""
    buf = ctypes.create_string_buffer(4096)
    call("bte_profile_read", buf, len(buf))
    out = {}
    for item in buf.value.decode().split(";"):
        if "=" in item:
            k, v = item.split("=")
            out[k] = float(v)
    return out


def probe_mac_rate(stream=None):
    v = ctypes.c_double(0.0)
    call("bte_probe_mac_rate", ctypes.byref(v), ctypes.c_void_p(stream or 0))
    return v.value
