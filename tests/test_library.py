"""CPU checks of the C ABI library: it is built, it loads, and it exports every function that
include/bte_hip.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

from helpers import ROOT

HEADER = os.path.join(ROOT, "include", "bte_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bte_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "bte_g1_msm" in names and "bte_g2_msm_indexed" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    import weighted_bte_amd

    lib = weighted_bte_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in lib.bte_version()


def test_python_binding_covers_header():
    import weighted_bte_amd

    missing = [n for n in declared_functions() if n not in weighted_bte_amd._SIGNATURES]
    assert not missing, missing


def test_error_path_without_gpu_work():
    """Argument validation returns a status instead of aborting (reference panics)."""
    import weighted_bte_amd

    lib = weighted_bte_amd.lib()
    rc = lib.bte_msm_set_window(7, 16, 3)
    assert rc == 1
    assert b"bad arguments" in lib.bte_last_error()


def test_cyclotomic_prefix_work_counter_matches_oracle_plans():
    """The library's count of the prefix inverse cyclotomic transform's operations (used for the
    decrypt roofline) equals the count from the oracle's own plans of the same twiddles
    (fft.rs:33-154 decomposition and width-4 NAF), for the bench shape and a small one.  Host-only
    call: no GPU needed."""
    import weighted_bte_amd
    from helpers import O

    lib = weighted_bte_amd.lib()
    for n, out in ((16, 5), (1024, 512)):
        counts = (ctypes.c_uint64 * 3)()
        assert lib.bte_cyclotomic_prefix_work(n, out, counts) == 0
        winv = O.Radix2Domain(n).generator_inverse
        logn = n.bit_length() - 1
        sqr = mul = frob = 0

        def pow_work(s):
            nonlocal sqr, mul, frob
            nafs, count = O.cyclotomic_scalar_plan(s)
            if count == 0:
                return
            nnz = sum(1 for d in nafs for x in d if x)
            sqr += count
            mul += 3 + nnz - 1
            frob += 12

        for lh in range(logn - 1):
            half = 1 << lh
            for k in range(half):
                for _ in range(n // (2 * half)):
                    if k:
                        pow_work(pow(winv, k << (logn - 1 - lh), O.R))
                    mul += 2
        for k in range(out):
            if k:
                pow_work(pow(winv, k, O.R))
            mul += 1
        assert list(counts) == [sqr, mul, frob], n


def test_tune_knobs_round_trip_and_range_checks():
    """bte_tune_set / bte_tune_get (launch-shape knobs, host-only): defaults, a round trip, and
    out-of-range values or unknown knobs refused with BTE_EINVAL and the value left unchanged."""
    import weighted_bte_amd
    from weighted_bte_amd import api

    lib = weighted_bte_amd.lib()
    for knob, default in api.TUNE_DEFAULTS.items():
        assert api.tune_get(knob) == default
    try:
        api.tune_set(api.TUNE_MSM_ACC_ROUNDS, 4)
        assert api.tune_get(api.TUNE_MSM_ACC_ROUNDS) == 4
        api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, 0)
        assert api.tune_get(api.TUNE_CYC_THREE_LANE_MAX) == 0
        for knob, bad in ((api.TUNE_MSM_ACC_ROUNDS, 0), (api.TUNE_MSM_ACC_ROUNDS, 17), (api.TUNE_MM_CHUNK_MAX, 65),
                          (api.TUNE_MM_MIN_THREADS, 0), (api.TUNE_CYC_THREE_LANE_MAX, -1)):
            before = api.tune_get(knob)
            assert lib.bte_tune_set(knob, bad) == 1
            assert api.tune_get(knob) == before
        assert lib.bte_tune_set(9, 1) == 1
        assert b"unknown knob" in lib.bte_last_error()
    finally:
        for knob, default in api.TUNE_DEFAULTS.items():
            api.tune_set(knob, default)
