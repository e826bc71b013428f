"""Autotuner for the launch shapes of the hot path.  None of them changes a result; blst and
rayon make the corresponding choices themselves in the reference (Pippenger window,
blst_utils.rs:57-64; chunk and pair scheduling of the transforms, fft.rs:370-386).

  * Pippenger window width c per (group, log2 n): the plain per-call MSM picks
    c = log2 n - 4 up to 2^12, then 8 / 10 / 13 / 16 for log2 n up to 14 / 16 / 18 / above.
        best, times = tune_msm(1, points, scalars)
  * bucket-accumulation rounds (BTE_TUNE_MSM_ACC_ROUNDS): chunk length of the MSM's sorted
    entries per thread, as full-occupancy rounds of the device.
        best, times = tune_acc_rounds(points, scalars)
  * three-lane switch of the cyclotomic stages (BTE_TUNE_CYC_THREE_LANE_MAX): the stage width
    below which each Fp12 runs on three lanes instead of one.
        best, table = tune_three_lane(n=1024)
  * multi-Miller chunk (BTE_TUNE_MM_CHUNK_MAX with BTE_TUNE_MM_MIN_THREADS): pairs sharing
    one Miller accumulator.
        (cmax, tmin), times = tune_mm_chunk(lines, points, seg_offsets)
  * largest G1/G2 transform run as one batched MSM instead of staged (bte_fft_set_dft_max_log,
    built-in 2^10).
        cutoff, times = tune_dft_cutoff(2, jacobian_points_of_size_2_to_the_10)

    save_table({("g1", 20): best}, knobs={...})     # keyed by device name
    apply_table()                                   # re-applies saved choices in a new process

Times are wall-clock per call on inputs already in HBM (best of `reps` after one warm-up), so
they include the host-side work the per-call entry points perform.
"""
from __future__ import annotations

import json
import math
import os
import time

import torch

from . import api

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "msm_windows.json")
C_MIN, C_MAX = 4, 16


def default_window(n: int) -> int:
    """The library's built-in window width for n points (msm.hip default_window_bits)."""
    lg = max(0, n.bit_length() - 1)
    if lg <= 12:
        return max(C_MIN, lg - 4)
    return 8 if lg <= 14 else 10 if lg <= 16 else 13 if lg <= 18 else 16


def _msm_fn(group):
    return api.g1_multi_exp_affine_bytes if group == 1 else api.g2_multi_exp_affine_bytes


def time_msm(group, points, scalars, reps=3):
    fn = _msm_fn(group)
    fn(points, scalars)
    torch.cuda.synchronize()
    best = math.inf
    for _ in range(reps):
        t = time.perf_counter()
        fn(points, scalars)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    return best * 1e3


def tune_msm(group, points, scalars, candidates=None, reps=3):
    """Times every candidate window for this (group, n), installs the fastest and returns
    (best c, {c: ms})."""
    n = scalars.numel() // 32
    c0 = default_window(n)
    cands = candidates or [c for c in range(c0 - 2, c0 + 3) if C_MIN <= c <= C_MAX]
    times = {}
    try:
        for c in cands:
            api.set_msm_window(group, n, c)
            times[c] = round(time_msm(group, points, scalars, reps), 4)
    finally:
        api.set_msm_window(group, n, 0)
    best = min(times, key=times.get)
    api.set_msm_window(group, n, best)
    return best, times


def _time(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    best = math.inf
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    return round(best * 1e3, 4)


def tune_knob(knob, candidates, workload, reps=3):
    """Times workload() under each candidate value of a bte_tune_set knob, installs the fastest
    and returns (best, {value: ms})."""
    old = api.tune_get(knob)
    times = {}
    try:
        for v in candidates:
            api.tune_set(knob, v)
            times[v] = _time(workload, reps)
    finally:
        api.tune_set(knob, old)
    best = min(times, key=times.get)
    api.tune_set(knob, best)
    return best, times


def tune_acc_rounds(points, scalars, group=1, candidates=(1, 2, 3, 4)):
    """Bucket-accumulation rounds for one MSM shape (the window already installed)."""
    fn = _msm_fn(group)
    return tune_knob(api.TUNE_MSM_ACC_ROUNDS, candidates, lambda: fn(points, scalars))


def tune_mm_chunk(lines, points, seg_offsets, candidates=(1, 2, 4, 8, 16), reps=3):
    """Pairs per multi-Miller accumulator on a prepared multi-pairing workload.  The chunk the
    library uses is min(BTE_TUNE_MM_CHUNK_MAX, pairs / BTE_TUNE_MM_MIN_THREADS), so each
    candidate c is timed with the thread floor at 1 (exactly c pairs per accumulator) beside the
    built-in pair of values; installs the fastest pair and returns
    ((chunk max, min threads), {"c/min": ms})."""
    knobs = (api.TUNE_MM_CHUNK_MAX, api.TUNE_MM_MIN_THREADS)
    old = tuple(api.tune_get(k) for k in knobs)
    default = tuple(api.TUNE_DEFAULTS[k] for k in knobs)
    settings = [default] + [(c, 1) for c in candidates]
    times = {}
    fn = lambda: api.multi_miller_loop_prepared(lines, points, seg_offsets)
    try:
        for pair in settings:
            for k, v in zip(knobs, pair):
                api.tune_set(k, v)
            times[pair] = _time(fn, reps)
    finally:
        for k, v in zip(knobs, old):
            api.tune_set(k, v)
    best = min(times, key=times.get)
    for k, v in zip(knobs, best):
        api.tune_set(k, v)
    return best, {"%d/%d" % pair: ms for pair, ms in times.items()}


def tune_three_lane(n=1024, counts=(2, 8, 16, 32, 64, 128), reps=2):
    """Times the batched prefix inverse cyclotomic transform of `count` size-n transforms with
    every stage on one lane per Fp12 and with every stage on three lanes, and installs the
    switch at the narrowest stage width where one lane wins.  Returns (threshold, rows)."""
    from . import call
    from .api import _ptr, _stream

    rows = []
    old = api.tune_get(api.TUNE_CYC_THREE_LANE_MAX)
    g = torch.Generator(device="cpu").manual_seed(5)
    try:
        for count in counts:
            # any Fp12 values time the same: below p limb-wise, top limb masked
            v = torch.randint(0, 1 << 31, (count * n * 144,), dtype=torch.int64, generator=g)
            v = v.to(torch.int32).view(count * n, 12, 12)
            v[:, :, 11] &= 0x0FFFFFFF
            base = v.reshape(-1).to("cuda")
            row = {"count": count, "stage_butterflies": n // 2 * count}
            for name, thr in (("one_lane_ms", 0), ("three_lane_ms", 1 << 30)):
                api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, thr)
                work = base.clone()
                row[name] = _time(lambda: call("bte_ifft_cyclotomic_prefix_unscaled_batched", _ptr(work), n, count,
                                               n // 2, _stream()), reps)
            rows.append(row)
    finally:
        api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, old)
    # the narrowest measured width from which one lane wins at every wider width; if three
    # lanes still win at the widest width measured, switch just above it
    rows.sort(key=lambda r: r["stage_butterflies"])
    thr = 2 * rows[-1]["stage_butterflies"] if rows else old
    for r in reversed(rows):
        if r["one_lane_ms"] > r["three_lane_ms"]:
            break
        thr = r["stage_butterflies"]
    api.tune_set(api.TUNE_CYC_THREE_LANE_MAX, thr)
    return thr, rows


DFT_MAX_LOG_DEFAULT = 10


def set_dft_max_log(log_n):
    """G1/G2 transforms of size up to 2^log_n (with n^2 x count <= 2^20) run as one batched MSM
    (bte_fft_set_dft_max_log; 0 = always staged)."""
    from . import call
    call("bte_fft_set_dft_max_log", int(log_n))


def tune_dft_cutoff(group, values, reps=3):
    """Times the forward transform of `values` (n Jacobian G1 or G2 points, n = 2^logn with
    2 <= logn <= 10) in the MSM form and in the staged form and installs the cutoff that takes the
    faster one: logn when the MSM form wins, logn - 1 otherwise (no result depends on it).
    Returns (cutoff, {"msm_ms": ..., "staged_ms": ...})."""
    from .api import G1_JAC, G2_JAC, Radix2Domain

    n = values.numel() // (G1_JAC if group == 1 else G2_JAC)
    logn = n.bit_length() - 1
    if n != 1 << logn or not 2 <= logn <= 10:
        raise ValueError("tune_dft_cutoff: size must be 2^2 .. 2^10")
    d = Radix2Domain(n)
    fwd = d.fft_g1 if group == 1 else d.fft_g2
    times = {}
    try:
        for name, cut in (("msm_ms", logn), ("staged_ms", logn - 1)):
            set_dft_max_log(cut)
            times[name] = _time(lambda: fwd(values), reps)
    finally:
        set_dft_max_log(DFT_MAX_LOG_DEFAULT)
    best = logn if times["msm_ms"] <= times["staged_ms"] else logn - 1
    set_dft_max_log(best)
    return best, times


KNOBS = {"msm_acc_rounds": 1, "cyc_three_lane_max": 2, "mm_chunk_max": 3, "mm_min_threads": 4}


def _device_key():
    return torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"


def save_table(choices, path=TABLE, knobs=None):
    """choices: {("g1" | "g2", log2 n): c}; knobs: {name in KNOBS or "dft_max_log": value};
    merged into the table of this device."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    dev = data.setdefault(_device_key(), {})
    for (group, logn), c in choices.items():
        dev["%s_%d" % (group, logn)] = int(c)
    for name, v in (knobs or {}).items():
        if name not in KNOBS and name != "dft_max_log":
            raise ValueError("unknown knob %r" % name)
        dev[name] = int(v)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def apply_table(path=TABLE):
    """Installs the saved windows of this device; returns how many were applied."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        data = json.load(f)
    applied = 0
    for key, c in data.get(_device_key(), {}).items():
        if key in KNOBS:
            api.tune_set(KNOBS[key], int(c))
        elif key == "dft_max_log":
            set_dft_max_log(int(c))
        else:
            group, logn = key.split("_")
            api.set_msm_window(1 if group == "g1" else 2, 1 << int(logn), int(c))
        applied += 1
    return applied
