"""Autotuner for the Pippenger window width (the plain per-call MSM picks
c = clamp(log2 n - 4, 4, 16); blst picks its own, blst_utils.rs:57-64, and the result does not
depend on it).

    best, times = tune_msm(1, points, scalars)      # sweeps c around the default, sets the best
    save_table({("g1", 20): best})                  # keyed by device name
    apply_table()                                   # re-applies saved choices in a new process

Times are wall-clock per call on inputs already in HBM (best of `reps` after one warm-up), so
they include the host-side window combination the per-call entry point performs.
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
    lg = max(0, n.bit_length() - 1)
    return max(C_MIN, min(C_MAX, lg - 4))


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


def _device_key():
    return torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"


def save_table(choices, path=TABLE):
    """choices: {("g1" | "g2", log2 n): c}; merged into the table of this device."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    dev = data.setdefault(_device_key(), {})
    for (group, logn), c in choices.items():
        dev["%s_%d" % (group, logn)] = int(c)
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
        group, logn = key.split("_")
        api.set_msm_window(1 if group == "g1" else 2, 1 << int(logn), int(c))
        applied += 1
    return applied
