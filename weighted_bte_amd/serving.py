"""Continuous-batching decryption service over the fused weighted_btx entry points.

Requests are batches of weighted_btx ciphertexts with the accepted shares of one committee.
The server keeps each committee's key material and prepared state (kernel line tables and
opening keys, per padded batch size) resident in HBM.  Every scheduling round it admits the
waiting requests that fit the HBM budget, groups the admitted ones by (committee, padded batch
size) and runs each group through one fused precompute + open.  A request of B ciphertexts is
padded to the next power of two with invalid slots (identity first component), which the
scheme opens to None; the padding is dropped from the result.  Requests that arrive while a
round runs join the next round.

This is host control flow (the reference has no service layer); all arithmetic runs in the
library's kernels.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import torch

from . import api

G1A, G1J, F12 = 96, 144, 576


def next_pow2(b):
    p = 1
    while p < b:
        p <<= 1
    return p


@dataclass
class Request:
    rid: int
    committee: int
    batch: int                 # B ciphertexts
    ct1: torch.Tensor          # B x G1 affine (device)
    ct2: torch.Tensor          # B x Fp12 (device)
    shares: torch.Tensor       # T x G1 Jacobian (device), committee order
    arrival: float = 0.0
    admitted: float = 0.0
    done: float = 0.0
    messages: torch.Tensor = None

    def footprint(self):
        """HBM bytes held while in flight: padded inputs, Miller/FFT scratch and outputs."""
        bp = next_pow2(self.batch)
        m = 2 * bp
        t = self.shares.numel() // G1J
        return bp * (G1A + 2 * F12) + m * (G1J + G1A + 3 * F12) + bp * t * (G1A + 192 + F12)


@dataclass
class Committee:
    material: torch.Tensor     # device G2 affine: negative blocks then positive blocks
    total_weight: int
    max_batch: int
    offsets: list
    accepted: list
    prepared: dict = field(default_factory=dict)  # padded B -> (lines, keys)

    def prepare(self, bp):
        if bp not in self.prepared:
            self.prepared[bp] = api.btx_prepare(self.material, self.total_weight, self.max_batch,
                                                self.offsets, self.accepted, bp)
        return self.prepared[bp]


class BtxServer:
    def __init__(self, committees, hbm_budget=8 << 30, max_fuse=128):
        self.committees = committees
        self.budget = hbm_budget
        self.max_fuse = max_fuse
        self.waiting = collections.deque()
        self.finished = []

    def submit(self, req: Request):
        if next_pow2(req.batch) > self.committees[req.committee].max_batch:
            raise ValueError("BatchTooLarge")
        self.waiting.append(req)

    def _admit(self):
        admitted, used = [], 0
        keep = collections.deque()
        while self.waiting:
            r = self.waiting.popleft()
            need = r.footprint()
            if admitted and used + need > self.budget:
                keep.append(r)
                continue
            admitted.append(r)
            used += need
        self.waiting = keep
        return admitted

    def _run_group(self, com, bp, reqs):
        lines, keys = com.prepare(bp)
        T = len(com.accepted)
        ct1 = torch.zeros(len(reqs) * bp * G1A, dtype=torch.uint8, device="cuda")
        ct2 = torch.zeros(len(reqs) * bp * F12, dtype=torch.uint8, device="cuda")
        valid = []
        for k, r in enumerate(reqs):
            ct1[k * bp * G1A:k * bp * G1A + r.batch * G1A] = r.ct1
            ct2[k * bp * F12:k * bp * F12 + r.batch * F12] = r.ct2
            valid += [True] * r.batch + [False] * (bp - r.batch)
        shares = torch.cat([r.shares for r in reqs])
        beta = api.btx_precompute_batched(ct1, valid, lines, bp)
        out = api.btx_open_batched(shares, T, keys, ct2, beta, valid, bp)
        for k, r in enumerate(reqs):
            r.messages = out[k * bp * F12:k * bp * F12 + r.batch * F12]

    def step(self, now_fn=time.perf_counter):
        """One scheduling round: admit, group, run; returns the requests completed."""
        admitted = self._admit()
        if not admitted:
            return []
        t_adm = now_fn()
        groups = collections.OrderedDict()
        for r in admitted:
            r.admitted = t_adm
            groups.setdefault((r.committee, next_pow2(r.batch)), []).append(r)
        for (cid, bp), reqs in groups.items():
            for i in range(0, len(reqs), self.max_fuse):
                self._run_group(self.committees[cid], bp, reqs[i:i + self.max_fuse])
        torch.cuda.synchronize()
        t = now_fn()
        for r in admitted:
            r.done = t
        self.finished += admitted
        return admitted

    def decrypt_alone(self, req: Request):
        """The same request through the plain per-call path (no fusion, no padding)."""
        com = self.committees[req.committee]
        lines, keys = api.btx_prepare(com.material, com.total_weight, com.max_batch, com.offsets,
                                      com.accepted, req.batch)
        valid = [True] * req.batch
        beta = api.btx_precompute(req.ct1, valid, lines)
        return api.btx_open(req.shares, keys, req.ct2, beta, valid)


def run_poisson(server, requests, rate, now_fn=time.perf_counter, sleep_fn=time.sleep):
    """Feeds `requests` at Poisson arrival times (mean `rate` requests/s) into the server and runs
    rounds until all are done.  Returns (p50 latency s, p99 latency s, ctxts/s, wall s)."""
    import random

    rnd = random.Random(17)
    t0 = now_fn()
    arrivals, t = [], 0.0
    for _ in requests:
        t += rnd.expovariate(rate)
        arrivals.append(t)
    pending = list(zip(arrivals, requests))
    done = 0
    while done < len(requests):
        now = now_fn() - t0
        while pending and pending[0][0] <= now:
            at, r = pending.pop(0)
            r.arrival = t0 + at
            server.submit(r)
        if not server.waiting:
            if pending:
                sleep_fn(max(0.0, pending[0][0] - (now_fn() - t0)))
            continue
        done += len(server.step(now_fn))
    wall = now_fn() - t0
    lat = sorted(r.done - r.arrival for r in requests)
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(0.99 * len(lat)))]
    total = sum(r.batch for r in requests)
    return p50, p99, total / wall, wall
