"""Continuous-batching decryption service over the fused weighted_btx and weighted_pfe entry
points.

Requests are batches of weighted_btx or weighted_pfe ciphertexts with the accepted shares of
one committee.  A weighted_pfe committee has one power-of-two batch size B; its requests are
completed to B slots (pad_pfe_batch) and all admitted requests of one committee run as one fused
pfe precompute + open.
The server keeps each committee's key material and prepared state (kernel line tables, opening
keys and the opening keys' line tables, per padded batch size) resident in HBM.  Whenever requests are waiting and HBM
budget and a launch slot are free, it admits the waiting requests that fit and runs all of them,
whatever their committees and batch sizes, through one fused precompute + open
(bte_btx_decrypt_requests): every stage is one launch over every admitted request.  A request
of B ciphertexts runs against the state prepared for min(nextpow2(B), L), the extra slots
being empty (the reference's None).  Up to `max_inflight` such launches overlap on the device,
each on its own stream; requests that arrive while others run are admitted as soon as budget
and a slot free up, and completion is observed with stream events.

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
_P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB


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
    valid: list = None         # B flags (None: all valid)
    padded: int = 0            # batch size of the prepared state it runs against (set by submit)

    def footprint(self):
        """HBM bytes held while in flight: padded inputs, Miller/FFT scratch and outputs."""
        bp = self.padded or next_pow2(self.batch)
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
    prepared: dict = field(default_factory=dict)  # padded B -> (lines, keys, key lines)

    def prepare(self, bp):
        if bp not in self.prepared:
            lines, keys = api.btx_prepare(self.material, self.total_weight, self.max_batch, self.offsets,
                                          self.accepted, bp)
            self.prepared[bp] = (lines, keys, api.prepare_g2_lines(keys))
        return self.prepared[bp]


@dataclass
class PfeCommittee:
    """A resident weighted_pfe committee: key material (d1 then d2), its power-of-two batch size B
    (weighted_pfe/src/decryption.rs:115-123: every batch has exactly B slots), alpha points,
    global key, and the prepared state (d1 by party and its line tables, total and mask lines)."""
    material: torch.Tensor
    total_weight: int
    batch_size: int
    offsets: list
    accepted: list
    alpha: list
    global_key: torch.Tensor
    state: tuple = None

    @property
    def max_batch(self):
        return self.batch_size

    def prepare(self, bp=None):
        if self.state is None:
            d1p, tl, ml = api.pfe_prepare(self.material, self.total_weight, self.batch_size, self.offsets,
                                          self.accepted, self.alpha, self.global_key)
            self.state = (d1p, tl, ml, api.prepare_g2_lines(d1p))
        return self.state


def pad_pfe_batch(ct1, ct2, batch_size):
    """A weighted_pfe request of b < B ciphertexts completed to the committee's B slots: the empty
    slots hold the identity in G1 and one in GT, so they add nothing to any party's share or to
    the Cauchy transforms of the real slots (the parties compute their shares over this padded
    batch); their messages are discarded."""
    b = ct1.numel() // G1A
    if b == batch_size:
        return ct1, ct2
    one = torch.zeros(F12, dtype=torch.uint8, device=ct2.device)
    one[:48] = torch.tensor(list(((1 << 384) % _P).to_bytes(48, "little")), dtype=torch.uint8, device=ct2.device)
    pad1 = torch.zeros((batch_size - b) * G1A, dtype=torch.uint8, device=ct1.device)
    pad2 = one.repeat(batch_size - b)
    return torch.cat([ct1, pad1]), torch.cat([ct2, pad2])


class BtxServer:
    def __init__(self, committees, hbm_budget=8 << 30, max_fuse=256, max_inflight=2):
        self.committees = committees
        self.budget = hbm_budget
        self.max_fuse = max_fuse
        self.max_inflight = max_inflight
        self.streams = [torch.cuda.Stream() for _ in range(max_inflight)]
        self.waiting = collections.deque()
        self.inflight = []   # (requests, event, bytes, stream index)
        self.finished = []
        self.launches = 0

    def submit(self, req: Request):
        """Queue a request.  Any 1 <= B <= L is accepted (validate_batch_size,
        weighted_btx/src/decryption.rs:711-722); it runs against the state prepared for
        min(nextpow2(B), L), whose FFT size nextpow2(2 B') is that of nextpow2(B)."""
        com = self.committees[req.committee]
        L = com.max_batch
        if req.batch < 1 or req.batch > L:
            raise ValueError("BatchTooLarge")
        if isinstance(com, PfeCommittee):
            if req.ct1.numel() != com.batch_size * G1A:
                raise ValueError("a weighted_pfe request carries the committee's B slots (pad_pfe_batch)")
            req.padded = com.batch_size
        else:
            req.padded = min(next_pow2(req.batch), L)
        self.waiting.append(req)

    def _used(self):
        return sum(b for _, _, b, _ in self.inflight)

    def _admit(self):
        admitted, used = [], self._used()
        keep = collections.deque()
        while self.waiting:
            r = self.waiting.popleft()
            need = r.footprint()
            if len(admitted) >= self.max_fuse or ((admitted or self.inflight) and used + need > self.budget):
                keep.append(r)
                continue
            admitted.append(r)
            used += need
        self.waiting = keep
        return admitted

    def _launch(self, reqs, stream):
        work, btx_reqs, pfe_groups = [], [], {}
        for r in reqs:
            com = self.committees[r.committee]
            if isinstance(com, PfeCommittee):
                pfe_groups.setdefault(r.committee, []).append(r)
                continue
            bp = r.padded
            lines, keys, key_lines = com.prepare(bp)
            btx_reqs.append(r)
            work.append({"ct1": r.ct1, "ct2": r.ct2, "valid": r.valid or [True] * r.batch, "batch_size": bp,
                         "lines": lines, "keys": keys, "key_lines": key_lines, "shares": r.shares,
                         "n_accepted": len(com.accepted)})
        for cid in pfe_groups:
            self.committees[cid].prepare()
        stream.wait_stream(torch.cuda.current_stream())  # inputs and prepared state
        with torch.cuda.stream(stream):
            if work:
                for r, o in zip(btx_reqs, api.btx_decrypt_requests(work)):
                    r.messages = o
            # weighted_pfe: every admitted request of one committee in one fused precompute + open
            for cid, group in pfe_groups.items():
                com = self.committees[cid]
                d1p, tl, ml, d1l = com.state
                B, T = com.batch_size, len(com.accepted)
                ct1 = torch.cat([r.ct1 for r in group])
                ct2 = torch.cat([r.ct2 for r in group])
                shares = torch.cat([r.shares for r in group])
                fc, pc = api.pfe_precompute_batched(ct1, tl, B)
                out = api.pfe_open_prepared_batched(shares, T, d1l, ml, ct1, ct2, fc, pc, com.alpha, B)
                for k, r in enumerate(group):
                    r.messages = out[k * B * F12:(k * B + r.batch) * F12]
            ev = torch.cuda.Event()
            ev.record(stream)
        return ev

    def poll(self, now_fn=time.perf_counter):
        """Retire finished launches; returns the requests completed."""
        done, keep = [], []
        for reqs, ev, b, si in self.inflight:
            if ev.query():
                t = now_fn()
                for r in reqs:
                    r.done = t
                done += reqs
            else:
                keep.append((reqs, ev, b, si))
        self.inflight = keep
        self.finished += done
        return done

    def step(self, now_fn=time.perf_counter, block=False):
        """Retire finished work, then launch the admissible waiting requests if a slot is free.
        With block=True waits for everything launched.  Returns the requests completed."""
        done = self.poll(now_fn)
        if self.waiting and len(self.inflight) < self.max_inflight:
            admitted = self._admit()
            if admitted:
                busy = {si for _, _, _, si in self.inflight}
                si = next(i for i in range(self.max_inflight) if i not in busy)
                t = now_fn()
                for r in admitted:
                    r.admitted = t
                ev = self._launch(admitted, self.streams[si])
                self.launches += 1
                self.inflight.append((admitted, ev, sum(r.footprint() for r in admitted), si))
        if block:
            for _, ev, _, _ in self.inflight:
                ev.synchronize()
            done += self.poll(now_fn)
        return done

    def drain(self, now_fn=time.perf_counter):
        done = []
        while self.waiting or self.inflight:
            done += self.step(now_fn, block=True)
        return done

    def decrypt_alone(self, req: Request):
        """The same request through the plain per-call path (no fusion; no padding for btx, the
        committee's fixed B for pfe)."""
        com = self.committees[req.committee]
        if isinstance(com, PfeCommittee):
            d1p, tl, ml, _ = com.prepare()
            fc, pc = api.pfe_precompute(req.ct1, tl)
            out = api.pfe_open(req.shares, d1p, ml, req.ct1, req.ct2, fc, pc, com.alpha)
            return out[:req.batch * F12]
        lines, keys = api.btx_prepare(com.material, com.total_weight, com.max_batch, com.offsets,
                                      com.accepted, req.batch)
        valid = req.valid or [True] * req.batch
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
    pending = collections.deque(zip(arrivals, requests))
    done = 0
    while done < len(requests):
        now = now_fn() - t0
        while pending and pending[0][0] <= now:
            at, r = pending.popleft()
            r.arrival = t0 + at
            server.submit(r)
        done += len(server.step(now_fn))
        if not server.waiting and not server.inflight and pending:
            sleep_fn(max(0.0, pending[0][0] - (now_fn() - t0)))
        elif server.inflight:
            sleep_fn(0.0005)
    wall = now_fn() - t0
    lat = sorted(r.done - r.arrival for r in requests)
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(0.99 * len(lat)))]
    total = sum(r.batch for r in requests)
    return p50, p99, total / wall, wall
