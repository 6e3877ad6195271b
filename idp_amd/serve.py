"""Continuous batching of animation-fix frames under a device-memory budget.

No counterpart in the reference, which advances one mesh per call
(Advance_One_Step_IE_Discrete_Shell, Library/FEM/Shell/IMPLICIT_EULER.h:152-891).

Slots.  A slot is one device-resident mesh of a given kind (a rest scene:
topology, rest positions, material, mass).  Every sequence of a kind has the
same mesh, so a frame of any sequence runs in any free slot of its kind; the
request brings the sequence's current positions, velocity (toward the
frame's target) and barrier stiffness, and the slot is reused by the next
request when the frame is done.

Admission.  The device footprint of one slot of each kind is measured on the
device: the library's device bytes (idp_device_memory) held by a one-slot
context after a two-iteration solve, when its buffers have grown to their
working sizes.  Slots are planned so that their summed footprint fits the
budget (at least one per kind, then one at a time to the kind with the most
demand per slot).  A request waits in its kind's FIFO queue until a slot of
that kind is free; it is then started between Newton iterations of the
others (idp_serve_start), and every idp_serve_iterate call advances all busy
slots by one Newton iteration.  A frame's result does not depend on the
slot, on what else is running, or on how long it waited: each instance's
arithmetic is the same alone and in any mix.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import capi, driver


def slot_footprint(scene, pcg_rtol=1e-8):
    """Device bytes one slot of `scene`'s kind holds once its buffers have
    reached their working sizes (a one-slot context, a two-iteration solve)."""
    before = capi.device_memory()[0]
    b = driver.BatchAnimFix([scene], pcg_rtol=pcg_rtol)
    try:
        b.advance_one_time_step(max_newton=2)
        return capi.device_memory()[0] - before
    finally:
        b.close()


def plan_slots(footprints, demand, budget):
    """Slots per kind: one each, then one at a time to the kind with the most
    demand per slot, while the summed footprint fits the budget."""
    kinds = list(footprints)
    n = {k: 1 for k in kinds}
    used = sum(footprints[k] for k in kinds)
    if used > budget:
        raise ValueError(f"budget {budget} B holds less than one slot of each kind ({used} B)")
    while True:
        cand = [k for k in kinds if used + footprints[k] <= budget and n[k] < demand.get(k, 0)]
        if not cand:
            return n
        k = max(cand, key=lambda q: demand.get(q, 0) / n[q])
        n[k] += 1
        used += footprints[k]


@dataclass
class Request:
    tag: object
    kind: str
    X: np.ndarray
    v: np.ndarray
    kappa3: np.ndarray
    submitted_iter: int = 0
    submitted_t: float = 0.0
    started_iter: int = -1
    slot: int = -1


@dataclass
class Done:
    tag: object
    X: np.ndarray
    v: np.ndarray
    kappa3: np.ndarray
    stats: dict
    slot: int
    waited_iters: int
    latency_s: float


class FrameServer:
    """Slots for each kind of mesh, planned from a device-memory budget, in one
    batched context; FIFO admission per kind.  kinds: {name: rest Scene};
    demand: {name: expected concurrent sequences} (the planner's weights);
    slots: {name: count} to skip planning."""

    def __init__(self, kinds, budget_bytes, demand=None, slots=None, pcg_rtol=1e-8, device=0):
        self.kinds = dict(kinds)
        demand = demand or {k: 1 for k in self.kinds}
        self.footprint = {k: slot_footprint(sc, pcg_rtol) for k, sc in self.kinds.items()}
        self.slots_per_kind = dict(slots) if slots else plan_slots(self.footprint, demand, budget_bytes)
        self.budget = budget_bytes
        self.slot_kind = [k for k in self.kinds for _ in range(self.slots_per_kind[k])]
        before = capi.device_memory()[0]
        self.batch = driver.BatchAnimFix([self.kinds[k] for k in self.slot_kind], device=device, pcg_rtol=pcg_rtol)
        self.ctx = self.batch.ctx
        self.ctx.serve_configure(self.batch.h, pcg_rtol=pcg_rtol)
        self.resident_bytes = capi.device_memory()[0] - before
        self.free = {k: deque(i for i, q in enumerate(self.slot_kind) if q == k) for k in self.kinds}
        self.queue = {k: deque() for k in self.kinds}
        self.running = {}
        self.iters = 0
        self.slot_users = [set() for _ in self.slot_kind]

    def planned_bytes(self):
        return sum(self.footprint[k] * n for k, n in self.slots_per_kind.items())

    def submit(self, tag, kind, X, v, kappa3):
        r = Request(tag, kind, np.ascontiguousarray(X, np.float64), np.ascontiguousarray(v, np.float64),
                    np.asarray(kappa3, np.float64).copy(), self.iters, time.perf_counter())
        self.queue[kind].append(r)
        return r

    def _admit(self):
        start = []
        for k in self.kinds:
            while self.queue[k] and self.free[k]:
                r = self.queue[k].popleft()
                r.slot = self.free[k].popleft()
                r.started_iter = self.iters
                start.append(r)
        if start:
            ids = [r.slot for r in start]
            b = self.batch
            self.ctx.serve_start(ids, [r.X for r in start], [r.v for r in start], b.tol[ids], b.dHat2[ids],
                                 np.stack([r.kappa3 for r in start]))
            for r in start:
                self.running[r.slot] = r
                self.slot_users[r.slot].add(r.tag)

    def pending(self):
        return sum(len(q) for q in self.queue.values()) + len(self.running)

    def iterate(self):
        """Admit what fits, advance every busy slot by one Newton iteration;
        returns the frames that finished."""
        self._admit()
        if not self.running:
            return []
        n, fin = self.ctx.serve_iterate()
        self.iters += 1
        now = time.perf_counter()
        out = []
        for s in np.nonzero(fin)[0]:
            s = int(s)
            r = self.running.pop(s)
            X, v, k3, st = self.ctx.serve_result(s)
            out.append(Done(r.tag, X, v, k3, st, s, r.started_iter - r.submitted_iter, now - r.submitted_t))
            self.free[r.kind].append(s)
        return out

    def close(self):
        self.batch.close()
