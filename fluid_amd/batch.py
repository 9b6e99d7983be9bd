"""Batched PBF: many independent scenes of uneven particle counts.

A plan assigns every scene to a group; each group is one multi-scene
context with one schedule (``flat``, ``graph`` or ``persistent`` with a
workgroup size).  Groups run concurrently on their own HIP streams.  The
autotuner times candidate plans on the actual scenes and keeps the fastest.

No schedule changes the arithmetic: every scene gives the same bits as the
plain single-scene path (the parity tests hold this).  The grid cell size is
therefore not a tuning knob here -- it decides the neighbour summation order.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from .params import Params
from .pbf import Particles


@dataclasses.dataclass(frozen=True)
class Group:
    schedule: str  # flat | graph | persistent
    block: int = 256
    max_count: int | None = None  # scenes with count <= max_count (None: all remaining)


@dataclasses.dataclass(frozen=True)
class Plan:
    groups: tuple

    def describe(self) -> str:
        parts = []
        for g in self.groups:
            s = g.schedule + (f"/{g.block}" if g.schedule == "persistent" else "")
            parts.append(s + (f"(n<={g.max_count})" if g.max_count else ""))
        return " + ".join(parts)


class BatchRunner:
    """Runs a list of scenes [(x, v, box_lo, box_hi)] with a plan."""

    def __init__(self, params: Params, items, plan: Plan):
        self.params = params
        self.plan = plan
        self.items = items
        counts = np.array([len(it[0]) for it in items])
        left = np.ones(len(items), bool)
        self.groups = []  # (ctx, scene indices)
        for g in plan.groups:
            sel = left if g.max_count is None else (left & (counts <= g.max_count))
            idx = np.nonzero(sel)[0]
            left &= ~sel
            if len(idx) == 0:
                continue
            spec = [(counts[i], items[i][2], items[i][3]) for i in idx]
            ctx = Particles(params, scenes=spec)
            ctx.set_schedule(g.schedule, g.block)
            ctx.load(np.concatenate([items[i][0] for i in idx]),
                     np.concatenate([items[i][1] for i in idx]))
            self.groups.append((ctx, idx))
        assert not left.any(), "plan leaves scenes unassigned"

    def step(self, nsteps: int) -> None:
        for ctx, _ in self.groups:
            ctx.step_async(nsteps)
        for ctx, _ in self.groups:
            ctx.sync()

    def read(self):
        """Per-scene (x, v, rho) in the original scene order."""
        out = [None] * len(self.items)
        for ctx, idx in self.groups:
            x, v, rho = ctx.read()
            off = 0
            for i in idx:
                n = len(self.items[i][0])
                out[i] = (x[off:off + n], v[off:off + n], rho[off:off + n])
                off += n
        return out

    def close(self) -> None:
        for ctx, _ in self.groups:
            ctx.close()
        self.groups = []


def candidate_plans(counts) -> list:
    plans = [Plan((Group("flat"),)), Plan((Group("graph"),))]
    for b in (256, 512, 1024):
        plans.append(Plan((Group("persistent", b),)))
    for thr in (4096, 8192, 16384):
        if (counts <= thr).any() and (counts > thr).any():
            for b in (256, 512):
                plans.append(Plan((Group("persistent", b, thr), Group("graph"))))
                plans.append(Plan((Group("persistent", b, thr), Group("flat"))))
            # small and large scenes as two flat contexts on concurrent streams
            plans.append(Plan((Group("flat", 256, thr), Group("flat"))))
    return plans


def autotune(params: Params, items, steps: int = 3, warmup: int = 1, log=None):
    """Times every candidate plan on the scenes; returns (best plan, table)."""
    counts = np.array([len(it[0]) for it in items])
    table = []
    for plan in candidate_plans(counts):
        r = BatchRunner(params, items, plan)
        r.step(warmup)
        t0 = time.perf_counter()
        r.step(steps)
        dt = time.perf_counter() - t0
        r.close()
        rate = len(items) * steps / dt
        table.append((plan, rate))
        if log:
            log(f"autotune {plan.describe():40s} {rate:12.1f} scene-steps/s")
    best = max(table, key=lambda t: t[1])[0]
    return best, table


def run_looped(params: Params, items, steps: int):
    """The plain single-scene path looped over the scenes; returns the wall
    time of the stepping only and the final per-scene states."""
    ctxs = []
    for x, v, lo, hi in items:
        c = Particles(params.replace(box_lo=lo, box_hi=hi))
        c.load(x, v)
        ctxs.append(c)
    t0 = time.perf_counter()
    for c in ctxs:
        c.time_step(nsteps=steps)
    wall = time.perf_counter() - t0
    states = [c.read() for c in ctxs]
    for c in ctxs:
        c.close()
    return wall, states
