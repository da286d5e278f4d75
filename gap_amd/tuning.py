"""Autotuner for the launch shape of the grid schedule.

The grid schedule runs one workgroup per network.  Its launch shape
(wavefronts per workgroup for ordinary networks, for big networks, and the
size from which a network counts as big) changes only the speed, never the
results, so it can be chosen by timing real solves on the plan at hand.
"""
from __future__ import annotations

import itertools
import time

from .capi import Plan

DEFAULT_CANDIDATES = {
    "grid_waves": (1, 2, 4),
    "grid_big_waves": (4, 8),
    "grid_big_work": (-1, 20_000, 40_000, 80_000, 160_000),
}


def autotune(plan: Plan, tolerance: float = 1e-8, max_iterations: int = 100, base_power: float = 100e6,
             candidates: dict | None = None, repeats: int = 2) -> dict:
    """Time every candidate tuning with `repeats` solves (after one warm-up
    solve) and leave the fastest set on the plan.  Returns
    {"tuning": best, "ms": its time per solve, "table": [(tuning, ms), ...]}.
    Plans on the level schedule are returned unchanged."""
    if plan.schedule != "grid":
        return {"tuning": plan.tuning(), "ms": None, "table": []}
    cand = dict(DEFAULT_CANDIDATES)
    cand.update(candidates or {})
    table = []
    seen = set()
    for gw, bw, work in itertools.product(cand["grid_waves"], cand["grid_big_waves"], cand["grid_big_work"]):
        if work < 0:
            bw = gw  # nobody is big: the big-network shape is irrelevant
        key = (gw, bw, work)
        if key in seen:
            continue
        seen.add(key)
        plan.set_tuning(gw, bw, work)
        plan.solve(tolerance, max_iterations, base_power, fetch_v=False)
        t0 = time.perf_counter()
        for _ in range(repeats):
            plan.solve(tolerance, max_iterations, base_power, fetch_v=False)
        table.append(({"grid_waves": gw, "grid_big_waves": bw, "grid_big_work": work},
                      1e3 * (time.perf_counter() - t0) / repeats))
    best, ms = min(table, key=lambda r: r[1])
    plan.set_tuning(**best)
    return {"tuning": best, "ms": ms, "table": table}


__all__ = ["autotune", "DEFAULT_CANDIDATES"]
