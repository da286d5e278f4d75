"""Autotuner: schedule and layout choices timed on the workload at hand.

Nothing here changes results beyond rounding, only speed, so every choice is
made by timing real solves.

Run-time knobs (``autotune``, on an existing plan; gaphip_tuning):
  * grid schedule: wavefronts per workgroup for ordinary and for big networks,
    and the size (front slots + factor entries) from which a network is big;
  * level schedule: whether the Newton step is replayed from a captured HIP
    graph, and for the large fronts (> 128 slots) the outer panel width of the
    two-level factorisation and the front size from which it applies.

Plan-time knobs (``autotune_plan``, one plan per candidate; gaphip_plan_options):
  * the relaxed supernode amalgamation thresholds (per-supernode fronts: how
    many explicit zeros a merge of a child into its parent may add);
  * the grid schedule's task layout: the smallest lanes-per-front class (the
    per-front lane mapping) and whether smaller fronts fill the free lane groups
    of a larger class's last task.
"""
from __future__ import annotations

import itertools
import time

from .capi import Plan

DEFAULT_CANDIDATES = {
    "grid_waves": (1, 2, 4),
    "grid_big_waves": (4, 8),
    "grid_big_work": (-1, 20_000, 40_000, 80_000, 160_000),
    "level_graph": (0, 1),
    "big_outer": (128, 256, 512),
    "big_two_level_min": (2048, 4096, 8192),
}

DEFAULT_PLAN_CANDIDATES = (
    {},
    {"grid_pack": 0},
    {"grid_min_class": 1},
    {"relax_abs": 2, "relax_frac": 0.05},
    {"relax_abs": 6, "relax_frac": 0.10},
)


def _time(plan: Plan, tolerance, max_iterations, base_power, repeats):
    plan.solve(tolerance, max_iterations, base_power, fetch_v=False)
    t0 = time.perf_counter()
    for _ in range(repeats):
        plan.solve(tolerance, max_iterations, base_power, fetch_v=False)
    return 1e3 * (time.perf_counter() - t0) / repeats


def _grid_candidates(cand):
    seen = set()
    for gw, bw, work in itertools.product(cand["grid_waves"], cand["grid_big_waves"], cand["grid_big_work"]):
        if work < 0:
            bw = gw  # nobody is big: the big-network shape is irrelevant
        if (gw, bw, work) in seen:
            continue
        seen.add((gw, bw, work))
        yield {"grid_waves": gw, "grid_big_waves": bw, "grid_big_work": work}


def _level_candidates(cand, max_front):
    if max_front <= 128:  # no blocked fronts: only the launch mode matters
        for g in cand["level_graph"]:
            yield {"level_graph": g}
        return
    seen = set()
    for g, outer, lim in itertools.product(cand["level_graph"], cand["big_outer"], cand["big_two_level_min"]):
        if lim > max_front:
            outer, lim = 256, 1 << 30  # no front is two-level: the width is irrelevant
        if (g, outer, lim) in seen:
            continue
        seen.add((g, outer, lim))
        yield {"level_graph": g, "big_outer": outer, "big_two_level_min": lim}


def autotune(plan: Plan, tolerance: float = 1e-8, max_iterations: int = 100, base_power: float = 100e6,
             candidates: dict | None = None, repeats: int = 2) -> dict:
    """Time every candidate of the plan's schedule with `repeats` solves (after
    one warm-up solve) and leave the fastest on the plan.  Returns
    {"tuning": the plan's tuning, "ms": its time per solve, "table": [(knobs, ms), ...]}."""
    cand = dict(DEFAULT_CANDIDATES)
    cand.update(candidates or {})
    if plan.schedule == "grid":
        cands = list(_grid_candidates(cand))
    else:
        cands = list(_level_candidates(cand, int(plan.stats()["max_front"])))
    table = []
    for knobs in cands:
        plan.set_tuning(**knobs)
        table.append((knobs, _time(plan, tolerance, max_iterations, base_power, repeats)))
    best, ms = min(table, key=lambda r: r[1])
    plan.set_tuning(**best)
    return {"tuning": plan.tuning(), "ms": ms, "table": table}


def autotune_plan(nets, precision="mixed", schedule="auto", tolerance: float = 1e-8, max_iterations: int = 100,
                  base_power: float = 100e6, plan_candidates=DEFAULT_PLAN_CANDIDATES, candidates: dict | None = None,
                  repeats: int = 2) -> dict:
    """Build one plan per plan-time candidate (gaphip_plan_options), tune each
    with ``autotune``, then time every plan once more with its best knobs in
    the reverse order (so that a drift of the device clock over the run does
    not favour the candidates timed first) and keep the plan with the lowest
    mean of the two timings.  Returns {"plan": the tuned plan, "options": its
    options, "tuning": its run-time knobs, "ms": its mean time per solve,
    "table": [(options, best run-time knobs, ms), ...]}; the other plans are
    released."""
    plans = []
    for opts in plan_candidates:
        plan = Plan(nets, precision, schedule=schedule, options=opts)
        r = autotune(plan, tolerance, max_iterations, base_power, candidates, repeats)
        plans.append([plan, r["ms"]])
    for entry in reversed(plans):
        entry[1] = 0.5 * (entry[1] + _time(entry[0], tolerance, max_iterations, base_power, repeats))
    table = [(p.options, p.tuning(), ms) for p, ms in plans]
    best = min(plans, key=lambda e: e[1])
    for p, _ in plans:
        if p is not best[0]:
            p.close()
    plan, ms = best
    return {"plan": plan, "options": plan.options, "tuning": plan.tuning(), "ms": ms, "table": table}


__all__ = ["autotune", "autotune_plan", "DEFAULT_CANDIDATES", "DEFAULT_PLAN_CANDIDATES"]
