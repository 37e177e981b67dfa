"""Launch-shape autotuner of the step (sdm_tuning in include/sdm.h).

The launch shapes do not change any result (they only schedule the same per-interaction work
differently), so they can be chosen by timing a workload's own steps: the context's state is
snapshotted, each candidate runs `steps` steps from that snapshot, and the fastest is kept.
The search is greedy, one knob at a time in the order of their expected effect:
  graph                 steady-state steps replayed from a captured hipGraph, or launched
  solver_blocks_per_cu  grid of the persistent solvers (0 = the occupancy maximum)
  sort_by_cost          Nelder-Mead list sorted by expected cost (longest solves first)
  misc_blocks_per_cu    grid of the wall / sphere kernel
"""
from __future__ import annotations

import time

SEARCH = (
    ("graph", (1, 0)),
    ("solver_blocks_per_cu", (0, 1, 2, 4)),
    ("sort_by_cost", (1, 0)),
    ("misc_blocks_per_cu", (2, 1, 4)),
)


def _time(ctx, snap, tuning, steps):
    ctx.restore(snap)
    ctx.set_tuning(**tuning)
    ctx.step(2)  # capture / first launches outside the timed steps
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    return (time.perf_counter() - t0) / steps


def autotune(ctx, steps=20, log=None):
    """Pick the fastest launch shapes for this context's workload; leaves the context in the
    state it had on entry, with the chosen tuning set.  Returns (tuning, trials)."""
    snap = ctx.snapshot()
    best = dict(ctx.tuning())
    best_t = _time(ctx, snap, best, steps)
    trials = [(dict(best), best_t)]
    for knob, values in SEARCH:
        for v in values:
            if v == best[knob]:
                continue
            cand = dict(best, **{knob: v})
            t = _time(ctx, snap, cand, steps)
            trials.append((cand, t))
            if log:
                log("autotune %s=%d: %.3f ms/step (best %.3f)" % (knob, v, t * 1e3, best_t * 1e3))
            if t < best_t:
                best, best_t = cand, t
    ctx.restore(snap)
    ctx.set_tuning(**best)
    return best, trials
