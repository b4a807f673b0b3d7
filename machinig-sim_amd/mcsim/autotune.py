"""Autotuner of the search kernel's launch schedule.

The search's makespan is set by two things: how many of the longest bisection trees run as
wide (four-wave) tasks, and how many persistent waves share the compute units.  The best
setting depends on the batch (its mix of trial lengths and tree depths), so it is picked by
timing candidates on the caller's own batch: a coarse pass over the wide-task threshold, then
a refinement around the best, each candidate timed on the device (HIP events around the
kernel) by the slowest of three launches.  The winner is installed with mc_set_search_tuning and returned with the table of
timings.  Results do not depend on the schedule; the tuner checks that every candidate
returns the same limits.
"""
import numpy as np

from . import lib


def _time(cells, opts, tuning, repeats=3, prior=None):
    """the slowest of `repeats` launches: the makespan depends on when each cell's check
    answers relative to its tree, so a schedule is judged by its bad runs, not its luckiest"""
    lib.set_search_tuning(**tuning)
    worst, res = None, None
    for _ in range(repeats):
        res = lib.prior_batch(cells, prior[0], prior[1], **opts) if prior is not None else lib.bisect_batch(cells, **opts)
        ms = lib.last_kernel_ms()
        worst = ms if worst is None else max(worst, ms)
    return worst, [(r.alim, r.lo, r.hi, r.trials) for r in res]


def tune_search(cells, opts, prior=None, fracs=(1e9, 0.6, 0.5, 0.4, 0.3), wide_max=(128, 256, 512),
                waves_per_cu=(8,), repeats=3):
    """time candidate schedules on this batch; install and return the fastest"""
    table = []
    ref = None
    for wpc in waves_per_cu:
        for f in fracs:
            t = dict(waves_per_cu=wpc, wide_max=max(wide_max), wide_frac=f, use_graph=1)
            ms, out = _time(cells, opts, t, repeats, prior)
            if ref is None:
                ref = out
            if out != ref:
                raise RuntimeError("search results changed with the launch schedule %r" % t)
            table.append((ms, t))
    best_ms, best = min(table, key=lambda x: x[0])
    if best["wide_frac"] < 1e9:
        for m in wide_max:
            if m == best["wide_max"]:
                continue
            t = dict(best, wide_max=m)
            ms, out = _time(cells, opts, t, repeats, prior)
            if out != ref:
                raise RuntimeError("search results changed with the launch schedule %r" % t)
            table.append((ms, t))
        best_ms, best = min(table, key=lambda x: x[0])
    lib.set_search_tuning(**best)
    return best, best_ms, [dict(t, kernel_ms=ms) for ms, t in table]
