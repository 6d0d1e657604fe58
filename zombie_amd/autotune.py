"""Launch-shape autotuner for the persistent walk kernel.

Tunes, on a sample of the caller's points and with the walk kernel's own HIP-event time:
  - BVH width {4, 8} and leaf size {1, 2, 4} (scene layout),
  - walks per lane (claim size) and the regroup interval (ballot/refill period),
  - the certified distance tolerance of the closest-point query.
The objective is completed walks per second: the number of walks needed for a target
error is fixed by the estimator's variance, so walks/s (not raw steps/s, which a looser
radius inflates) is what shortens the time to a given RMSE.  Coordinate descent: layout
first, then tolerance, then the launch shape.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

from . import engine as E


@dataclass
class TuneResult:
    bvh_width: int
    leaf_size: int
    walks_per_lane: int
    regroup_interval: int
    step_interval: int
    distance_tolerance: float
    walks_per_s: float
    steps_per_s: float
    table: list = field(default_factory=list)


def autotune(problem, points: E.DevicePoints, scenes: dict, widths=(4, 8), leaf_sizes=(1, 2, 4),
             walks_per_lane=(1, 4), regroups=(4, 8, 16, 32), step_intervals=(2, 4, 8, 16),
             tolerances=(0.0, 0.1, 0.2, 0.3),
             sample: int = 65536, walks: int = 64, seed: int = 1234, solve=None) -> TuneResult:
    """scenes: (width, leaf) -> E.Scene built from `problem` (built on demand when missing)."""
    import torch
    solve = solve or E.wos_solve
    n = min(sample, points.n)
    sub = E.DevicePoints(pt=points.pt[:n].contiguous(), dist_to_absorbing=points.dist_to_absorbing[:n].contiguous(),
                         n_walks=torch.full((n,), walks, dtype=torch.int32, device=points.pt.device),
                         n_walks_max=walks,
                         type=None if points.type is None else points.type[:n].contiguous(),
                         normal=None if points.normal is None else points.normal[:n].contiguous())
    table = []

    def scene_for(width, leaf):
        key = (width, leaf)
        if key not in scenes:
            scenes[key] = E.Scene(problem.positions, problem.indices, problem.is_reflecting, problem.domain_min,
                                  problem.domain_max, bvh_width=width, leaf_size=leaf,
                                  solve_double_sided=problem.solve_double_sided)
        return scenes[key]

    def run(width, leaf, wpl, rg, si, tol):
        sc = scene_for(width, leaf)
        opts = E.SolveOptions(seed=seed, walks_per_lane=wpl, regroup_interval=rg, step_interval=si,
                              distance_tolerance=tol)
        rep = solve(sc, problem.pde, problem.settings, sub, E.DeviceStats.zeros(n), opts)
        t = max(rep.kernel_ms * 1e-3, 1e-9)
        row = {"bvh_width": width, "leaf_size": leaf, "walks_per_lane": wpl, "regroup_interval": rg,
               "step_interval": si, "distance_tolerance": tol, "walks_per_s": rep.total_walks / t, "steps_per_s": rep.total_steps / t,
               "mean_walk_length": rep.total_steps / max(rep.total_walks, 1), "kernel_ms": rep.kernel_ms}
        table.append(row)
        return row

    best = {"bvh_width": widths[0], "leaf_size": leaf_sizes[-1], "walks_per_lane": walks_per_lane[0],
            "regroup_interval": regroups[min(1, len(regroups) - 1)], "step_interval": step_intervals[min(2, len(step_intervals) - 1)],
            "distance_tolerance": tolerances[min(2, len(tolerances) - 1)]}
    # warm every kernel variant once
    for width in widths:
        run(width, leaf_sizes[-1], 1, 16, 4, 0.0)
    table.clear()

    def cfg(**kw):
        c = dict(best)
        c.update(kw)
        return run(c["bvh_width"], c["leaf_size"], c["walks_per_lane"], c["regroup_interval"], c["step_interval"],
                   c["distance_tolerance"])

    def better(row, cur):
        return cur is None or row["walks_per_s"] > cur["walks_per_s"]

    cur = None
    for width, leaf in itertools.product(widths, leaf_sizes):
        row = cfg(bvh_width=width, leaf_size=leaf)
        if better(row, cur):
            cur = row
    best.update(bvh_width=cur["bvh_width"], leaf_size=cur["leaf_size"])
    for tol in tolerances:
        row = cfg(distance_tolerance=tol)
        if better(row, cur):
            cur = row
    best.update(distance_tolerance=cur["distance_tolerance"])
    for si in step_intervals:
        row = cfg(step_interval=si)
        if better(row, cur):
            cur = row
    best.update(step_interval=cur["step_interval"])
    for wpl, rg in itertools.product(walks_per_lane, regroups):
        row = cfg(walks_per_lane=wpl, regroup_interval=rg)
        if better(row, cur):
            cur = row
    return TuneResult(cur["bvh_width"], cur["leaf_size"], cur["walks_per_lane"], cur["regroup_interval"],
                      cur["step_interval"], cur["distance_tolerance"], cur["walks_per_s"], cur["steps_per_s"], table)


def autotune_batch(problems, points: E.DevicePoints, instance_of_point, leaf_size: int = 2, widths=(4, 8),
                   walks_per_lane=(1, 2, 4), regroups=(8, 16, 32), step_interval: int = 4,
                   distance_tolerance: float = 0.1, sample: int = 65536, walks: int = 32, seed: int = 4321,
                   batches: dict | None = None) -> TuneResult:
    """Exhaustive search over BVH width x walks per lane x regroup interval for a batch of
    instances (one E.Problem each) solved by one launch.  `batches` caches width -> (scenes,
    E.Batch) across calls.  The sample is the first `sample` points of the concatenation."""
    import torch
    batches = {} if batches is None else batches
    n = min(sample, points.n)
    sub = E.DevicePoints(pt=points.pt[:n].contiguous(), dist_to_absorbing=points.dist_to_absorbing[:n].contiguous(),
                         n_walks=torch.full((n,), walks, dtype=torch.int32, device=points.pt.device),
                         n_walks_max=walks)
    inst = instance_of_point[:n].contiguous()

    def batch_for(width):
        if width not in batches:
            scenes = [E.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max, bvh_width=width,
                              leaf_size=leaf_size) for p in problems]
            batches[width] = (scenes, E.Batch(scenes, [p.pde for p in problems]))
        return batches[width][1]

    table = []
    settings = problems[0].settings
    for width in widths:   # warm every kernel variant once
        E.batch_wos_solve(batch_for(width), settings, sub, inst, E.DeviceStats.zeros(n), E.SolveOptions(seed=seed))
    best = None
    for width, wpl, rg in itertools.product(widths, walks_per_lane, regroups):
        opts = E.SolveOptions(seed=seed, walks_per_lane=wpl, regroup_interval=rg, step_interval=step_interval,
                              distance_tolerance=distance_tolerance)
        rep = E.batch_wos_solve(batch_for(width), settings, sub, inst, E.DeviceStats.zeros(n), opts)
        t = max(rep.kernel_ms * 1e-3, 1e-9)
        row = {"bvh_width": width, "leaf_size": leaf_size, "walks_per_lane": wpl, "regroup_interval": rg,
               "step_interval": step_interval, "distance_tolerance": distance_tolerance,
               "walks_per_s": rep.total_walks / t, "steps_per_s": rep.total_steps / t, "kernel_ms": rep.kernel_ms}
        table.append(row)
        if best is None or row["walks_per_s"] > best["walks_per_s"]:
            best = row
    return TuneResult(best["bvh_width"], leaf_size, best["walks_per_lane"], best["regroup_interval"], step_interval,
                      distance_tolerance, best["walks_per_s"], best["steps_per_s"], table)
