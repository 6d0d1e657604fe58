#!/usr/bin/env python3
"""Benchmark: 3D Poisson Walk-on-Spheres on one MI355X, plain and with the learned control variate.

Workload (the headline configuration): icosphere level 6 (81,920 triangles) radially
displaced by real spherical harmonics l <= 4, u = sin 2x sin 2y sin 2z, f = 12 u,
Dirichlet g = u, 1M interior points, 256 walks per point per pass, epsilon 1e-4,
FP32 geometry, FP64 per-point accumulation.  A "step" is one 256-walk pass over all
points (one zg_wos_solve call; inputs resident in HBM).

Phases: setup (mesh, 1M interior points, distances) -> launch-shape autotune -> instrumented
sample (BVH visits per step, exact-radius walk length) -> timed plain passes -> learned
control variate (2M training pairs of 32-walk estimates, 5k Adam steps at batch 65,536, then
the same timed passes with the variate) -> CPU baseline (the oracle on the host cores, on a
bounded sample).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "s to RMSE 1e-3 vs analytic, 1M pts: plain vs learned-CV WoS; walk-steps/s"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
TARGET_RMSE = 1e-3
T_START = time.time()


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sphere3d", choices=["sphere3d", "star2d", "mixed2d", "batch"],
                    help="sphere3d: the headline 3D Poisson workload (default); star2d: 2D Laplace on the star "
                         "domain; mixed2d: 2D walk on stars with Neumann slits; batch: 256 screened-Poisson "
                         "instances in one graph-captured launch")
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--walks", type=int, default=None, help="walks per point per pass (workload default)")
    ap.add_argument("--instances", type=int, default=256, help="batch workload: number of instances")
    ap.add_argument("--chunk-walks", type=int, default=128, help="mixed2d: walks per point per solve call")
    ap.add_argument("--wost-interleave", action="store_true",
                    help="mixed2d: the interleaved walk-on-stars kernel (every lane's query advances one BVH "
                         "item per iteration) instead of the default run-to-completion kernel")
    ap.add_argument("--wost-step-interval", type=int, default=8,
                    help="mixed2d with --wost-interleave: iterations between the step logic runs")
    ap.add_argument("--bvh-width", type=int, default=4)
    ap.add_argument("--leaf-size", type=int, default=2)
    ap.add_argument("--walks-per-lane", type=int, default=1)
    ap.add_argument("--regroup", type=int, default=32)
    ap.add_argument("--step-interval", type=int, default=16)
    ap.add_argument("--tolerance", type=float, default=0.2, help="distance tolerance when not autotuning")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--tune", action="store_true",
                    help="run the launch-shape autotuner first (the defaults are its usual choice for this "
                         "workload, and the configuration of the committed kernel profile)")
    ap.add_argument("--no-tune", action="store_true", help="(default) use the launch shape given by the flags")
    ap.add_argument("--no-learned-cv", action="store_true")
    ap.add_argument("--cv-train-points", type=int, default=2_000_000)
    ap.add_argument("--cv-train-walks", type=int, default=32)
    ap.add_argument("--cv-steps", type=int, default=5000)
    ap.add_argument("--cv-batch", type=int, default=65536)
    ap.add_argument("--cv-grid", type=int, default=64)
    ap.add_argument("--to-rmse", type=float, default=1e-3,
                    help="after the timed passes, run learned-CV passes until the true RMSE vs the analytic "
                         "solution reaches this value, and report the measured time (0: skip)")
    ap.add_argument("--to-rmse-max-passes", type=int, default=400)
    ap.add_argument("--budget-s", type=float, default=420.0,
                    help="soft wall-time budget of the whole run: the plain run of equal wall time stops "
                         "early (and says so) when it would exceed it")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device copy (reference kernel for the PMC byte counters)")
    ap.add_argument("--cpu-points", type=int, default=16000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--point-order", default="kd", choices=["kd", "random"],
                    help="solve the points in kd-tree order (getSpatiallySortedPointIndices; batch: within "
                         "each instance) or as drawn; star2d solves its grid points in grid order")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def interior_points(torch, scene, prob, n, seed, dev):
    from zombie_amd import problems as PB
    pts, have = [], 0
    while have < n:
        cand = torch.from_numpy(PB.sample_box(prob.domain_min, prob.domain_max, max(4 * n, 1024), seed)).to(dev)
        seed += 1
        keep = cand[scene.inside_domain(cand).bool()]
        pts.append(keep)
        have += keep.shape[0]
    return torch.cat(pts)[:n].contiguous()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    from zombie_amd import autotune as AT
    from zombie_amd import engine as E
    from zombie_amd import learned_cv as LC
    from zombie_amd import problems as PB

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    if args.workload != "sphere3d":
        result = {"star2d": bench_star2d, "mixed2d": bench_mixed2d, "batch": bench_batch}[args.workload](args, torch, dev)
        print(json.dumps(result), flush=True)
        return
    if args.walks is None:
        args.walks = 256
    t_setup = time.time()
    prob = PB.sphere3d(seed=args.seed, n_walks=args.walks)
    scene = E.Scene(prob.positions, prob.indices, None, prob.domain_min, prob.domain_max,
                    bvh_width=args.bvh_width, leaf_size=args.leaf_size, sah_bins=16)
    log(f"scene: {scene.info.n_absorbing_prims} tris, {scene.info.n_absorbing_nodes} nodes, "
        f"depth {scene.info.absorbing_depth}, {scene.info.device_bytes/1e6:.2f} MB")
    P = interior_points(torch, scene, prob, args.points, args.seed, dev)
    if args.point_order == "kd":
        # the same points in kd-tree order: neighbouring lanes walk from nearby points
        P = P[torch.from_numpy(E.spatially_sorted_point_indices(P.cpu().numpy())).to(dev).long()].contiguous()
    dist, _, _, _ = scene.closest_point(P, want_point=False)
    n = P.shape[0]
    dp = E.DevicePoints(pt=P, dist_to_absorbing=dist.contiguous(),
                        n_walks=torch.full((n,), args.walks, dtype=torch.int32, device=dev), n_walks_max=args.walks)
    exact = prob.exact(P.cpu().numpy())
    torch.cuda.synchronize()
    log(f"setup {time.time()-t_setup:.1f}s, {n} points")

    # ---------------- launch shape ----------------
    tune = None
    scenes = {(args.bvh_width, args.leaf_size): scene}
    leaf, tol = args.leaf_size, args.tolerance
    if args.tune and not args.no_tune:
        tune = AT.autotune(prob, dp, scenes)
        for row in tune.table:
            log("tune: width {bvh_width} leaf {leaf_size} wpl {walks_per_lane:2d} regroup {regroup_interval:2d} "
                "step {step_interval:2d} tol {distance_tolerance:.2f} -> {walks_per_s:10.4g} walks/s "
                "{steps_per_s:10.4g} steps/s len {mean_walk_length:.2f}".format(**row))
        args.bvh_width, leaf = tune.bvh_width, tune.leaf_size
        args.walks_per_lane, args.regroup, tol = tune.walks_per_lane, tune.regroup_interval, tune.distance_tolerance
        args.step_interval = tune.step_interval
        scene = scenes[(args.bvh_width, leaf)]
        log(f"tuned: width {args.bvh_width} leaf {leaf} wpl {args.walks_per_lane} regroup {args.regroup} "
            f"step {args.step_interval} tol {tol}")

    def opts(step, count=False, tolerance=None, seed=None):
        return E.SolveOptions(seed=args.seed if seed is None else seed, stream_offset=step * args.walks,
                              walks_per_lane=args.walks_per_lane, regroup_interval=args.regroup,
                              step_interval=args.step_interval, waves_per_cu=args.waves_per_cu, count_visits=count,
                              distance_tolerance=tol if tolerance is None else tolerance)

    # ---------------- instrumented sample ----------------
    # a strided sample over the whole point set (in kd order the first points share one region)
    stride = max(1, n // 65536)
    sub = len(range(0, n, stride))
    dps = E.DevicePoints(pt=P[::stride].contiguous(), dist_to_absorbing=dp.dist_to_absorbing[::stride].contiguous(),
                         n_walks=dp.n_walks[::stride].contiguous(), n_walks_max=args.walks)
    # the algorithmic work: exact radii (the reference's walk) and, at every query point of those
    # walks, the nodes and triangles an exact nearest-first traversal of this BVH reads
    # (count_visits 2); the kernel's own visits at the tuned tolerance are reported beside them
    rx = E.wos_solve(scene, prob.pde, prob.settings, dps, E.DeviceStats.zeros(sub), opts(10_000, 2, 0.0))
    rc = E.wos_solve(scene, prob.pde, prob.settings, dps, E.DeviceStats.zeros(sub), opts(10_000, 1))
    len_exact = rx.total_steps / max(rx.total_walks, 1)
    nodes_per_step = rx.node_visits / max(rx.total_steps, 1)
    prims_per_step = rx.prim_visits / max(rx.total_steps, 1)
    bytes_per_step = nodes_per_step * scene.info.node_bytes + prims_per_step * scene.info.prim_bytes + 48.0
    own_nodes_per_step = rc.node_visits / max(rc.total_steps, 1)
    own_prims_per_step = rc.prim_visits / max(rc.total_steps, 1)
    visit_source = ("exact walks (tolerance 0) on a strided 65,536-point sample; at each of their query points "
                    "an exact nearest-first traversal from the root (count_visits 2, pinned against the oracle's "
                    "packed-BVH traversal by tests/test_gpu_roofline.py)")
    log(f"algorithmic visits/exact step: nodes {nodes_per_step:.2f} prims {prims_per_step:.2f} -> "
        f"{bytes_per_step:.0f} B/step; kernel's own visits/step at tolerance {tol}: nodes {own_nodes_per_step:.2f} "
        f"prims {own_prims_per_step:.2f}; walk length exact {len_exact:.3f}, tuned "
        f"{rc.total_steps / max(rc.total_walks, 1):.3f}")

    if args.calibrate_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del src, dst

    # ---------------- timed passes ----------------
    ex_dev = torch.from_numpy(exact).to(dev)

    class Progressive:
        """Progressive solves of all points (fresh walk streams per pass, statistics merged)."""

        def __init__(self, points, seed, label):
            self.points, self.seed, self.label = points, seed, label
            self.stats = E.DeviceStats.zeros(n)
            self.passes, self.solve_s, self.walks, self.steps = 0, 0.0, 0, 0
            self.kernel_ms = []
            self.curve = []

        def run(self, count):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = []
            for _ in range(count):
                reps.append(E.wos_solve(scene, prob.pde, prob.settings, self.points, self.stats,
                                        opts(self.passes, seed=self.seed)))
                self.passes += 1
                log(f"{self.label} pass {self.passes}: {reps[-1].total_steps} steps, {reps[-1].total_walks} walks, "
                    f"kernel {reps[-1].kernel_ms:.1f} ms")
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            self.solve_s += el
            self.walks += sum(r.total_walks for r in reps)
            self.steps += sum(r.total_steps for r in reps)
            self.kernel_ms += [r.kernel_ms for r in reps]
            return reps, el

        def rmse(self):
            r = float(torch.sqrt(torch.mean((self.stats.mean - ex_dev) ** 2)).item())
            self.curve.append((self.passes * args.walks, round(self.solve_s, 2), r))
            return r

        def host(self):
            return dict(mean=self.stats.mean.cpu().numpy(), var=self.stats.variance().cpu().numpy(),
                        n=self.stats.n.cpu().numpy())

    def warm(points, seed):
        for w in range(args.warmup):
            E.wos_solve(scene, prob.pde, prob.settings, points, E.DeviceStats.zeros(n), opts(1000 + w, seed=seed))
        torch.cuda.synchronize()

    warm(dp, args.seed)
    plain_run = Progressive(dp, args.seed, "plain")
    reports, elapsed = plain_run.run(args.steps)
    plain = plain_run.host()
    total_steps = sum(r.total_steps for r in reports)
    total_walks = sum(r.total_walks for r in reports)
    kernel_ms = sum(r.kernel_ms for r in reports) / len(reports)
    walks_per_s = total_walks / elapsed
    raw_steps_per_s = total_steps / elapsed
    # reported in the reference's unit: walks completed per second times the mean walk length
    # that exact radii give (equal to the raw rate when the tolerance is 0)
    steps_per_s = walks_per_s * len_exact if tol > 0 else raw_steps_per_s
    rmse = plain_run.rmse()
    se2 = float(np.mean(plain["var"] / np.maximum(plain["n"], 1)))
    se_rmse = math.sqrt(se2)
    var_plain = float(np.mean(plain["var"]))
    s_to_rmse_plain = var_plain / TARGET_RMSE ** 2 * n / walks_per_s

    # algorithmic bytes of one launch: its walks x the exact walk length x bytes per exact step
    exact_steps_per_pass = total_walks / len(reports) * len_exact
    achieved_gbs = bytes_per_step * exact_steps_per_pass / (kernel_ms * 1e-3) / 1e9
    # what the kernel itself reads: its own visits per (relaxed, raw) step
    own_bytes_per_step = own_nodes_per_step * scene.info.node_bytes + own_prims_per_step * scene.info.prim_bytes + 48.0
    own_read_gbs = own_bytes_per_step * (total_steps / len(reports)) / (kernel_ms * 1e-3) / 1e9
    # measured read ceilings: a 16 MB buffer stays in the XCDs' L2 caches, a 4 GiB one streams HBM
    E.probe_read_bandwidth(16 << 20, 20)
    l2_gbs = E.probe_read_bandwidth(16 << 20, 400)
    hbm_gbs = E.probe_read_bandwidth(4 << 30, 4)
    log(f"read probe: L2-resident {l2_gbs:.0f} GB/s, HBM {hbm_gbs:.0f} GB/s")
    # traffic past L2 per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    # the same kernel configuration, scaled to this launch's step count
    traffic = None
    issue = {}
    prof = REPO / "profiles" / "wos3d_sphere81920_1M.json"
    if prof.exists():
        pj = json.loads(prof.read_text())
        c = pj["config"]
        # traffic depends on the scene layout and the radius tolerance, not on the launch shape
        if (c["bvh_width"], c["leaf_size"], c["distance_tolerance"]) == (args.bvh_width, leaf, tol):
            traffic = pj["traffic_bytes"] / pj["raw_steps"] * (total_steps / len(reports))
            issue = {"valu_busy": pj.get("valu_busy"), "valu_lane_utilisation": pj.get("valu_lane_utilisation"),
                     "write_bytes_per_launch": pj.get("write_bytes"), "issue_source": pj["source"]}

    # ---------------- learned control variate ----------------
    learned = None
    if not args.no_learned_cv:
        t0 = time.perf_counter()
        TX = interior_points(torch, scene, prob, args.cv_train_points, args.seed + 7919, dev)
        td, _, _, _ = scene.closest_point(TX, want_point=False)
        tp = E.DevicePoints(pt=TX, dist_to_absorbing=td.contiguous(), n_walks=None, n_walks_max=0)
        X, Y = LC.generate_training_data(scene, prob, tp, walks=args.cv_train_walks, opts=opts(0, seed=args.seed + 1))
        torch.cuda.synchronize()
        t_data = time.perf_counter() - t0
        t0 = time.perf_counter()
        net = LC.MLP(LC.MLPConfig(in_dim=3, seed=args.seed), max_batch=args.cv_batch)
        losses = LC.train(net, X, Y, steps=args.cv_steps, batch=args.cv_batch, lr=1e-3, seed=args.seed,
                          log_every=max(1, args.cv_steps // 5))
        torch.cuda.synchronize()
        t_fit = time.perf_counter() - t0
        _, g0 = net.evaluate(P, value=False)
        grid = LC.bake_gradient_grid(net, prob.domain_min, prob.domain_max, args.cv_grid)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        train_flops = args.cv_steps * args.cv_batch * net.flops_per_sample()
        log(f"learned cv: data {t_data:.1f}s, train {t_fit:.1f}s ({train_flops / t_fit / 1e12:.1f} TFLOP/s), "
            f"train+bake {t_train:.1f}s, losses {losses}")
        dcv = E.DevicePoints(pt=P, dist_to_absorbing=dp.dist_to_absorbing, n_walks=dp.n_walks,
                             n_walks_max=args.walks, cv_gradient=g0.contiguous(), cv_gradient_grid=grid)
        cv_run = Progressive(dcv, args.seed + 2, "cv")   # independent streams from the plain passes (same kernel: warm)
        cv_reports, cv_el = cv_run.run(args.steps)
        cvh = cv_run.host()
        cv_wps = sum(r.total_walks for r in cv_reports) / cv_el
        var_cv = float(np.mean(cvh["var"]))
        s_to_rmse_cv = var_cv / TARGET_RMSE ** 2 * n / cv_wps + t_data + t_train
        # 4-SE agreement of the two estimators on a held-out 10k-point set
        h = slice(n - 10000, n)
        se = np.sqrt(plain["var"][h] / np.maximum(plain["n"][h], 1) + cvh["var"][h] / np.maximum(cvh["n"][h], 1))
        ok = se > 0
        z = (plain["mean"][h][ok] - cvh["mean"][h][ok]) / se[ok]
        rmse_cv_k = cv_run.rmse()
        learned = {
            "train_data_s": t_data, "train_s": t_train, "fit_s": t_fit, "final_loss": losses[-1][1] if losses else None,
            "training_mfma": {"flops": train_flops, "flops_per_sample": net.flops_per_sample(),
                              "achieved_tflops": train_flops / t_fit / 1e12,
                              "peak_tflops": BF16_PEAK_TFLOPS, "frac": train_flops / t_fit / 1e12 / BF16_PEAK_TFLOPS,
                              "note": "forward + backward + Adam of the 5k training steps (bf16 MFMA, fp32 "
                                      "accumulate) over their wall time"},
            "per_walk_variance_plain": var_plain, "per_walk_variance_cv": var_cv,
            "variance_reduction": var_plain / var_cv, "walks_per_s_cv": cv_wps,
            "rmse_vs_analytic_cv": rmse_cv_k,
            "projected_s_to_rmse_1e-3_cv": s_to_rmse_cv,
            "heldout_points": int(ok.sum()), "heldout_beyond_4se": int((np.abs(z) > 4).sum()),
            "heldout_z_mean": float(z.mean()), "heldout_z_std": float(z.std()),
            "variate": "first step: exact grad c(x0) per point; later steps: grad c baked on a "
                       f"{args.cv_grid}^3 grid (trilinear); MLP 3 -> 64 Fourier features -> 4x128 SiLU -> 1, "
                       "bf16 MFMA, Adam lr 1e-3",
        }

        # ---------------- the headline: time to RMSE 1e-3, measured ----------------
        if args.to_rmse > 0:
            rmse_cv = rmse_cv_k
            while rmse_cv > args.to_rmse and cv_run.passes < args.to_rmse_max_passes:
                cv_run.run(1)
                rmse_cv = cv_run.rmse()
            cv_total = cv_run.solve_s + t_data + t_train
            learned["measured_to_rmse"] = {
                "target_rmse": args.to_rmse, "reached": rmse_cv <= args.to_rmse, "rmse": rmse_cv,
                "passes": cv_run.passes, "walks_per_point": cv_run.passes * args.walks,
                "solve_s": cv_run.solve_s, "train_s": t_data + t_train, "total_s": cv_total,
                "curve": cv_run.curve[:: max(1, len(cv_run.curve) // 12)] + cv_run.curve[-1:]}
            log(f"learned cv reached rmse {rmse_cv:.3e} after {cv_run.passes} passes, {cv_total:.1f} s")
            # plain walk on spheres for the same wall time (bounded by the bench's time budget),
            # continuing the plain passes above; its measured RMSE, and the time plain needs for
            # the target along its measured 1/t curve
            left = args.budget_s - (time.time() - T_START) - 20.0   # keep room for the CPU leg
            per_pass = plain_run.solve_s / max(plain_run.passes, 1)
            target_s = min(cv_total, plain_run.solve_s + max(0.0, left))
            extra = max(0, int(math.floor((target_s - plain_run.solve_s) / per_pass)))
            for _ in range(extra):
                plain_run.run(1)
                if plain_run.passes % 8 == 0:
                    plain_run.rmse()
            rmse_plain = plain_run.rmse()
            t_plain = plain_run.solve_s * (rmse_plain / args.to_rmse) ** 2
            learned["plain_same_time"] = {
                "solve_s": plain_run.solve_s, "passes": plain_run.passes,
                "walks_per_point": plain_run.passes * args.walks, "rmse": rmse_plain,
                "equal_wall_time": plain_run.solve_s >= 0.98 * cv_total,
                "s_to_rmse_1e-3_on_measured_curve": t_plain,
                "curve": plain_run.curve,
                "note": "plain passes continue the timed ones; RMSE^2 falls as 1/t (no bias: "
                        "rmse_sq_over_se_sq ~ 1), so the time to the target is the measured time times "
                        "(rmse / target)^2"}
            learned["speedup_vs_plain_at_equal_rmse"] = t_plain / cv_total
            log(f"plain: {plain_run.solve_s:.1f} s -> rmse {rmse_plain:.3e}; plain to {args.to_rmse:g}: "
                f"{t_plain:.1f} s; learned cv {cv_total:.1f} s; speed-up {t_plain / cv_total:.2f}")
        else:
            learned["speedup_vs_plain_at_equal_rmse"] = s_to_rmse_plain / s_to_rmse_cv

    result = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "walk-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 geometry, f64 accumulation",
        "data": "synthetic (manufactured solution, seeded mesh and points)",
        "config": {"workload": "3D Poisson WoS, SH-displaced icosphere 81,920 tris, 1M points, "
                               f"{args.walks} walks/pt/pass, eps 1e-4",
                   "bvh_width": args.bvh_width, "leaf_size": leaf, "walks_per_lane": args.walks_per_lane,
                   "regroup_interval": args.regroup, "step_interval": args.step_interval,
                   "distance_tolerance": tol, "points": n,
                   "walks_per_point_per_step": args.walks, "autotuned": tune is not None,
                   "point_order": args.point_order},
        "value_definition": "completed walks/s x mean walk length with exact radii (reference walk-steps/s)",
        "walks_per_s": walks_per_s,
        "raw_walk_steps_per_s": raw_steps_per_s,
        "mean_walk_length_exact": len_exact,
        "rmse_vs_analytic": rmse,
        "estimated_se_rmse": se_rmse,
        "rmse_sq_over_se_sq": rmse ** 2 / se2,
        "walks_per_point_done": int(plain["n"].mean()),
        "mean_walk_length": total_steps / max(total_walks, 1),
        "projected_s_to_rmse_1e-3_plain": s_to_rmse_plain,
        "learned_cv": learned,
        "roofline": {"bound": "l2", "achieved": achieved_gbs, "peak": l2_gbs, "unit": "GB/s",
                     "frac": achieved_gbs / l2_gbs, "traffic": traffic,
                     "bytes_per_step": bytes_per_step, "nodes_per_step": nodes_per_step,
                     "prims_per_step": prims_per_step, "visit_count": visit_source,
                     "exact_steps_per_launch": exact_steps_per_pass,
                     "own_nodes_per_step": own_nodes_per_step, "own_prims_per_step": own_prims_per_step,
                     "own_read_gbs": own_read_gbs, "own_read_frac_of_l2": own_read_gbs / l2_gbs,
                     "kernel_ms": kernel_ms, **issue,
                     "hbm_peak": HBM_PEAK_GBS, "hbm_read_measured": hbm_gbs,
                     "traffic_past_l2_gbs": None if traffic is None else traffic / (kernel_ms * 1e-3) / 1e9,
                     "hbm_frac": None if traffic is None else traffic / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "note": "algorithmic bytes = BVH node + triangle bytes an exact nearest-first traversal "
                             "from the root reads per step (counted at the walks' own positions) + 48 B of walk "
                             "state; the 7 MB scene lives in L2 / MALL, so the binding memory level is L2 and "
                             "peak is the measured L2 read ceiling; traffic past L2 is a fraction of HBM peak "
                             "(hbm_frac); the kernel itself is bound by VALU issue under lane divergence "
                             "(valu_busy, valu_lane_utilisation from the committed SQ counter profile)"},
    }
    if not args.no_cpu_baseline:
        cs = max(1, n // args.cpu_points)   # strided: spread over the domain whatever the point order
        result["cpu_baseline"] = cpu_baseline(prob, P[::cs][: args.cpu_points].cpu().numpy(),
                                              dist[::cs][: args.cpu_points].cpu().numpy(), args.walks)
        result["speedup_vs_cpu_baseline"] = steps_per_s / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


def cpu_baseline(prob, pts, dist, walks):
    """The oracle (CPU restatement of the reference TBB solve) on a bounded sample."""
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    sc = O.Scene(prob.positions, prob.indices, None, prob.domain_min, prob.domain_max)
    st = O.Stats(len(pts))
    t = time.perf_counter()
    rep = O.solve(sc, prob.pde, prob.settings, pts, dist, walks, st, seed=7, rng_mode=O.RNG_PER_POINT,
                  threads=threads)
    el = time.perf_counter() - t
    return {"value": rep["total_steps"] / el, "unit": "walk-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(pts)} points x {walks} walks of the same workload ({rep['total_steps']} steps, "
                      f"{el:.1f} s, exact radii, one PCG32 per point as in the reference)"}


# ------------------------------------------------------------------------------------
# secondary workloads (parity-test configurations measured at their full sizes)
# ------------------------------------------------------------------------------------
def _result(metric_label, workload, value, elapsed, args, extra):
    r = {"metric": f"walk-steps/s ({metric_label})", "value": value, "unit": "walk-steps/s", "n_gpus": 1,
         "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
         "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 geometry, f64 accumulation",
         "data": "synthetic (manufactured solution, seeded geometry and points)", "config": workload}
    r.update(extra)
    return r


def _timed(torch, args, run):
    for w in range(args.warmup):
        r = run(1000 + w, None)
        log(f"warmup pass {w}: {r.total_walks} walks, {r.total_steps} steps, kernel {r.kernel_ms:.1f} ms")
    torch.cuda.synchronize()
    reps = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        reps.append(run(s, s))
        log(f"pass {s}: {reps[-1].total_walks} walks, {reps[-1].total_steps} steps, kernel {reps[-1].kernel_ms:.1f} ms")
    torch.cuda.synchronize()
    return reps, time.perf_counter() - t0


def _roofline(rep_count, node_bytes, prim_bytes, kernel_ms_per_pass, steps_per_pass, note, l2_gbs=None):
    """Closest-point bytes per step (the kernel's own visits, counted on an instrumented pass)
    against the measured L2 read ceiling when the scene is L2-resident (l2_gbs given), else
    against the published HBM peak."""
    nodes = rep_count.node_visits / max(rep_count.total_steps, 1)
    prims = rep_count.prim_visits / max(rep_count.total_steps, 1)
    bps = nodes * node_bytes + prims * prim_bytes + 48.0
    ach = bps * steps_per_pass / (kernel_ms_per_pass * 1e-3) / 1e9
    bound, peak = ("l2", l2_gbs) if l2_gbs else ("hbm", HBM_PEAK_GBS)
    return {"bound": bound, "achieved": ach, "peak": peak, "unit": "GB/s", "frac": ach / peak,
            "traffic": None, "bytes_per_step": bps, "nodes_per_step": nodes, "prims_per_step": prims,
            "kernel_ms": kernel_ms_per_pass, "hbm_peak": HBM_PEAK_GBS, "note": note}


def _roofline_batch(rep_exact, rep_own, info, kernel_ms_per_pass, exact_steps_per_pass, l2_gbs):
    """Batched workload: the algorithmic bytes per exact step (an exact nearest-first traversal
    from the root of the instance's BVH at every query point of exact walks, count_visits 2) x
    the exact-step equivalent of the timed replays (completed walks x exact walk length) / their
    graph time.  The instances' points are solved instance by instance, so the working set at a
    time is a few instances' BVHs (0.2-17 MB each, by triangle count): the bytes come from L2 and the MALL,
    and peak is the measured L2 read ceiling, as for the headline workload."""
    nodes = rep_exact.node_visits / max(rep_exact.total_steps, 1)
    prims = rep_exact.prim_visits / max(rep_exact.total_steps, 1)
    bps = nodes * info.node_bytes + prims * info.prim_bytes + 48.0
    ach = bps * exact_steps_per_pass / (kernel_ms_per_pass * 1e-3) / 1e9
    s = max(rep_own.total_steps, 1)
    return {"bound": "l2", "achieved": ach, "peak": l2_gbs, "unit": "GB/s", "frac": ach / l2_gbs,
            "traffic": None, "bytes_per_step": bps, "nodes_per_step": nodes, "prims_per_step": prims,
            "own_nodes_per_raw_step": rep_own.node_visits / s, "own_prims_per_raw_step": rep_own.prim_visits / s,
            "kernel_ms": kernel_ms_per_pass, "hbm_peak": HBM_PEAK_GBS,
            "note": "exact nearest-first traversal reads per exact step (count_visits 2 on exact walks over a "
                    "strided sample) x exact-step equivalent of the timed replays / their graph time (HIP events "
                    "on the launch stream); the 256 scenes (0.5 GB) are in HBM, solved instance by instance"}


def _roofline_wost(rep_count, info, kernel_ms_per_pass, steps_per_pass, l2_gbs):
    """Walk-on-stars bytes per step: every BVH node and primitive the step's queries read (the
    kernel's own visits, counted by the visit-counting variant of the interleaved kernel on a
    sample): absorbing-boundary distance and projection, reflecting-boundary ray, visibility,
    boundary-sample descent and Robin bound, silhouette star radius; + 48 B of walk state.  The
    scenes of this workload live in L2, so peak is the measured L2 read ceiling."""
    s = max(rep_count.total_steps, 1)
    sil_bytes = 32 if info.dim == 2 else 64
    parts = {
        "absorbing": (rep_count.node_visits / s, rep_count.prim_visits / s, info.prim_bytes),
        "reflecting": (rep_count.reflecting_node_visits / s, rep_count.reflecting_prim_visits / s, info.prim_bytes),
        "silhouette": (rep_count.silhouette_node_visits / s, rep_count.silhouette_prim_visits / s, sil_bytes),
    }
    bps = 48.0 + sum(n * info.node_bytes + p * pb for n, p, pb in parts.values())
    ach = bps * steps_per_pass / (kernel_ms_per_pass * 1e-3) / 1e9
    return {"bound": "l2", "achieved": ach, "peak": l2_gbs, "unit": "GB/s", "frac": ach / l2_gbs, "traffic": None,
            "bytes_per_step": bps, "kernel_ms": kernel_ms_per_pass,
            "per_step": {k: {"nodes": n, "prims": p} for k, (n, p, _) in parts.items()},
            "hbm_peak": HBM_PEAK_GBS,
            "note": "algorithmic bytes = nodes and primitives read by the step's absorbing, reflecting and "
                    "silhouette queries (the kernel's own visits on a strided sample), node " +
                    f"{info.node_bytes} B, segment {info.prim_bytes} B, silhouette candidate {sil_bytes} B; "
                    "peak is the measured L2 read ceiling (the 2D scenes are L2-resident)"}


def _oracle_baseline(label, run, cores):
    t = time.perf_counter()
    rep = run()
    el = time.perf_counter() - t
    return {"value": rep["total_steps"] / el, "unit": "walk-steps/s", "cores": cores, "kind": "port",
            "sample": f"{label} ({rep['total_steps']} steps, {el:.1f} s, one PCG32 per point as in the reference)"}


def bench_star2d(args, torch, dev):
    """2D Laplace on the star domain r = 1 + 0.3 cos 5t (1,024 segments), g = e^x cos y,
    the 34,627 interior points of a 256^2 grid, 512 walks, epsilon 1e-3."""
    import numpy as np
    import oracle as O
    from zombie_amd import engine as E
    from zombie_amd import problems as PB
    walks = args.walks or 512
    prob = PB.star2d(n_walks=walks)
    scene = E.Scene(prob.positions, prob.indices, None, prob.domain_min, prob.domain_max, bvh_width=args.bvh_width,
                    leaf_size=args.leaf_size)
    P = torch.from_numpy(prob.meta["points"]).to(dev)
    n = P.shape[0]
    dist, _, _, _ = scene.closest_point(P, want_point=False)
    dp = E.DevicePoints(pt=P, dist_to_absorbing=dist.contiguous(),
                        n_walks=torch.full((n,), walks, dtype=torch.int32, device=dev), n_walks_max=walks)
    stats = E.DeviceStats.zeros(n)

    def opts(s, count=False):
        return E.SolveOptions(seed=args.seed, stream_offset=s * walks, walks_per_lane=args.walks_per_lane,
                              regroup_interval=args.regroup, step_interval=args.step_interval, count_visits=count)

    def run(s, keep):
        return E.wos_solve(scene, prob.pde, prob.settings, dp, stats if keep is not None else E.DeviceStats.zeros(n),
                           opts(s))
    reps, el = _timed(torch, args, run)
    steps = sum(r.total_steps for r in reps)
    walks_done = sum(r.total_walks for r in reps)
    kms = sum(r.kernel_ms for r in reps) / len(reps)
    rc = E.wos_solve(scene, prob.pde, prob.settings, dp, E.DeviceStats.zeros(n), opts(99, True))
    E.probe_read_bandwidth(16 << 20, 20)
    l2_gbs = E.probe_read_bandwidth(16 << 20, 400)
    mean = stats.mean.cpu().numpy()
    rmse = float(np.sqrt(np.mean((mean - prob.exact(prob.meta["points"])) ** 2)))
    res = _result("2D Laplace WoS, star domain", {
        "workload": f"2D Laplace WoS, star domain 1,024 segments, {n} grid points, {walks} walks/pt/pass, eps 1e-3",
        "bvh_width": args.bvh_width, "leaf_size": args.leaf_size, "walks_per_lane": args.walks_per_lane,
        "regroup_interval": args.regroup, "step_interval": args.step_interval, "points": n}, steps / el, el, args,
        {"walks_per_s": walks_done / el, "mean_walk_length": steps / max(walks_done, 1), "rmse_vs_analytic": rmse,
         "roofline": _roofline(rc, scene.info.node_bytes, scene.info.prim_bytes, kms, steps / len(reps),
                               "closest-point bytes the kernel's own queries read per step; the 1,024-segment "
                               "scene is L2-resident, so peak is the measured L2 read ceiling", l2_gbs)})
    if not args.no_cpu_baseline:
        pts, d = prob.meta["points"], dist.cpu().numpy()
        threads = min(os.cpu_count() or 1, 16)
        osc = O.Scene(prob.positions, prob.indices, None, prob.domain_min, prob.domain_max)
        res["cpu_baseline"] = _oracle_baseline(f"{len(pts)} points x {walks} walks", lambda: O.solve(
            osc, prob.pde, prob.settings, pts, d, walks, O.Stats(len(pts)), seed=7, rng_mode=O.RNG_PER_POINT,
            threads=threads), threads)
        res["speedup_vs_cpu_baseline"] = res["value"] / res["cpu_baseline"]["value"]
    return res


def bench_mixed2d(args, torch, dev):
    """2D walk on stars: the star curve with 4,096 segments, Neumann on the upper half and on
    32 interior slits (double-sided), u = x^3 - 3xy^2; 200k interior points, 1,024 walks."""
    import numpy as np
    import oracle as O
    from zombie_amd import engine as E
    from zombie_amd import problems as PB
    walks = args.walks or 1024
    t0 = time.time()
    prob = PB.mixed2d(n_walks=walks)
    scene = E.Scene(prob.positions, prob.indices, prob.is_reflecting, prob.domain_min, prob.domain_max,
                    bvh_width=args.bvh_width, leaf_size=args.leaf_size, solve_double_sided=True)
    V, I = prob.meta["outer"]
    n_target = args.points if args.points != 1_000_000 else 200_000
    cand = PB.sample_box(prob.domain_min, prob.domain_max, 2 * n_target, args.seed)
    pts = cand[PB.points_in_polygon(cand, V, I)][:n_target]
    if args.point_order == "kd":
        pts = pts[E.spatially_sorted_point_indices(pts)]
    P = torch.from_numpy(pts).to(dev)
    n = P.shape[0]
    dist, _, _, _ = scene.closest_point(P, boundary=0, want_point=False)
    dp = E.DevicePoints(pt=P, dist_to_absorbing=dist.contiguous(),
                        n_walks=torch.full((n,), walks, dtype=torch.int32, device=dev), n_walks_max=walks)
    stats = E.DeviceStats.zeros(n)
    log(f"mixed2d setup {time.time() - t0:.1f}s: {n} points, {walks} walks")

    def opts(s, count=False, tol=None):
        return E.SolveOptions(seed=args.seed, stream_offset=s * walks, walks_per_lane=args.walks_per_lane,
                              regroup_interval=args.regroup, step_interval=args.wost_step_interval,
                              count_visits=count, distance_tolerance=args.tolerance if tol is None else tol,
                              wost_interleave_queries=args.wost_interleave)

    # a pass of `walks` walks per point runs as progressive calls of `chunk` walks (fresh stream
    # offsets; equal to one call of `walks` walks), so a long pass reports its progress
    chunk = min(walks, args.chunk_walks)
    dpc = E.DevicePoints(pt=P, dist_to_absorbing=dist.contiguous(),
                         n_walks=torch.full((n,), chunk, dtype=torch.int32, device=dev), n_walks_max=chunk)

    def run(s, keep):
        st = stats if keep is not None else E.DeviceStats.zeros(n)
        tot = None
        for c in range(0, walks, chunk):
            o = opts(0)
            o.stream_offset = s * walks + c
            r = E.wost_solve(scene, prob.pde, prob.settings, dpc, st, o)
            log(f"  chunk {c // chunk}: {r.total_walks} walks, {r.total_steps} steps, kernel {r.kernel_ms:.1f} ms")
            if tot is None:
                tot = r
            else:
                tot = E.SolveReport(tuple(a + b for a, b in zip(tot.completion_codes, r.completion_codes)),
                                    tot.total_steps + r.total_steps, tot.total_walks + r.total_walks,
                                    tot.node_visits + r.node_visits, tot.prim_visits + r.prim_visits,
                                    tot.queries + r.queries, tot.kernel_ms + r.kernel_ms)
        return tot
    reps, el = _timed(torch, args, run)
    steps = sum(r.total_steps for r in reps)
    walks_done = sum(r.total_walks for r in reps)
    kms = sum(r.kernel_ms for r in reps) / len(reps)
    stride = max(1, n // 16384)   # a strided sample spread over the domain whatever the point order
    sub = len(range(0, n, stride))
    dps = E.DevicePoints(pt=P[::stride].contiguous(), dist_to_absorbing=dist[::stride].contiguous(),
                         n_walks=torch.full((sub,), 64, dtype=torch.int32, device=dev), n_walks_max=64)
    # visit counts of every BVH the steps read (absorbing, reflecting, silhouette)
    rc = E.wost_solve(scene, prob.pde, prob.settings, dps, E.DeviceStats.zeros(sub), opts(99, True))
    E.probe_read_bandwidth(16 << 20, 20)
    l2_gbs = E.probe_read_bandwidth(16 << 20, 400)
    # exact-distance walk length on the sample: the value is reported in the reference's unit
    rx = E.wost_solve(scene, prob.pde, prob.settings, dps, E.DeviceStats.zeros(sub), opts(98, tol=0.0))
    len_exact = rx.total_steps / max(rx.total_walks, 1)
    mean = stats.mean.cpu().numpy()
    rmse = float(np.sqrt(np.mean((mean - prob.exact(pts)) ** 2)))
    value = walks_done / el * len_exact if args.tolerance > 0 else steps / el
    res = _result("2D mixed Dirichlet/Neumann walk on stars", {
        "workload": f"2D WoSt, star curve 4,096 segments + 32 Neumann slits, {n} points, {walks} walks/pt/pass, "
                    "eps 1e-3", "bvh_width": args.bvh_width, "leaf_size": args.leaf_size,
        "walks_per_lane": args.walks_per_lane, "regroup_interval": args.regroup,
        "step_interval": args.wost_step_interval,
        "kernel": "interleaved" if args.wost_interleave else "run-to-completion",
        "distance_tolerance": args.tolerance, "points": n, "point_order": args.point_order}, value, el, args,
        {"value_definition": "completed walks/s x mean walk length with exact distances (reference walk-steps/s)",
         "walks_per_s": walks_done / el, "raw_walk_steps_per_s": steps / el, "mean_walk_length_exact": len_exact,
         "mean_walk_length": steps / max(walks_done, 1), "rmse_vs_analytic": rmse,
         "completion_codes": [int(sum(r.completion_codes[k] for r in reps)) for k in range(4)],
         "roofline": _roofline_wost(rc, scene.info, kms, steps / len(reps), l2_gbs)})
    if not args.no_cpu_baseline:
        k, kw = 2048, 64
        ks = max(1, n // k)   # strided sample
        threads = min(os.cpu_count() or 1, 16)
        osc = O.Scene(prob.positions, prob.indices, prob.is_reflecting, prob.domain_min, prob.domain_max,
                      solve_double_sided=True)
        res["cpu_baseline"] = _oracle_baseline(
            f"{k} points x {kw} walks (binary-BVH star radius, sampling, visibility and distance queries)",
            lambda: O.solve(osc, prob.pde, prob.settings, pts[::ks][:k], dist[::ks][:k].cpu().numpy(), kw, O.Stats(k), seed=7,
                            wost=True, rng_mode=O.RNG_PER_POINT, threads=threads), threads)
        res["speedup_vs_cpu_baseline"] = res["value"] / res["cpu_baseline"]["value"]
    return res


def bench_batch(args, torch, dev):
    """256 screened-Poisson instances (SH-displaced geodesic spheres, triangles log-uniform in
    [2k, 200k], points log-uniform in [1k, 1e5], sigma ~ U[0, 10], u = exp(a.x) with
    |a|^2 = sigma), 128 walks, one persistent launch per pass replayed from a hipGraph."""
    import concurrent.futures as cf
    import numpy as np
    import oracle as O
    from zombie_amd import autotune as AT
    from zombie_amd import engine as E
    from zombie_amd import problems as PB
    walks = args.walks or 128
    t0 = time.time()
    spec = PB.screened_batch_spec(args.instances, seed=args.seed)

    def build(sp):
        pr = PB.screened_instance(sp["n_tris"], sp["sigma"], sp["direction"], sp["seed"], n_walks=walks)
        return pr
    with cf.ThreadPoolExecutor(16) as ex:
        probs = list(ex.map(build, spec))

    def scenes_for(width, leaf):
        with cf.ThreadPoolExecutor(16) as ex:
            return list(ex.map(lambda pr: E.Scene(pr.positions, pr.indices, None, pr.domain_min, pr.domain_max,
                                                  bvh_width=width, leaf_size=leaf), probs))
    scenes = scenes_for(args.bvh_width, args.leaf_size)
    pts, dists, inst, exact = [], [], [], []
    for i, (sp, pr, sc) in enumerate(zip(spec, probs, scenes)):
        need, got, seed = sp["n_points"], [], args.seed + 100 * i
        while sum(len(g) for g in got) < need:
            c = torch.from_numpy(PB.sample_box(pr.domain_min, pr.domain_max, 2 * need + 64, seed)).to(dev)
            seed += 1
            got.append(c[sc.inside_domain(c).bool()])
        p = torch.cat(got)[:need].contiguous()
        if args.point_order == "kd":
            # each instance's points in kd-tree order: neighbouring lanes walk from nearby points
            p = p[torch.from_numpy(E.spatially_sorted_point_indices(p.cpu().numpy())).to(dev).long()].contiguous()
        d, _, _, _ = sc.closest_point(p, want_point=False)
        pts.append(p); dists.append(d); inst.append(torch.full((need,), i, dtype=torch.int32, device=dev))
        exact.append(PB.exp_linear(p.cpu().numpy(), pr.meta["a"]))
    P = torch.cat(pts).contiguous(); D = torch.cat(dists).contiguous(); IOP = torch.cat(inst).contiguous()
    exact = np.concatenate(exact)
    n = P.shape[0]
    n_tris = sum(int(pr.indices.shape[0]) for pr in probs)
    log(f"batch setup {time.time()-t0:.1f}s: {len(probs)} instances, {n_tris} triangles, {n} points")
    dp = E.DevicePoints(pt=P, dist_to_absorbing=D, n_walks=torch.full((n,), walks, dtype=torch.int32, device=dev),
                        n_walks_max=walks)
    batches = {args.bvh_width: (scenes, E.Batch(scenes, [pr.pde for pr in probs]))}
    wpl, rg, width = args.walks_per_lane, args.regroup, args.bvh_width
    tune = None
    if not args.no_tune:
        # the sample must cover many instances: a strided gather of the concatenation
        idx = torch.arange(0, n, max(1, n // 131072), device=dev)[:131072]
        sdp = E.DevicePoints(pt=P[idx].contiguous(), dist_to_absorbing=D[idx].contiguous(), n_walks=None,
                             n_walks_max=0)
        for w in (4, 8):
            if w not in batches:
                sc_w = scenes_for(w, args.leaf_size)
                batches[w] = (sc_w, E.Batch(sc_w, [pr.pde for pr in probs]))
        tune = AT.autotune_batch(probs, sdp, IOP[idx].contiguous(), leaf_size=args.leaf_size, batches=batches,
                                 distance_tolerance=args.tolerance, step_interval=args.step_interval)
        for row in tune.table:
            log("tune: width {bvh_width} wpl {walks_per_lane} regroup {regroup_interval:2d} -> "
                "{walks_per_s:10.4g} walks/s".format(**row))
        width, wpl, rg = tune.bvh_width, tune.walks_per_lane, tune.regroup_interval
    batch = batches[width][1]
    stats = E.DeviceStats.zeros(n)
    o = E.SolveOptions(seed=args.seed, stream_offset=0, walks_per_lane=wpl, regroup_interval=rg,
                       step_interval=args.step_interval, distance_tolerance=args.tolerance)
    graph = E.SolveGraph(batch, probs[0].settings, dp, IOP, stats, o)
    warm_stats = E.DeviceStats.zeros(n)
    for w in range(args.warmup):
        E.batch_wos_solve(batch, probs[0].settings, dp, IOP, warm_stats, E.SolveOptions(
            seed=args.seed + 1, walks_per_lane=wpl, regroup_interval=rg, step_interval=args.step_interval,
            distance_tolerance=args.tolerance))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    steps = walks_done = 0
    replay_ms = 0.0
    for s in range(args.steps):
        # events on the stream the graph is launched on (torch's current stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        graph.launch()
        e1.record()
        r = graph.report()          # synchronises: counters of this replay
        steps += r.total_steps; walks_done += r.total_walks
        e1.synchronize()
        replay_ms += e0.elapsed_time(e1)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    graph.close()
    # on a strided sample: exact walks (tolerance 0, the reference's walk) whose query points
    # each also run a counted exact nearest-first traversal from the root of their instance's BVH
    # (count_visits 2: the algorithmic reads per step, and the exact walk length: the value is
    # reported in the reference's unit, completed walks/s x exact-radius mean walk length); and
    # the kernel's own visits at the timed tolerance (count_visits 1), reported beside them
    idx = torch.arange(0, n, max(1, n // 262144), device=dev)[:262144]
    sx = E.DevicePoints(pt=P[idx].contiguous(), dist_to_absorbing=D[idx].contiguous(),
                        n_walks=torch.full((idx.numel(),), 32, dtype=torch.int32, device=dev), n_walks_max=32)
    sop = dict(walks_per_lane=wpl, regroup_interval=rg, step_interval=args.step_interval)
    rx = E.batch_wos_solve(batch, probs[0].settings, sx, IOP[idx].contiguous(), E.DeviceStats.zeros(idx.numel()),
                           E.SolveOptions(seed=args.seed + 4, distance_tolerance=0.0, count_visits=2, **sop))
    rc = E.batch_wos_solve(batch, probs[0].settings, sx, IOP[idx].contiguous(), E.DeviceStats.zeros(idx.numel()),
                           E.SolveOptions(seed=args.seed + 3, distance_tolerance=args.tolerance, count_visits=1, **sop))
    E.probe_read_bandwidth(16 << 20, 20)
    l2_gbs = E.probe_read_bandwidth(16 << 20, 400)
    len_exact = rx.total_steps / max(rx.total_walks, 1)
    mean = stats.mean.cpu().numpy()
    var = stats.variance().cpu().numpy()
    nn = np.maximum(stats.n.cpu().numpy(), 1)
    err = mean - exact
    pooled = float(err.sum() / np.sqrt((var / nn).sum()))
    res = _result("batched screened Poisson, one graph-captured launch", {
        "workload": f"{len(probs)} screened-Poisson instances, {n_tris} triangles, {n} points, {walks} walks/pt/pass, "
                    "eps 1e-4", "bvh_width": width, "leaf_size": args.leaf_size, "walks_per_lane": wpl,
        "regroup_interval": rg, "step_interval": args.step_interval, "distance_tolerance": args.tolerance,
        "points": n, "instances": len(probs), "autotuned": tune is not None, "point_order": args.point_order},
        walks_done / el * len_exact if args.tolerance > 0 else steps / el, el, args,
        {"value_definition": "completed walks/s x mean walk length with exact radii (reference walk-steps/s)",
         "walks_per_s": walks_done / el, "raw_walk_steps_per_s": steps / el,
         "mean_walk_length": steps / max(walks_done, 1), "mean_walk_length_exact": len_exact,
         "rmse_vs_analytic": float(np.sqrt(np.mean(err ** 2))), "pooled_z_vs_analytic": pooled,
         "roofline": _roofline_batch(rx, rc, batches[width][0][0].info, replay_ms / args.steps,
                                     walks_done / args.steps * len_exact, l2_gbs)})
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        i0 = int(np.argmin([abs(sp["n_tris"] - 43000) for sp in spec]))   # an instance of typical size
        sel = (IOP == i0).nonzero().flatten()[:20000]
        pr = probs[i0]
        osc = O.Scene(pr.positions, pr.indices, None, pr.domain_min, pr.domain_max)
        pts0, d0 = P[sel].cpu().numpy(), D[sel].cpu().numpy()
        res["cpu_baseline"] = _oracle_baseline(
            f"instance {i0} ({spec[i0]['n_tris']} triangles, sigma {spec[i0]['sigma']:.2f}): {len(pts0)} points x "
            f"{walks} walks, exact radii", lambda: O.solve(osc, pr.pde, pr.settings, pts0, d0, walks, O.Stats(len(pts0)),
                                                         seed=7, rng_mode=O.RNG_PER_POINT, threads=threads), threads)
        res["speedup_vs_cpu_baseline"] = res["value"] / res["cpu_baseline"]["value"]
    return res


if __name__ == "__main__":
    main()
