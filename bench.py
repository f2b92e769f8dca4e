#!/usr/bin/env python3
"""LM iterations/s of the HIP bundle adjuster on the 2,000-camera / 1,000,000-point scene.

Workload: the large synthetic scene (seed 0xBA0003): 2,000 cameras with free f, k1, k2; 1M points;
~6.1M observations with geometric track lengths (mean 6, clipped to [2, 50]); 2 % outliers; Huber
loss with delta 1 px; inexact LM with block-Jacobi PCG on the implicit Schur system (forcing
tolerance 0.1). One step is one LM iteration: linearization, point elimination, block assembly,
the PCG solve and the candidate pass (sparse/self_calibration/mod.rs:212-639).

The timed region is the LM loop as the library reports it (`seconds_iterations`): it starts after
the problem has been lowered and uploaded, so every input is resident in HBM, and every iteration
ends in a synchronous scalar readback. The printed JSON line carries:

* ``roofline`` for the dominant kernel, from HIP events the library records around each of that
  kernel's launches on the solve's stream, and that kernel's algorithmic bytes per launch;
* ``cpu_baseline``: the oracle (a single-threaded C++ restatement of the same LM path, used here
  only as the CPU reference timing) on a bounded subsample of the same scene generator, scaled
  to the full scene by observation count;
* ``time_to_cost``: seconds until the loop holds a state within 1e-6 relative of the converged
  FP64 cost of a full run.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "openmvg-rs_amd"))

HBM_PEAK_GBS = 8000.0
PCG_SEG = 128  # FP64 camera-major positions per segment of the PCG camera pass


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def lowered_counts(problem) -> dict:
    """Counts of the lowered problem that the byte formulas use (every point of the synthetic
    scenes is kept, every camera has one intrinsic group)."""
    n_obs = problem.n_observations
    n_pts = problem.n_points
    n_cams = problem.n_poses
    # camera windows of the PCG camera pass: from each camera's first camera-major position
    # rounded down to a pair of observations (16-byte FP64 groups), PCG_SEG positions each
    per_cam = np.bincount(problem.obs_pose, minlength=n_cams).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(per_cam)])
    span = np.where(per_cam > 0, off[1:] - off[:-1] // 2 * 2, 0)
    segments = int(np.sum((span + PCG_SEG - 1) // PCG_SEG))
    return {"obs": n_obs, "points": n_pts, "cameras": n_cams, "segments": segments}


def schur_pairs(problem) -> int:
    """Camera pairs c < j of the explicit reduced system: cameras that share a point (every point
    of the synthetic scenes is free)."""
    order = np.lexsort((problem.obs_pose, problem.obs_point))
    cam = problem.obs_pose[order].astype(np.int64)
    pt = problem.obs_point[order]
    starts = np.flatnonzero(np.r_[True, pt[1:] != pt[:-1]])
    lens = np.diff(np.r_[starts, len(pt)])
    nc = int(problem.n_poses)
    keys = []
    for k in range(2, int(lens.max()) + 1 if len(lens) else 2):
        s = starts[lens == k]
        if len(s) == 0:
            continue
        c = cam[s[:, None] + np.arange(k)[None, :]]
        a, b = np.triu_indices(k, 1)
        lo, hi = np.minimum(c[:, a], c[:, b]), np.maximum(c[:, a], c[:, b])
        keys.append((lo * nc + hi).ravel())
    return len(np.unique(np.concatenate(keys))) if keys else 0


def schur_counts(problem) -> dict:
    """Stored slots (the diagonal and upper blocks), camera pairs and observation pairs of the
    explicit reduced system."""
    pairs = schur_pairs(problem)
    pt = np.sort(problem.obs_point)
    k = np.diff(np.r_[np.flatnonzero(np.r_[True, pt[1:] != pt[:-1]]), len(pt)]).astype(np.int64)
    return {"slots": int(problem.n_poses) + pairs, "pairs": pairs, "terms": int(np.sum(k * (k - 1) // 2))}


def algorithmic_bytes(kernel: str, counts: dict, D: int) -> float:
    """Algorithmic HBM bytes of one launch of `kernel` in FP64: each input read once, each output
    written once; vectors indexed by camera or point are counted once per launch.

    The camera pass rebuilds the Jacobians from each observation's compact linearisation state
    (n = (x/z, y/z), sqrt(w)/z, sqrt(w): 32 B) and the camera records (rotation, intrinsics and the
    product input, 192 B per camera); the point pass uses J~_X with n and sqrt(w).
    pcg_points (k_pcg_points: w_p = V^-1 sum_o J~_X,o^T J~_c,o v, persistent wavefronts over
      64-observation point chunks): per observation J~_X (48 B), n and sqrt(w) (24 B) and its camera
      and point indices (8 B); per point its offset (4 B), fixed flag (1 B), packed V^-1 (48 B), X
      (24 B) and the written w_p record (32 B); per camera its world-frame record (72 B), read once.
    pcg_cameras (k_pcg_cameras: per camera segment of 128 positions the sums of
      J~_c^T (J~_c v - J~_X w_p)): per observation the compact state and rotated point of the
      camera-major copy (56 B) and its point index (4 B); per point the w_p record once (32 B); per
      segment its camera (4 B) and the written sums (72 B); per camera offsets (12 B) and its
      record (192 B).
    linearize (k_linearize): per observation pixel (16 B), camera and point index (8 B), the
      written r~, compact state and J~_X (96 B), cost, squared residual and flag (17 B); per point
      X (24 B); per camera R, t, f, k1, k2 (120 B).
    pcg_product (k_pcg_product_upper: the first pass of q = S v over the stored diagonal and
      upper blocks of the explicit reduced system, one workgroup per camera row): per stored slot
      the 9 x 9 block (648 B) and its neighbour index (4 B); per camera pair the written column
      sums (72 B); per camera the 9 input values (72 B), its offsets (16 B) and the written row
      sums (72 B).
    schur_pairs (k_schur_pairs: the off-diagonal blocks of the explicit S): each observation's Z
      record once (27 values, 216 B), 8 B per observation pair, and per camera pair the written
      9 x 9 block and its indices (656 B).
    camera_assembly (k_camera_assembly with the explicit system): per observation the
      camera-major rows read (72 B) and point index (4 B) and the Z record written (216 B); per
      point its packed record (128 B); per segment its camera (4 B) and written sums (576 B); per
      camera R, f, k1, k2 (120 B).
    """
    no, npt, nc = counts["obs"], counts["points"], counts["cameras"]
    if kernel == "schur_pairs":
        return no * 216 + counts["terms"] * 8 + counts["pairs"] * (648 + 8)
    if kernel == "camera_assembly":
        return no * (72 + 4 + 216) + npt * 128 + counts["segments"] * (4 + 576) + nc * 120
    if kernel == "pcg_product":
        return counts["slots"] * (648 + 4) + counts["pairs"] * 72 + nc * (72 + 16 + 72)
    if kernel == "pcg_points":
        return no * (48 + 24 + 8) + npt * (4 + 1 + 48 + 24 + 32) + nc * 72
    if kernel == "pcg_cameras":
        return no * (56 + 4) + npt * 32 + counts["segments"] * (4 + 72) + nc * (12 + 192)
    if kernel == "linearize":
        return no * (16 + 8 + 96 + 17) + npt * 24 + nc * 120
    raise ValueError(kernel)


def measured_traffic(kernel: str, per_launch: float):
    """HBM bytes per launch of `kernel` from the committed counter summary
    profiles/<kernel>_large_fp64_pmc.md (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    same counter's reading for a copy of known size).  None when there is no summary, or when it
    was measured for a different algorithmic byte count (an older kernel)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_large_fp64_pmc.md")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for ln in fh:
            for key in ("traffic_bytes_per_launch", "algorithmic_bytes_per_launch"):
                if ln.startswith(key + ":"):
                    vals[key] = float(ln.split(":", 1)[1])
    if "traffic_bytes_per_launch" not in vals or abs(vals.get("algorithmic_bytes_per_launch", -1) - per_launch) > 1:
        return None
    return vals["traffic_bytes_per_launch"]


def cpu_baseline(scene_fn, n_points_sample: int, n_obs_full: int, iterations: int) -> dict:
    """The oracle's LM iterations/s on a subsample, scaled to the full scene by observations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import load_oracle

    from openmvg_ba import SolveConfig, _abi as A
    from openmvg_ba.problem import solve_with

    oracle = load_oracle()
    scene = scene_fn(n_points=n_points_sample)
    cfg = SolveConfig(max_iterations=iterations, linear_solver=A.SOLVER_ITERATIVE)
    t0 = time.perf_counter()
    sol = solve_with(oracle, "bao_solve", scene.problem, cfg)
    wall = time.perf_counter() - t0
    tr = sol.trace
    n_it = len(tr["seconds"])
    loop_s = float(tr["seconds"][-1]) if n_it else float("nan")
    rate_sample = n_it / loop_s
    n_obs = scene.problem.n_observations
    return {
        "value": rate_sample * n_obs / n_obs_full,
        "unit": "LM iter/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (single-threaded C++ restatement), same generator with {n_points_sample} points / "
                   f"{n_obs} observations, {n_it} LM iterations in {loop_s:.2f} s "
                   f"({rate_sample:.3f} iter/s, {int(tr['pcg_iterations'].sum())} PCG iterations), scaled by "
                   f"{n_obs}/{n_obs_full} observations; whole call {wall:.1f} s"),
    }


def bench_mixed(ba, prob, base, steps: int, ttc: dict | None) -> dict:
    """The mixed-precision form on the same workload: LM iterations/s over `steps` iterations and
    the loop time until an accepted state is within 1e-6 of the converged FP64 cost."""
    from openmvg_ba import _abi as A

    cfg = base.with_(precision=A.PRECISION_MIXED)
    ba.solve(prob, cfg.with_(max_iterations=1))
    sol = ba.solve(prob, cfg.with_(max_iterations=steps))
    out = {"lm_iter_per_s": sol.report.iterations / sol.report.seconds_iterations,
           "ms_per_step": 1e3 * sol.report.seconds_iterations / max(sol.report.iterations, 1),
           "pcg_iterations": int(sol.report.pcg_iterations)}
    if ttc is not None:
        conv = ba.solve(prob, cfg.with_(max_iterations=ttc["fp64_iterations"] + 20))
        target = ttc["converged_cost"] * (1.0 + 1e-6)
        tr = conv.trace
        hit = None
        for i in range(len(tr["seconds"])):
            if tr["accepted"][i] and tr["candidate_cost"][i] <= target:
                hit = float(tr["seconds"][i])
                break
        out.update(time_to_cost_seconds=hit, iterations=int(conv.report.iterations),
                   final_cost=float(conv.report.final_cost),
                   final_cost_rel_error_vs_fp64=abs(conv.report.final_cost - ttc["converged_cost"]) / ttc["converged_cost"],
                   speedup_vs_fp64=(ttc["fp64_seconds"] / hit) if (hit and ttc["fp64_seconds"]) else None)
    return out


def bench_medium(ba) -> dict:
    """Direct form on the 200-camera / 50,000-point scene (D = 1,793): LM iterations/s over ten
    iterations in FP64 and mixed precision, and the dense Cholesky (kernel d) rate on the matrix
    cores from HIP events around its launches (D^3/3 + 2 D^2 flops per factor and solve)."""
    from openmvg_ba import SolveConfig, _abi as A, synthetic

    scene = synthetic.medium_scene()
    prob = scene.problem
    cfg = SolveConfig(max_iterations=10, linear_solver=A.SOLVER_DIRECT)
    out = {"workload": f"{prob.n_poses} cameras, {prob.n_points} points, {prob.n_observations} observations, "
                       "free f/k1/k2, squared loss, direct Schur + dense Cholesky, seed 0xBA0002"}
    ba.solve(prob, cfg.with_(max_iterations=1))
    for name, prec in (("fp64", A.PRECISION_FP64), ("mixed", A.PRECISION_MIXED)):
        sol = ba.solve(prob, cfg.with_(precision=prec, profile_kernel=A.PROFILE_CHOLESKY))
        rep = sol.report
        D = rep.reduced_dimension
        avg = rep.profile_seconds / max(rep.profile_launches, 1)
        flops = D ** 3 / 3.0 + 2.0 * D * D
        out[name] = {"lm_iter_per_s": rep.iterations / rep.seconds_iterations,
                     "ms_per_step": 1e3 * rep.seconds_iterations / max(rep.iterations, 1),
                     "cholesky_ms": 1e3 * avg, "cholesky_tflops": flops / avg / 1e12,
                     "cholesky_peak_tflops": 78.6 if name == "fp64" else 157.3,
                     "final_cost": rep.final_cost, "precision_fallbacks": int(rep.precision_fallbacks)}
    return out


def bench_batched(ba, sample_windows: int = 64) -> dict:
    """512 local windows (seed 0xBA0004: U{5..40} cameras, U{200..5000} points, 2-8 views), ten LM
    iterations each with the direct step: one batched solve (graph-replayed iterations) in FP64 and
    mixed precision, against one solve per window (the first `sample_windows` windows, scaled)."""
    from openmvg_ba import SolveConfig, _abi as A, synthetic

    probs = [w.problem for w in synthetic.window_scenes()]
    cfg = SolveConfig(max_iterations=10, linear_solver=A.SOLVER_DIRECT)
    out = {"workload": f"{len(probs)} windows, {sum(p.n_observations for p in probs)} observations, "
                       "10 LM iterations each"}
    for name, prec in (("fp64", A.PRECISION_FP64), ("mixed", A.PRECISION_MIXED)):
        sols, info = ba.solve_batched(probs, cfg.with_(precision=prec), mode=A.BATCH_GRAPH)
        its = sum(s.report.iterations for s in sols)
        out[name] = {"loop_seconds": info["seconds_loop"], "window_iterations_per_s": its / info["seconds_loop"],
                     "tile": info["tile"], "autotune_seconds": list(info["autotune_seconds"])}
    loop = 0.0
    its = 0
    for p in probs[:sample_windows]:
        s = ba.solve(p, cfg)
        loop += s.report.seconds_iterations
        its += s.report.iterations
    out["per_window_launches"] = {"window_iterations_per_s": its / loop,
                                  "sample": f"first {sample_windows} windows solved one by one (FP64)"}
    out["speedup_batched_vs_per_window"] = out["fp64"]["window_iterations_per_s"] / (its / loop)
    return out


def bench_differentiation(steps: int) -> dict:
    """Forward+backward observations/s of the differentiable reprojection cost: 1,000 cameras,
    200,000 points, 1,000,000 observations, Huber 1 px (seed 0xBA0005), FP32 and FP64, through
    the torch.autograd.Function (forward call + backward call) on torch's current stream."""
    import torch

    from openmvg_ba import _abi as A, synthetic
    from openmvg_ba.autodiff import ReprojectionTopology, reprojection_cost

    cams, pts, oc, op, uv = synthetic.bal_scene()
    topo = ReprojectionTopology(oc, op, len(cams), len(pts))
    out = {"workload": f"{len(cams)} cameras, {len(pts)} points, {len(oc)} observations, Huber 1px, seed 0xBA0005"}
    for name, dtype in (("fp32", torch.float32), ("fp64", torch.float64)):
        tc = torch.tensor(cams, device="cuda", dtype=dtype, requires_grad=True)
        tp = torch.tensor(pts, device="cuda", dtype=dtype, requires_grad=True)
        tuv = torch.tensor(uv, device="cuda", dtype=dtype)

        def step():
            tc.grad = None
            tp.grad = None
            reprojection_cost(tc, tp, tuv, topo, A.LOSS_HUBER, 1.0).backward()

        for _ in range(3):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3 / steps
        out[name] = {"obs_per_s": len(oc) / sec, "ms_per_forward_backward": sec * 1e3}
    topo.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed LM iterations")
    ap.add_argument("--warmup", type=int, default=2, help="untimed LM iterations (a separate solve)")
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--roofline-kernel", default="auto",
                    choices=["auto", "pcg_product", "pcg_points", "pcg_cameras", "linearize"],
                    help="auto: the PCG product kernel of the form the solve uses")
    ap.add_argument("--converge-iterations", type=int, default=60,
                    help="iteration cap of the run that fixes the converged cost (0 skips time-to-cost)")
    ap.add_argument("--cpu-points", type=int, default=200_000, help="points of the CPU-baseline subsample")
    ap.add_argument("--cpu-iterations", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the other forms' measurements")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES between two device buffers three times (a known-size kernel that "
                         "scales HBM counters in a profiled run)")
    ap.add_argument("--only-form", choices=["medium_direct", "batched", "differentiation", "mixed_large"],
                    help="run only this secondary measurement and print its JSON (no headline line)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU point-sharded design is described, not built)", file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU (torch.cuda.is_available() is False)", file=sys.stderr)
        return 1
    from openmvg_ba import HipBundleAdjuster, SolveConfig, _abi as A
    from openmvg_ba import synthetic

    if args.only_form:
        fn = {"medium_direct": lambda: bench_medium(HipBundleAdjuster(0)),
              "mixed_large": lambda: bench_mixed(HipBundleAdjuster(0), synthetic.large_scene(n_points=args.points).problem,
                                                 SolveConfig(linear_solver=A.SOLVER_ITERATIVE), args.steps, None),
              "batched": lambda: bench_batched(HipBundleAdjuster(0)),
              "differentiation": lambda: bench_differentiation(20)}[args.only_form]
        print(json.dumps({args.only_form: fn()}))
        return 0

    if args.copy_calibration > 0:
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {4 * n} bytes")
        del src, dst

    t0 = time.perf_counter()
    scene = synthetic.large_scene(n_points=args.points)
    prob = scene.problem
    log(f"scene: {prob.n_poses} cameras, {prob.n_points} points, {prob.n_observations} observations "
        f"({time.perf_counter() - t0:.1f} s)")
    ba = HipBundleAdjuster(0)
    base = SolveConfig(linear_solver=A.SOLVER_ITERATIVE)

    # warmup: a separate short solve (module load, first-touch allocations)
    if args.warmup > 0:
        t0 = time.perf_counter()
        w = ba.solve(prob, base.with_(max_iterations=args.warmup))
        log(f"warmup: {w.report.iterations} iterations, loop {w.report.seconds_iterations:.3f} s, "
            f"setup {w.report.seconds_setup:.2f} s, call {time.perf_counter() - t0:.1f} s")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol = ba.solve(prob, base.with_(max_iterations=args.steps))
    torch.cuda.synchronize()
    call_s = time.perf_counter() - t0
    rep = sol.report
    steps = int(rep.iterations)
    loop_s = float(rep.seconds_iterations)
    D = int(rep.reduced_dimension)
    log(f"timed: {steps} iterations in {loop_s:.3f} s ({steps / loop_s:.2f} iter/s), setup {rep.seconds_setup:.2f} s, "
        f"call {call_s:.1f} s, PCG iterations {rep.pcg_iterations}, cost {rep.initial_cost:.6e} -> "
        f"{rep.final_cost:.6e}, status {rep.status}")

    form = {A.SCHUR_EXPLICIT: "explicit", A.SCHUR_IMPLICIT: "implicit"}.get(
        ba.handle.lib.ba_hip_last_schur_product(ba.handle.ptr), "unknown")
    if args.roofline_kernel == "auto":
        args.roofline_kernel = "pcg_product" if form == "explicit" else "pcg_points"
    # roofline of the dominant kernel: HIP events around each of its launches
    kernel_ids = {"pcg_points": A.PROFILE_PCG_POINTS, "pcg_cameras": A.PROFILE_PCG_CAMERAS,
                  "linearize": A.PROFILE_LINEARIZE, "pcg_product": A.PROFILE_PCG_PRODUCT}
    # the profiled kernel's form for every iteration of that solve (AUTO keeps the first,
    # few-PCG-iteration LM iterations matrix-free)
    prof_mode = {"pcg_product": A.SCHUR_EXPLICIT, "pcg_points": A.SCHUR_IMPLICIT,
                 "pcg_cameras": A.SCHUR_IMPLICIT}.get(args.roofline_kernel, A.SCHUR_AUTO)
    ba.handle.lib.ba_hip_set_schur_product(ba.handle.ptr, prof_mode)
    prof = ba.solve(prob, base.with_(max_iterations=max(2, min(args.steps, 4)),
                                     profile_kernel=kernel_ids[args.roofline_kernel]))
    ba.handle.lib.ba_hip_set_schur_product(ba.handle.ptr, A.SCHUR_AUTO)
    launches = int(prof.report.profile_launches)
    avg_s = float(prof.report.profile_seconds) / max(launches, 1)
    counts = lowered_counts(prob)
    if args.roofline_kernel == "pcg_product":
        counts.update(schur_counts(prob))
    per_launch = algorithmic_bytes(args.roofline_kernel, counts, D)
    achieved = per_launch / avg_s / 1e9
    log(f"roofline {args.roofline_kernel}: {launches} launches, avg {avg_s * 1e6:.1f} us, "
        f"{per_launch / 1e6:.1f} MB/launch, {achieved:.1f} GB/s")

    # time-to-cost: first point at which an accepted state is within 1e-6 of the converged cost
    ttc = None
    if args.converge_iterations > 0:
        conv = ba.solve(prob, base.with_(max_iterations=args.converge_iterations))
        c_star = float(conv.report.final_cost)
        tr = conv.trace
        target = c_star * (1.0 + 1e-6)
        hit = None
        for i in range(len(tr["seconds"])):
            if tr["accepted"][i] and tr["candidate_cost"][i] <= target:
                hit = float(tr["seconds"][i])
                break
        ttc = {"fp64_seconds": hit, "fp64_iterations": int(conv.report.iterations),
               "converged_cost": c_star, "status": conv.report.status,
               "loop_seconds": float(conv.report.seconds_iterations), "mixed_seconds": None,
               "reference_cost": "FP64 GPU run to convergence (the FP64 path matches the oracle per iteration)"}
        log(f"converge: {conv.report.iterations} iterations, {conv.report.status}, cost {c_star:.9e}, "
            f"time-to-cost {hit}")

    forms = {}
    if not args.no_forms:
        forms["mixed"] = bench_mixed(ba, prob, base, args.steps, ttc)
        log(f"mixed: {forms['mixed']}")
        forms["medium_direct"] = bench_medium(ba)
        log(f"medium direct: {forms['medium_direct']}")
        forms["batched"] = bench_batched(ba)
        log(f"batched: {forms['batched']}")
        forms["differentiation"] = bench_differentiation(20)
        log(f"differentiation: {forms['differentiation']}")

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(synthetic.large_scene, args.cpu_points, prob.n_observations, args.cpu_iterations)
        log(f"cpu baseline: {cpu['value']:.4f} iter/s ({time.perf_counter() - t0:.1f} s)")

    value = steps / loop_s
    line = {
        "metric": "LM iter/s (2k-cam/1M-pt BA, FP64 GPU)",
        "value": value,
        "unit": "LM iter/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * loop_s / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"large BA: {prob.n_poses} cameras (f,k1,k2 free), {prob.n_points} points, "
                         f"{prob.n_observations} observations, geometric tracks, 2% outliers, Huber 1px, "
                         "LM + block-Jacobi PCG (rtol 0.1) on the Schur system: explicit in the LM iterations "
                         "after a PCG solve of 10 or more iterations, matrix-free otherwise (AUTO), seed 0xBA0003"),
            "reduced_dimension": D,
            "precision": "fp64",
            "schur_product": "auto" + (" (explicit system used)" if form == "explicit" else " (matrix-free only)"),
            "pcg_iterations": int(rep.pcg_iterations),
        },
        "roofline": {
            "kernel": args.roofline_kernel,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.roofline_kernel, per_launch),
            "launches": launches,
            "avg_launch_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": per_launch,
        },
        "cpu_baseline": cpu,
        "time_to_cost": ttc,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
