#!/usr/bin/env python3
"""LM iterations/s of the HIP bundle adjuster on the 2,000-camera / 1,000,000-point scene.

Workload: the large synthetic scene (seed 0xBA0003): 2,000 cameras with free f, k1, k2; 1M points;
~6.1M observations with geometric track lengths (mean 6, clipped to [2, 50]); 2 % outliers; Huber
loss with delta 1 px; inexact LM with block-Jacobi PCG on the Schur system (forcing tolerance 0.1,
the native cap of min(600, D) PCG iterations per solve). One step is one LM iteration:
linearization, point elimination, block assembly, the PCG solve and the candidate pass
(sparse/self_calibration/mod.rs:212-639).

The timed region is the LM loop as the library reports it (`seconds_iterations`): it starts after
the problem has been lowered and uploaded, so every input is resident in HBM, and every iteration
ends in a synchronous scalar readback. The printed JSON line carries:

* ``kernels``: every hot kernel's launches, average duration (HIP events the library records
  around that kernel's launches on the solve's stream, one profiled solve per kernel with the same
  iterations and product choice as the timed one), total time, algorithmic bytes and fraction of
  the HBM peak; ``roofline`` is the kernel with the largest total time;
* ``cpu_baseline``: the oracle (the C++ restatement of the same LM path, OpenMP over point shards
  fixed by the problem, so the result does not depend on the thread count) on the full scene on
  all host cores: LM iterations/s over the same first iterations, the time to reach its own
  converged cost, and one iteration on one thread;
* ``time_to_cost``: seconds until an accepted state is within 1e-6 relative of the oracle's
  converged cost C* (from the CPU leg, or the committed oracle fixture when it is skipped).
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
    point_elim (k_point_elim): per observation r~ and J~_X (64 B), cost and squared residual
      (16 B) and the validity flag (1 B); per point its offset and fixed flag (5 B) and the written
      six scalars (48 B), y = V^-1 g_p and the damping diagonal (24 B each), V^ and V^-1 (48 B
      each) and the packed 128-byte record of the camera-side assembly: 325 B.
    candidate (k_candidate): per observation the camera and point index (8 B), pixel (16 B),
      J~_X and three values of the compact state (72 B); per point its offset and fixed flag
      (5 B), X read and the candidate written (24 B each), V^-1 (48 B), y and the damping
      diagonal (24 B each) and the written scalars (32 B): 181 B.
    pcg_product (k_pcg_product_upper: the first pass of q = S v over the stored diagonal and
      upper blocks of the explicit reduced system, one workgroup per camera row): per stored slot
      the 9 x 9 block (648 B) and its neighbour index (4 B); per camera pair the written column
      sums (72 B); per camera the 9 input values (72 B), its offsets (16 B) and the written row
      sums (72 B).
    schur_pairs (k_schur_pairs: the off-diagonal blocks of the explicit S, rebuilt from compact
      records): each observation's 16-value record once (128 B), 8 B per observation pair, and per
      camera pair the written 9 x 9 block and its indices (656 B).
    camera_assembly (k_camera_assembly with the explicit system): per observation the
      camera-major rows read (72 B) and point index (4 B) and the compact record written (128 B);
      per point its packed record (128 B); per segment its camera (4 B) and written sums (576 B);
      per camera R, f, k1, k2 (120 B).
    """
    no, npt, nc = counts["obs"], counts["points"], counts["cameras"]
    if kernel == "schur_pairs":
        return no * 128 + counts["terms"] * 8 + counts["pairs"] * (648 + 8)
    if kernel == "camera_assembly":
        return no * (72 + 4 + 128) + npt * 128 + counts["segments"] * (4 + 576) + nc * 120
    if kernel == "pcg_product":
        return counts["slots"] * (648 + 4) + counts["pairs"] * 72 + nc * (72 + 16 + 72)
    if kernel == "pcg_points":
        return no * (48 + 24 + 8) + npt * (4 + 1 + 48 + 24 + 32) + nc * 72
    if kernel == "pcg_cameras":
        return no * (56 + 4) + npt * 32 + counts["segments"] * (4 + 72) + nc * (12 + 192)
    if kernel == "linearize":
        return no * (16 + 8 + 96 + 17) + npt * 24 + nc * 120
    if kernel == "point_elim":
        return no * 81 + npt * 325
    if kernel == "candidate":
        return no * 96 + npt * 181
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


def host_cpu() -> dict:
    """The host's logical CPUs and model name (the CPU baseline's cores)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count() or 1, "model": model}


def omp_threads(n: int | None = None) -> int:
    """Set (when n is given) and return the OpenMP thread count of this process's runtime, the one
    the oracle library uses."""
    import ctypes

    gomp = ctypes.CDLL("libgomp.so.1")
    if n is not None:
        gomp.omp_set_num_threads(int(n))
    return int(gomp.omp_get_max_threads())


def load_oracle_lib():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import load_oracle

    return load_oracle()


def first_hit(trace: dict, target: float):
    """Loop seconds at the first accepted iteration whose candidate cost is at most `target`."""
    for i in range(len(trace["seconds"])):
        if trace["accepted"][i] and trace["candidate_cost"][i] <= target:
            return float(trace["seconds"][i]), i + 1
    return None, None


def cpu_baseline(prob, steps: int, converge_iterations: int) -> dict:
    """The oracle on the full scene: the run to convergence on all host cores (LM iterations/s over
    the first `steps` iterations, the same ones the GPU line times; its converged cost C*; the
    time to within 1e-6 of C*) and the first iteration on one thread."""
    from openmvg_ba import SolveConfig, _abi as A
    from openmvg_ba.problem import solve_with

    oracle = load_oracle_lib()
    cpu = host_cpu()
    threads = omp_threads(cpu["nproc"]) if "OMP_NUM_THREADS" not in os.environ else omp_threads()
    cfg = SolveConfig(max_iterations=converge_iterations, linear_solver=A.SOLVER_ITERATIVE,
                      certification=A.CERTIFY_LOOP_ONLY)
    t0 = time.perf_counter()
    sol = solve_with(oracle, "bao_solve", prob, cfg)
    wall = time.perf_counter() - t0
    tr = sol.trace
    n = min(steps, len(tr["seconds"]))
    c_star = float(sol.report.final_cost)
    ttc, ttc_it = first_hit(tr, c_star * (1.0 + 1e-6))
    omp_threads(1)
    one = solve_with(oracle, "bao_solve", prob,
                     SolveConfig(max_iterations=1, linear_solver=A.SOLVER_ITERATIVE, certification=A.CERTIFY_LOOP_ONLY))
    omp_threads(threads)
    one_s = float(one.trace["seconds"][0])
    return {
        "value": n / float(tr["seconds"][n - 1]),
        "unit": "LM iter/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (C++ FP64 restatement of the same LM path, OpenMP over point shards fixed by the "
                   f"problem: same result at any thread count) on the full scene, {threads} threads on "
                   f"{cpu['nproc']} logical CPUs ({cpu['model']}); value = the first {n} LM iterations of its run to "
                   f"convergence ({int(tr['pcg_iterations'][:n].sum())} PCG iterations); whole call {wall:.1f} s"),
        "nproc": cpu["nproc"],
        "cpu_model": cpu["model"],
        "converged_cost": c_star,
        "converge_iterations": int(sol.report.iterations),
        "status": str(sol.report.status),
        "time_to_cost_seconds": ttc,
        "time_to_cost_iterations": ttc_it,
        "loop_seconds": float(tr["seconds"][-1]),
        "one_thread_first_iteration_seconds": one_s,
        "all_threads_first_iteration_seconds": float(tr["seconds"][0]),
    }


def kernel_table(ba, prob, base, steps: int, D: int) -> list:
    """Per hot kernel: a solve like the timed one (same iterations, AUTO product choice) with HIP
    events around that kernel's launches; its launches, average and total time, algorithmic bytes
    per launch and achieved fraction of the HBM peak.  PCG launches issued after the PCG state
    turned done return at once and are not counted."""
    from openmvg_ba import _abi as A

    counts = lowered_counts(prob)
    counts.update(schur_counts(prob))
    ids = {"schur_pairs": A.PROFILE_SCHUR_PAIRS, "camera_assembly": A.PROFILE_PAIRS,
           "pcg_product": A.PROFILE_PCG_PRODUCT, "pcg_points": A.PROFILE_PCG_POINTS,
           "pcg_cameras": A.PROFILE_PCG_CAMERAS, "linearize": A.PROFILE_LINEARIZE,
           "point_elim": A.PROFILE_POINT_ELIM, "candidate": A.PROFILE_CANDIDATE}
    rows = []
    for name, kid in ids.items():
        sol = ba.solve(prob, base.with_(max_iterations=steps, profile_kernel=kid))
        launches = int(sol.report.profile_launches)
        total = float(sol.report.profile_seconds)
        if launches == 0:
            continue
        avg = total / launches
        per_launch = algorithmic_bytes(name, counts, D)
        achieved = per_launch / avg / 1e9
        rows.append({"kernel": name, "launches": launches, "avg_launch_us": avg * 1e6, "total_ms": total * 1e3,
                     "algorithmic_bytes_per_launch": per_launch, "achieved_gbs": achieved,
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(name, per_launch)})
        log(f"kernel {name}: {launches} launches, avg {avg * 1e6:.1f} us, total {total * 1e3:.2f} ms, "
            f"{achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.3f})")
    rows.sort(key=lambda r: -r["total_ms"])
    return rows


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
        c_star = ttc["oracle_converged_cost"]
        hit, hit_it = first_hit(conv.trace, c_star * (1.0 + 1e-6))
        out.update(time_to_cost_seconds=hit, time_to_cost_iterations=hit_it, iterations=int(conv.report.iterations),
                   final_cost=float(conv.report.final_cost),
                   final_cost_rel_error_vs_oracle=abs(conv.report.final_cost - c_star) / c_star,
                   speedup_vs_fp64=(ttc["fp64_seconds"] / hit) if (hit and ttc["fp64_seconds"]) else None)
    return out


def bench_medium(ba, cpu: bool = False) -> dict:
    """Direct form on the 200-camera / 50,000-point scene (D = 1,793): LM iterations/s over ten
    iterations in FP64 and mixed precision, and the dense Cholesky (kernel d) rate on the matrix
    cores from HIP events around its launches (D^3/3 + 2 D^2 flops per factor and solve)."""
    from openmvg_ba import SolveConfig, _abi as A, synthetic

    scene = synthetic.medium_scene()
    prob = scene.problem
    cfg = SolveConfig(max_iterations=10, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY)
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
    if cpu:
        from openmvg_ba.problem import solve_with

        sol = solve_with(load_oracle_lib(), "bao_solve", prob, cfg)
        loop = float(sol.trace["seconds"][-1])
        out["cpu"] = {"lm_iter_per_s": sol.report.iterations / loop, "threads": omp_threads(),
                      "final_cost": float(sol.report.final_cost), "kind": "port",
                      "sample": "oracle, the same 10-iteration direct solve, all host cores"}
    return out


def bench_calibrated(ba, scene, steps: int, cpu: bool = False) -> dict:
    """The calibrated pixel path on the large scene with every intrinsic fixed (native/calibrated.rs):
    Auto takes the retained PCG step by the workload rule (at least 250,000 observations and
    D >= 768); LM iterations/s over `steps` iterations in FP64, and with `cpu` the oracle's same
    solve on all host cores and the largest per-iteration cost difference between the two."""
    from openmvg_ba import SolveConfig, _abi as A, synthetic

    prob = synthetic.fixed_intrinsics(scene).problem
    cfg = SolveConfig(max_iterations=steps, certification=A.CERTIFY_LOOP_ONLY)
    ba.solve(prob, cfg.with_(max_iterations=2))
    sol = ba.solve(prob, cfg)
    rep = sol.report
    out = {"workload": f"{prob.n_poses} cameras (intrinsics fixed: calibrated path), {prob.n_points} points, "
                       f"{prob.n_observations} observations, Huber 1px, Auto camera step, seed 0xBA0003",
           "reduced_dimension": int(rep.reduced_dimension),
           "linear_solver": "pcg" if rep.linear_solver_used == A.SOLVER_ITERATIVE else "direct",
           "plan_reason": int(rep.plan_reason),
           "fp64": {"lm_iter_per_s": rep.iterations / rep.seconds_iterations,
                    "ms_per_step": 1e3 * rep.seconds_iterations / max(rep.iterations, 1),
                    "pcg_iterations": int(rep.pcg_iterations), "final_cost": float(rep.final_cost)}}
    if cpu:
        from openmvg_ba.problem import solve_with

        want = solve_with(load_oracle_lib(), "bao_solve", prob, cfg)
        loop = float(want.trace["seconds"][-1])
        n = min(len(want.trace["cost"]), len(sol.trace["cost"]))
        c0, c1 = np.asarray(sol.trace["cost"][:n]), np.asarray(want.trace["cost"][:n])
        out["cpu"] = {"lm_iter_per_s": want.report.iterations / loop, "threads": omp_threads(), "kind": "port",
                      "final_cost": float(want.report.final_cost),
                      "sample": f"oracle, the same {steps}-iteration solve, all host cores"}
        out["max_iteration_cost_rel_diff_vs_oracle"] = float(np.max(np.abs(c0 - c1) / np.abs(c1)))
        out["same_accept_sequence"] = bool(np.array_equal(sol.trace["accepted"][:n], want.trace["accepted"][:n]))
    return out


def bench_batched(ba, sample_windows: int = 64, cpu: bool = False) -> dict:
    """512 local windows (seed 0xBA0004: U{5..40} cameras, U{200..5000} points, 2-8 views), ten LM
    iterations each with the direct step: one batched solve (graph-replayed iterations) in FP64 and
    mixed precision, against one solve per window (the first `sample_windows` windows, scaled)."""
    from openmvg_ba import SolveConfig, _abi as A, synthetic

    probs = [w.problem for w in synthetic.window_scenes()]
    cfg = SolveConfig(max_iterations=10, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY)
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
    if cpu:
        # the oracle solves the windows one by one, one window per host thread (each window's solve
        # is sequential: fewer than 32,768 observations make one point shard)
        from concurrent.futures import ThreadPoolExecutor

        from openmvg_ba.problem import solve_with

        oracle = load_oracle_lib()
        workers = os.cpu_count() or 1
        t0 = time.perf_counter()
        with ThreadPoolExecutor(workers) as pool:
            sols = list(pool.map(lambda p: solve_with(oracle, "bao_solve", p, cfg), probs))
        wall = time.perf_counter() - t0
        its = sum(x.report.iterations for x in sols)
        out["cpu"] = {"window_iterations_per_s": its / wall, "wall_seconds": wall, "threads": workers,
                      "kind": "port", "sample": "oracle, every window's 10-iteration direct solve, one window per "
                                                f"host thread ({workers} threads), whole calls"}
    return out


def bench_differentiation(steps: int, cpu: bool = False) -> dict:
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
    if cpu:
        import ctypes as C

        oracle = load_oracle_lib()
        c64, p64, uv64 = (np.ascontiguousarray(x, dtype=np.float64) for x in (cams, pts, uv))
        oc32, op32 = np.ascontiguousarray(oc, dtype=np.uint32), np.ascontiguousarray(op, dtype=np.uint32)
        gc, gp = np.zeros_like(c64), np.zeros_like(p64)
        cost, inv = C.c_double(), C.c_uint32()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            oracle.bao_cost_grad(len(c64), len(p64), len(oc32), c64.ctypes.data, p64.ctypes.data, oc32.ctypes.data,
                                 op32.ctypes.data, uv64.ctypes.data, A.LOSS_HUBER, 1.0, C.byref(cost), C.byref(inv),
                                 gc.ctypes.data, gp.ctypes.data)
        sec = (time.perf_counter() - t0) / reps
        out["cpu"] = {"obs_per_s": len(oc) / sec, "threads": 1, "kind": "port",
                      "sample": "oracle cost and analytic gradient (FP64), one thread, mean of 3 calls"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed LM iterations")
    ap.add_argument("--warmup", type=int, default=2, help="untimed LM iterations (a separate solve)")
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--profile-kernel", default="all",
                    choices=["all", "none", "schur_pairs", "camera_assembly", "pcg_product", "pcg_points",
                             "pcg_cameras", "linearize", "point_elim", "candidate"],
                    help="kernels to time with HIP events (all: the per-kernel table; the roofline is the one "
                         "with the largest total time)")
    ap.add_argument("--schur-product", default="auto", choices=["auto", "explicit", "implicit"],
                    help="PCG product form of every solve (auto: the library's per-iteration choice)")
    ap.add_argument("--converge-iterations", type=int, default=60,
                    help="iteration cap of the runs to convergence (0 skips time-to-cost)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle's CPU legs")
    ap.add_argument("--no-forms", action="store_true", help="skip the other forms' measurements")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES between two device buffers three times (a known-size kernel that "
                         "scales HBM counters in a profiled run)")
    ap.add_argument("--only-form", choices=["medium_direct", "batched", "differentiation", "mixed_large",
                                            "calibrated_large"],
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

    product = {"auto": A.SCHUR_AUTO, "explicit": A.SCHUR_EXPLICIT, "implicit": A.SCHUR_IMPLICIT}[args.schur_product]

    def adjuster():
        b = HipBundleAdjuster(0)
        b.handle.lib.ba_hip_set_schur_product(b.handle.ptr, product)
        return b

    if args.only_form:
        fn = {"medium_direct": lambda: bench_medium(adjuster(), cpu=not args.no_cpu),
              "mixed_large": lambda: bench_mixed(adjuster(), synthetic.large_scene(n_points=args.points).problem,
                                                 SolveConfig(linear_solver=A.SOLVER_ITERATIVE,
                                                             certification=A.CERTIFY_LOOP_ONLY), args.steps, None),
              "calibrated_large": lambda: bench_calibrated(adjuster(), synthetic.large_scene(n_points=args.points),
                                                           args.steps, cpu=not args.no_cpu),
              "batched": lambda: bench_batched(adjuster(), cpu=not args.no_cpu),
              "differentiation": lambda: bench_differentiation(20, cpu=not args.no_cpu)}[args.only_form]
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
    ba = adjuster()
    base = SolveConfig(linear_solver=A.SOLVER_ITERATIVE, certification=A.CERTIFY_LOOP_ONLY)

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

    # per-kernel table (HIP events) and the roofline of the kernel with the largest total time
    table = []
    if args.profile_kernel == "all":
        table = kernel_table(ba, prob, base, steps, D)
    elif args.profile_kernel != "none":
        table = [r for r in kernel_table(ba, prob, base, steps, D) if r["kernel"] == args.profile_kernel]
    roofline = None
    if table:
        top = table[0]
        roofline = {"kernel": top["kernel"], "bound": "hbm", "achieved": top["achieved_gbs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": top["frac"], "traffic": top["traffic"], "launches": top["launches"],
                    "avg_launch_us": top["avg_launch_us"],
                    "algorithmic_bytes_per_launch": top["algorithmic_bytes_per_launch"],
                    "total_ms_in_timed_iterations": top["total_ms"]}

    # the oracle's run on all host cores: CPU baseline and the converged cost C*
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(prob, steps, args.converge_iterations or 60)
        log(f"cpu baseline: {cpu['value']:.4f} iter/s, C* {cpu['converged_cost']:.10e} in {cpu['converge_iterations']} "
            f"iterations, time-to-cost {cpu['time_to_cost_seconds']} s ({time.perf_counter() - t0:.1f} s)")
    if cpu is not None:
        c_star, c_src = cpu["converged_cost"], "oracle run to convergence in this bench (CPU leg)"
    else:
        with open(os.path.join(ROOT, "tests", "golden", "large_pcg.json")) as fh:
            c_star = float(json.load(fh)["converged"]["final_cost"])
        c_src = "oracle run to convergence, committed fixture tests/golden/large_pcg.json"

    # time-to-cost: first accepted state within 1e-6 of the oracle's converged cost
    ttc = None
    if args.converge_iterations > 0:
        conv = ba.solve(prob, base.with_(max_iterations=args.converge_iterations))
        hit, hit_it = first_hit(conv.trace, c_star * (1.0 + 1e-6))
        ttc = {"fp64_seconds": hit, "fp64_iterations_to_cost": hit_it, "fp64_iterations": int(conv.report.iterations),
               "fp64_final_cost": float(conv.report.final_cost), "oracle_converged_cost": c_star,
               "fp64_final_cost_rel_error_vs_oracle": abs(conv.report.final_cost - c_star) / c_star,
               "status": conv.report.status, "loop_seconds": float(conv.report.seconds_iterations),
               "reference_cost": c_src,
               "cpu_seconds": cpu["time_to_cost_seconds"] if cpu else None}
        log(f"converge: {conv.report.iterations} iterations, {conv.report.status}, cost "
            f"{conv.report.final_cost:.10e} (C* {c_star:.10e}), time-to-cost {hit}")

    forms = {}
    if not args.no_forms:
        forms["mixed"] = bench_mixed(ba, prob, base, args.steps, ttc)
        log(f"mixed: {forms['mixed']}")
        if ttc is not None:
            ttc["mixed_seconds"] = forms["mixed"].get("time_to_cost_seconds")
        forms["calibrated"] = bench_calibrated(ba, scene, args.steps, cpu=not args.no_cpu)
        log(f"calibrated: {forms['calibrated']}")
        forms["medium_direct"] = bench_medium(ba, cpu=not args.no_cpu)
        log(f"medium direct: {forms['medium_direct']}")
        forms["batched"] = bench_batched(ba, cpu=not args.no_cpu)
        log(f"batched: {forms['batched']}")
        forms["differentiation"] = bench_differentiation(20, cpu=not args.no_cpu)
        log(f"differentiation: {forms['differentiation']}")

    value = steps / loop_s
    pcg_cap = min(600, D)
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
                         "LM + block-Jacobi PCG (forcing tolerance 0.1) on the Schur system, seed 0xBA0003"),
            "reduced_dimension": D,
            "precision": "fp64",
            "pcg_cap": f"min(600, D) = {pcg_cap} PCG iterations per solve (the native cap; "
                       "self_calibration_retained.rs:527-531)",
            "schur_product": f"{args.schur_product} (" + ("explicit system used" if form == "explicit"
                                                         else "matrix-free only") + ")",
            "pcg_iterations": int(rep.pcg_iterations),
            "seconds_setup": float(rep.seconds_setup),
        },
        "roofline": roofline,
        "kernels": table,
        "cpu_baseline": cpu,
        "time_to_cost": ttc,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
