"""Writes the golden fixtures in tests/golden/ from the oracle.

    python tests/golden/make_golden.py            # the ring-scene fixtures (seconds)
    python tests/golden/make_golden.py --scale    # the medium, large and window fixtures (minutes;
                                                  # OMP_NUM_THREADS threads, same bits at any count)

Every fixture is oracle-generated (the oracle restates the reference algorithm; its header lists
the reference files it follows) on the seeded 10-camera ring scene, and records a SHA-256 of the
scene's input arrays so that a test regenerating the scene can tell it is the same one.  Files:

  ring_index_maps.json     lowered layout and index maps (integers)
  ring_linearize.json      per-observation r~, J~_X, J~_c, cost and validity, first 64
                           observations, squared and Huber (delta 1) losses
  ring_schur.json          reduced system S, reduced and full gradient, damping diagonal and the
                           assembly scalars at lambda = 1e-4 (squared loss)
  ring_solve.json          LM trajectories (cost, damping, accepted) and final parameters of the
                           direct and the PCG solve, 50 iterations (squared loss)
  medium_direct.json       200-camera / 50,000-point scene (synthetic.medium_scene), direct step,
                           10 LM iterations: trajectory, final cost, final camera parameters and
                           every 97th point; and the run to convergence (at most 50 iterations)
  large_pcg.json           2,000-camera / 1M-point scene (synthetic.large_scene), PCG step: the
                           first 5 LM iterations (trajectory, PCG iterations, final cost) and the
                           run to convergence (at most 60 iterations, the bench's converged-cost
                           run): its trajectory and converged cost C*
  windows.json             the 512-window batch (synthetic.window_scenes): direct-step solves of 10
                           iterations of the 4 largest windows and 4 seeded others
"""

from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "openmvg-rs_amd"))

from golden_lib import scene_digest  # noqa: E402
from helpers import host_view, native_host, oracle_assemble  # noqa: E402
from oracle_lib import load_oracle  # noqa: E402
from openmvg_ba import SolveConfig, synthetic  # noqa: E402
from openmvg_ba import _abi as A  # noqa: E402
from openmvg_ba.problem import solve_with  # noqa: E402

N_OBS = 64
SOLVE_ITERATIONS = 50


def floats(a) -> list:
    return [float(x) for x in np.asarray(a, dtype=np.float64).ravel()]


def ints(a) -> list:
    return [int(x) for x in np.asarray(a).ravel()]


def linearize(oracle, nat, arrays, loss, scale) -> dict:
    no = nat.sizes["n_observations"]
    hv = host_view(nat, arrays, loss, scale)
    r, j, c, k, v = np.zeros(2 * no), np.zeros(6 * no), np.zeros(18 * no), np.zeros(no), np.zeros(no, np.uint8)
    oracle.bao_linearize(C.byref(hv), *(x.ctypes.data_as(C.c_void_p) for x in (r, j, c, k, v)))
    n = min(N_OBS, no)
    return dict(residual=floats(r[:2 * n]), point_jacobian=floats(j[:6 * n]), camera_jacobian=floats(c[:18 * n]),
                cost=floats(k[:n]), valid=ints(v[:n]))


def solve(oracle, problem, solver) -> dict:
    sol = solve_with(oracle, "bao_solve", problem, SolveConfig(max_iterations=SOLVE_ITERATIONS, linear_solver=solver))
    return dict(cost=floats(sol.trace["cost"]), damping=floats(sol.trace["damping"]),
                accepted=ints(sol.trace["accepted"]), final_cost=float(sol.report.final_cost),
                status=str(sol.report.status), iterations=int(sol.report.iterations),
                pose_rotation=floats(sol.pose_rotation), pose_center=floats(sol.pose_center),
                intrinsic_focal=floats(sol.intrinsic_focal), point_xyz=floats(sol.point_xyz))


def main() -> None:
    oracle = load_oracle()
    scene = synthetic.ring_scene()
    nat, arrays = native_host(oracle, scene.problem)
    head = dict(scene="synthetic.ring_scene()", scene_sha256=scene_digest(scene.problem),
                made_by="tests/golden/make_golden.py", source="oracle (C++ FP64 restatement)")

    maps = dict(head, sizes={k: int(v) for k, v in nat.sizes.items()}, gauge_camera=int(nat.gauge_camera),
                scene_scale=float(nat.scene_scale))
    for k in ("camera_offset", "camera_dim", "group_offset", "group_dim", "camera_group", "point_offsets",
              "obs_camera", "point_fixed"):
        maps[k] = ints(arrays[k])

    lin = dict(head, n_observations=min(N_OBS, nat.sizes["n_observations"]),
               squared=linearize(oracle, nat, arrays, A.LOSS_SQUARED, 0.0),
               huber=linearize(oracle, nat, arrays, A.LOSS_HUBER, 1.0))

    D = nat.sizes["dimension"]
    hv = host_view(nat, arrays, A.LOSS_SQUARED, 0.0)
    S, g, fg, dg, sc = oracle_assemble(oracle, hv, D, nat.sizes["n_points"], nat.sizes["n_point_parameters"], 1e-4)
    schur = dict(head, damping=1e-4, dimension=int(D), S=floats(S), gradient=floats(g), full_gradient=floats(fg),
                 diagonal=floats(dg), cost=float(sc.cost), point_model_reduction=float(sc.point_model_reduction),
                 scaled_gradient_inf_norm=float(sc.scaled_gradient_inf_norm),
                 weak_point_count=int(sc.weak_point_count))

    solves = dict(head, max_iterations=SOLVE_ITERATIONS, direct=solve(oracle, scene.problem, A.SOLVER_DIRECT),
                  iterative=solve(oracle, scene.problem, A.SOLVER_ITERATIVE))

    for name, obj in (("ring_index_maps", maps), ("ring_linearize", lin), ("ring_schur", schur),
                      ("ring_solve", solves)):
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(obj, fh, separators=(",", ":"))
        print(name, os.path.getsize(os.path.join(HERE, name + ".json")), "bytes")


def trajectory(sol) -> dict:
    tr = sol.trace
    return dict(cost=floats(tr["cost"]), damping=floats(tr["damping"]), candidate_cost=floats(tr["candidate_cost"]),
                accepted=ints(tr["accepted"]), pcg_iterations=ints(tr["pcg_iterations"]),
                seconds=floats(tr["seconds"]), final_cost=float(sol.report.final_cost),
                status=str(sol.report.status), iterations=int(sol.report.iterations),
                rejected_candidate=int(sol.report.rejected_candidate))


def scale_fixtures() -> None:
    import time

    oracle = load_oracle()
    threads = os.environ.get("OMP_NUM_THREADS", "default")
    out = {}
    # medium: direct step, 10 iterations
    scene = synthetic.medium_scene()
    cfg = SolveConfig(max_iterations=10, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY)
    t0 = time.time()
    sol = solve_with(oracle, "bao_solve", scene.problem, cfg)
    obj = dict(scene="synthetic.medium_scene()", scene_sha256=scene_digest(scene.problem),
               made_by="tests/golden/make_golden.py --scale", source="oracle (C++ FP64 restatement)",
               config=dict(max_iterations=10, linear_solver="direct"), omp_threads=threads,
               wall_seconds=time.time() - t0, **trajectory(sol),
               pose_rotation=floats(sol.pose_rotation), pose_center=floats(sol.pose_center),
               intrinsic_focal=floats(sol.intrinsic_focal), intrinsic_k1=floats(sol.intrinsic_k1),
               intrinsic_k2=floats(sol.intrinsic_k2), point_stride=97, point_xyz=floats(sol.point_xyz[::97]))
    t0 = time.time()
    conv = solve_with(oracle, "bao_solve", scene.problem, SolveConfig(max_iterations=50, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY))
    obj["converged"] = dict(config=dict(max_iterations=50, linear_solver="direct"), wall_seconds=time.time() - t0,
                            **trajectory(conv), pose_center=floats(conv.pose_center),
                            intrinsic_focal=floats(conv.intrinsic_focal), intrinsic_k1=floats(conv.intrinsic_k1),
                            intrinsic_k2=floats(conv.intrinsic_k2), point_xyz=floats(conv.point_xyz[::97]))
    out["medium_direct"] = obj
    print("medium", obj["final_cost"], obj["wall_seconds"], "converged", conv.report.final_cost,
          conv.report.iterations, flush=True)
    # large: PCG step, 5 iterations, then the converged run
    scene = synthetic.large_scene()
    head = dict(scene="synthetic.large_scene()", scene_sha256=scene_digest(scene.problem),
                made_by="tests/golden/make_golden.py --scale", source="oracle (C++ FP64 restatement)",
                omp_threads=threads)
    t0 = time.time()
    five = solve_with(oracle, "bao_solve", scene.problem, SolveConfig(max_iterations=5, linear_solver=A.SOLVER_ITERATIVE, certification=A.CERTIFY_LOOP_ONLY))
    t1 = time.time()
    print("large 5", five.report.final_cost, t1 - t0, flush=True)
    conv = solve_with(oracle, "bao_solve", scene.problem, SolveConfig(max_iterations=60, linear_solver=A.SOLVER_ITERATIVE, certification=A.CERTIFY_LOOP_ONLY))
    t2 = time.time()
    print("large conv", conv.report.final_cost, conv.report.iterations, t2 - t1, flush=True)
    out["large_pcg"] = dict(head, first=dict(config=dict(max_iterations=5, linear_solver="iterative"),
                                             wall_seconds=t1 - t0, **trajectory(five)),
                            converged=dict(config=dict(max_iterations=60, linear_solver="iterative"),
                                           wall_seconds=t2 - t1, **trajectory(conv)))
    # windows: the 4 largest and 4 seeded others
    wins = synthetic.window_scenes()
    sizes = np.array([w.problem.n_observations for w in wins])
    pick = list(np.argsort(-sizes, kind="stable")[:4])
    rng = np.random.default_rng(0xBA0004)
    pick += [int(i) for i in rng.choice(np.setdiff1d(np.arange(len(wins)), pick), 4, replace=False)]
    cfg = SolveConfig(max_iterations=10, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY)
    wl = []
    for i in pick:
        sol = solve_with(oracle, "bao_solve", wins[i].problem, cfg)
        wl.append(dict(window=int(i), scene_sha256=scene_digest(wins[i].problem),
                       n_observations=int(wins[i].problem.n_observations), **trajectory(sol),
                       point_xyz_digest=float(np.sum(sol.point_xyz * sol.point_xyz))))
    out["windows"] = dict(scene="synthetic.window_scenes()", made_by="tests/golden/make_golden.py --scale",
                          source="oracle (C++ FP64 restatement)", config=dict(max_iterations=10, linear_solver="direct"),
                          windows=wl)
    for name, obj in out.items():
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(obj, fh, separators=(",", ":"))
        print(name, os.path.getsize(os.path.join(HERE, name + ".json")), "bytes")


if __name__ == "__main__":
    if "--scale" in sys.argv:
        scale_fixtures()
    else:
        main()
