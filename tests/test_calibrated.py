"""The calibrated (fixed-intrinsic) pixel path on the CPU: the oracle's restatement pinned, the
product's host lowering and Auto plan against it, and the declines (no GPU needed).

The path is crates/openmvg-ba/src/native/mod.rs:336-455 (solve_calibrated) with
native/calibrated.rs:180-400 (PreparedCalibratedProblem), native/gauge.rs:123-229
(calibrated_gauge_plan) and kernel/multi_view/sparse/calibrated/mod.rs:169-589.  It shares the
self-calibrating loop's arithmetic except for three rules, each pinned here by what the
mathematics fixes rather than by restating the formula:

* the camera step limit (optimizer.rs:677-702): a camera rolled 0.6 rad about its optical axis
  makes the first direct step's rotation correction exceed 0.5, and the accepted rotation change
  of that camera is then exactly 0.5 rad (the retraction R' = exp(w) R turns by |w|);
* the direct step's model reduction from the damped quadratic form
  (calibrated/camera_step.rs:40-62): on a noise-free scene the LM gain ratio of every step is 1
  to within the quadratic model's error (a dropped 1/2 or sign gives 0.5, 2 or a negative ratio);
* the gauge plan: a free-scale scene keeps the anchor fixed and the farthest camera on the S^2
  baseline chart (D = 6 (n - 2) + 5); survey controls without a fixed pose leave every camera free
  (D = 6 n).
The noise-free scene is also solved to its exact minimiser (final cost ~ 0) with the anchor and the
gauge camera's baseline kept.
"""

from __future__ import annotations

import numpy as np
import pytest

from openmvg_ba import SolveConfig, Unsupported, _abi as A, _lib, synthetic
from openmvg_ba.problem import BundleProblem, prepare_with, plan_with, solve_with


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _rotation_angles(Ra, Rb):
    cos = (np.einsum("nii->n", Ra @ np.transpose(Rb, (0, 2, 1))) - 1.0) / 2.0
    return np.arccos(np.clip(cos, -1.0, 1.0))


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_noise_free_scene_reaches_its_minimiser(oracle, solver):
    sc = synthetic.calibrated_ring_scene(noise=0.0)
    p = sc.problem
    cfg = SolveConfig(max_iterations=50, linear_solver=solver, function_tolerance=1e-14)
    r = solve_with(oracle, "bao_solve", p, cfg)
    assert r.report.initial_cost > 1e5
    assert r.report.final_cost < 1e-12
    assert r.report.reduced_dimension == 6 * (p.n_poses - 2) + 5
    # the anchor (pose 0, fixed by the plan) keeps its input pose; the farthest camera's baseline
    # to it is the gauge
    assert np.array_equal(r.pose_rotation[0], p.pose_rotation[0])
    assert np.array_equal(r.pose_center[0], p.pose_center[0])
    d_in = np.linalg.norm(p.pose_center - p.pose_center[0], axis=1)
    d_out = np.linalg.norm(r.pose_center - r.pose_center[0], axis=1)
    far = int(np.argmax(d_in))
    assert abs(d_out[far] - d_in[far]) <= 1e-12 * d_in[far]
    # the intrinsics are returned as given
    assert np.array_equal(r.intrinsic_focal, np.full(p.n_poses, 1000.0))


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_gain_ratio_is_one_on_a_noise_free_scene(oracle, solver):
    sc = synthetic.calibrated_ring_scene(noise=0.0)
    r = solve_with(oracle, "bao_solve", sc.problem,
                   SolveConfig(max_iterations=50, linear_solver=solver, function_tolerance=1e-14))
    t = r.trace
    acc = t["accepted"].astype(bool) & (t["cost"] > 1e-6)
    assert acc.sum() >= 3
    assert np.all(np.abs(t["gain_ratio"][acc] - 1.0) < 1e-3), t["gain_ratio"]


def test_step_limit_holds_the_largest_block_to_one_half(oracle):
    sc = synthetic.step_limit_scene(roll=0.6, camera=3)
    p = sc.problem
    cfg = SolveConfig(max_iterations=1, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY)
    r = solve_with(oracle, "bao_solve", p, cfg)
    assert r.trace["accepted"][0]
    ang = _rotation_angles(r.pose_rotation, p.pose_rotation)
    assert abs(ang[3] - 0.5) < 1e-12, ang
    assert np.all(np.delete(ang, 3) < 0.1)
    # without the roll, no block reaches the limit
    r0 = solve_with(oracle, "bao_solve", synthetic.calibrated_ring_scene().problem, cfg)
    assert _rotation_angles(r0.pose_rotation, synthetic.calibrated_ring_scene().problem.pose_rotation).max() < 0.1


def test_gauge_plans(lib, oracle):
    ring = synthetic.calibrated_ring_scene().problem
    a = prepare_with(oracle, "bao_prepare", ring)
    assert a.sizes["dimension"] == 6 * (ring.n_poses - 2) + 5 and a.gauge_camera >= 1
    ctl = synthetic.controls_scene().problem
    b = prepare_with(oracle, "bao_prepare", ctl)
    assert b.sizes["dimension"] == 6 * ctl.n_poses and b.gauge_camera == -1
    r = solve_with(oracle, "bao_solve", ctl, SolveConfig(max_iterations=30))
    assert r.report.final_cost < 2.0 * r.report.observation_count  # at the noise level (sigma 0.5)
    # the controls stay where they were given
    assert np.array_equal(r.point_xyz[:40], ctl.point_xyz[:40])


@pytest.mark.parametrize("name", ["ring", "brown", "controls", "limit"])
def test_host_lowering_is_bit_identical(lib, oracle, name):
    scene = {"ring": synthetic.calibrated_ring_scene, "brown": synthetic.calibrated_brown_scene,
             "controls": synthetic.controls_scene, "limit": synthetic.step_limit_scene}[name]()
    a = prepare_with(lib, "ba_hip_prepare", scene.problem)
    b = prepare_with(oracle, "bao_prepare", scene.problem)
    assert a.sizes == b.sizes
    assert a.gauge_camera == b.gauge_camera
    assert a.scene_scale == b.scene_scale
    for k in a.arrays:
        assert np.array_equal(a.arrays[k], b.arrays[k]), k
    assert not b.arrays["group_dim"].any()  # every intrinsic fixed: no intrinsic blocks


def test_auto_plan_follows_the_calibrated_rule(lib, oracle):
    """select_calibrated (execution_policy.rs:123-183): at least 250,000 observations and D >= 768
    prefer the retained PCG step; a small scene takes the direct step."""
    small = synthetic.calibrated_ring_scene().problem
    big = synthetic.fixed_intrinsics(synthetic.medium_scene()).problem
    for problem, solver, reason in ((small, A.SOLVER_DIRECT, A.PLAN_DIRECT_FITS_BUDGET),
                                    (big, A.SOLVER_ITERATIVE, A.PLAN_WORKLOAD_PREFERS_RETAINED)):
        got = plan_with(lib, "ba_hip_plan", problem, SolveConfig())
        want = plan_with(oracle, "bao_plan", problem, SolveConfig())
        assert got["linear_solver"] == want["linear_solver"] == solver
        assert got["reason"] == want["reason"] == reason
        assert got["estimated_retained_bytes"] == want["estimated_retained_bytes"]
        assert got["reduced_dimension"] == want["reduced_dimension"]
    # the self-calibrating medium scene has no workload rule: it takes the direct step
    got = plan_with(lib, "ba_hip_plan", synthetic.medium_scene().problem, SolveConfig())
    assert got["linear_solver"] == A.SOLVER_DIRECT
    # a budget below the retained estimate: the lowest-memory choice
    got = plan_with(lib, "ba_hip_plan", small, SolveConfig(workspace_budget_bytes=1000))
    want = plan_with(oracle, "bao_plan", small, SolveConfig(workspace_budget_bytes=1000))
    assert got["linear_solver"] == want["linear_solver"] == A.SOLVER_ITERATIVE
    assert got["reason"] == want["reason"] == A.PLAN_LOWEST_MEMORY


def _two_view():
    sc = synthetic.calibrated_ring_scene()
    p = sc.problem
    keep = np.isin(p.obs_pose, [0, 5])
    pts, inv = np.unique(p.obs_point[keep], return_inverse=True)
    counts = np.bincount(inv)
    obs_keep = np.flatnonzero(keep)[counts[inv] == 2]
    pts2, inv2 = np.unique(p.obs_point[obs_keep], return_inverse=True)
    pose = np.where(p.obs_pose[obs_keep] == 0, 0, 1)
    var = np.array([0, 1], dtype=np.uint8)
    return BundleProblem(p.pose_rotation[[0, 5]], p.pose_center[[0, 5]], [p.intrinsics[0], p.intrinsics[5]],
                         p.point_xyz[pts2], pose, pose, inv2, p.obs_pixel[obs_keep], pose_var_rotation=var,
                         pose_var_translation=var.copy())


def _declines():
    ring = synthetic.calibrated_ring_scene().problem
    from dataclasses import replace
    aniso = replace(ring, intrinsics=[replace(i, focal_y=1001.0) for i in ring.intrinsics])
    tang = replace(ring, intrinsics=[replace(i, model=A.CAMERA_BROWN_K1K2, p1=1e-4) for i in ring.intrinsics])
    k3 = replace(ring, intrinsics=[replace(i, model=A.CAMERA_BROWN_K1K2K3, k3=1e-3) for i in ring.intrinsics])
    fisheye = replace(ring, intrinsics=[replace(i, model=A.CAMERA_FISHEYE) for i in ring.intrinsics])
    vt = np.ones(ring.n_poses, dtype=np.uint8)
    vr = vt.copy()
    vr[4] = 0  # translation only
    masks = replace(ring, pose_var_rotation=vr, pose_var_translation=vt)
    angular = replace(ring, residual_space=A.RESIDUAL_ANGULAR)
    return {"two_view": _two_view(), "anisotropic": aniso, "tangential": tang, "k3": k3, "fisheye": fisheye,
            "translation_only": masks, "angular": angular}


@pytest.mark.parametrize("case", ["two_view", "anisotropic", "tangential", "k3", "fisheye", "translation_only",
                                  "angular"])
def test_outside_the_slice_is_declined(lib, oracle, case):
    """Problems the reference solves by paths this library does not restate return
    SolveError::Unsupported, the sanctioned decline (openmvg-sfm/src/ba/native.rs:79-85)."""
    problem = _declines()[case]
    for l, fn in ((lib, "ba_hip_plan"), (oracle, "bao_plan")):
        with pytest.raises(Unsupported):
            plan_with(l, fn, problem, SolveConfig())
