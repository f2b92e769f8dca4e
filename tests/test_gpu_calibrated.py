"""The calibrated (fixed-intrinsic) pixel path on an MI355X against the oracle
(tests/test_calibrated.py pins the oracle's step limit, model reduction and gauge plans).

Bars of the plain form: the same accept/reject sequence, per-iteration cost within 1e-10
relative, final parameters within 1e-8 relative, the same certificate.  Scenes: the pinhole ring
(squared and Huber loss), the Brown (k1, k2) grid, the rolled-camera scene whose first direct steps
are held by the step limit, the survey-control scene with every camera free, and the 200-camera
scene with every intrinsic fixed, which Auto sends to PCG (at least 250,000 observations and
D >= 768), in both forms of the Schur product.  The mixed form's error against the FP64 oracle is
stated for the ring.
"""

from __future__ import annotations

import numpy as np
import pytest

from openmvg_ba import SolveConfig, _abi as A, synthetic
from openmvg_ba.problem import HipBundleAdjuster, solve_with
from oracle_lib import load_oracle
from test_gpu_parity import _compare_solutions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def adjuster():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return HipBundleAdjuster(0)


@pytest.fixture(scope="module")
def oracle():
    return load_oracle()


SCENES = {
    "ring": lambda: synthetic.calibrated_ring_scene().problem,
    "ring_huber": lambda: synthetic.calibrated_ring_scene(loss=A.LOSS_HUBER, loss_scale=1.0).problem,
    "brown": lambda: synthetic.calibrated_brown_scene().problem,
    "step_limit": lambda: synthetic.step_limit_scene().problem,
    "controls": lambda: synthetic.controls_scene().problem,
}


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
@pytest.mark.parametrize("scene", list(SCENES))
def test_calibrated_solve_matches_oracle(adjuster, oracle, scene, solver):
    prob = SCENES[scene]()
    cfg = SolveConfig(max_iterations=50, linear_solver=solver)
    want = solve_with(oracle, "bao_solve", prob, cfg, trace_capacity=200)
    got = adjuster.solve(prob, cfg, trace_capacity=200)
    r = got.report
    print(f"calibrated {scene} solver {solver}: D {r.reduced_dimension}, {r.iterations} iterations, status "
          f"{r.status}, cost {r.initial_cost:.6g} -> {r.final_cost:.10g} (oracle {want.report.final_cost:.10g}), "
          f"observability {r.observability}")
    assert r.reduced_dimension == want.report.reduced_dimension
    _compare_solutions(got, want)


@pytest.mark.parametrize("product", [A.SCHUR_IMPLICIT, A.SCHUR_EXPLICIT])
def test_calibrated_pcg_products_match_oracle(adjuster, oracle, product):
    prob = SCENES["step_limit"]()
    cfg = SolveConfig(max_iterations=50, linear_solver=A.SOLVER_ITERATIVE)
    want = solve_with(oracle, "bao_solve", prob, cfg, trace_capacity=200)
    h = adjuster.handle
    assert h.lib.ba_hip_set_schur_product(h.ptr, product) == A.BA_OK
    try:
        got = adjuster.solve(prob, cfg, trace_capacity=200)
        assert h.lib.ba_hip_last_schur_product(h.ptr) == product
    finally:
        h.lib.ba_hip_set_schur_product(h.ptr, A.SCHUR_AUTO)
    _compare_solutions(got, want)


def test_calibrated_medium_scene_takes_pcg_and_matches_oracle(adjuster, oracle):
    """The 200-camera scene with its intrinsics fixed (D = 1,193, ~298k observations): Auto picks
    the retained PCG step by the workload rule; 8 LM iterations against the oracle."""
    prob = synthetic.fixed_intrinsics(synthetic.medium_scene()).problem
    cfg = SolveConfig(max_iterations=8, certification=A.CERTIFY_LOOP_ONLY)
    want = solve_with(oracle, "bao_solve", prob, cfg, trace_capacity=50)
    got = adjuster.solve(prob, cfg, trace_capacity=50)
    r = got.report
    assert r.linear_solver_used == A.SOLVER_ITERATIVE and r.plan_reason == A.PLAN_WORKLOAD_PREFERS_RETAINED
    print(f"calibrated medium: D {r.reduced_dimension}, {r.iterations} iterations, {r.pcg_iterations} PCG "
          f"iterations, cost {r.initial_cost:.8g} -> {r.final_cost:.10g}, loop {r.seconds_iterations:.4f} s")
    # eight inexact PCG solves carry rounding into the weakly observed camera depths: parameters
    # within 1e-7 (the medium scene's bar for the two product forms)
    _compare_solutions(got, want, tol_param=1e-7)


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_calibrated_mixed_error_against_oracle(adjuster, oracle, solver):
    prob = SCENES["ring"]()
    want = solve_with(oracle, "bao_solve", prob, SolveConfig(max_iterations=50, linear_solver=solver))
    got = adjuster.solve(prob, SolveConfig(max_iterations=50, linear_solver=solver, precision=A.PRECISION_MIXED))
    rel = abs(got.report.final_cost - want.report.final_cost) / want.report.final_cost
    perr = np.max(np.abs(got.point_xyz - want.point_xyz) / np.maximum(np.abs(want.point_xyz), 1.0))
    print(f"calibrated mixed solver {solver}: final cost rel error {rel:.2e}, point error {perr:.2e}")
    assert rel <= 1e-9
    assert perr <= 1e-6
