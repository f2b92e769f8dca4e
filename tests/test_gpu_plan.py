"""The boundary's budget and report contract on an MI355X, through the C ABI.

* An explicit workspace budget that the solve's device workspace exceeds is refused with
  ResourceExhausted before the first iteration (check_workspace_budget, native/mod.rs:557-570);
  a sufficient one is honoured and the report states the workspace.
* AUTO takes the camera step of ba_hip_plan (execution_policy.rs:190-227).
* The report's inlier count and inlier RMSE (solver.rs:1353-1371) equal the oracle's, which
  evaluates them in the caller's frame from the returned state.
"""

from __future__ import annotations

import numpy as np
import pytest

from openmvg_ba import SolveConfig, _abi as A, synthetic
from openmvg_ba.problem import HipBundleAdjuster, ResourceExhausted, plan_with, solve_with
from oracle_lib import load_oracle

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


def test_explicit_budget_refuses_a_larger_workspace(adjuster):
    problem = synthetic.ring_scene().problem
    for solver in (A.SOLVER_DIRECT, A.SOLVER_ITERATIVE):
        ok = adjuster.solve(problem, SolveConfig(max_iterations=3, linear_solver=solver))
        used = int(ok.report.workspace_bytes)
        assert used > 0
        with pytest.raises(ResourceExhausted, match="exceeds the workspace budget"):
            adjuster.solve(problem, SolveConfig(max_iterations=3, linear_solver=solver,
                                                workspace_budget_bytes=used - 1))
        fit = adjuster.solve(problem, SolveConfig(max_iterations=3, linear_solver=solver,
                                                  workspace_budget_bytes=used))
        assert fit.report.workspace_bytes == used
        np.testing.assert_array_equal(fit.point_xyz, ok.point_xyz)


def test_auto_takes_the_planned_step(adjuster):
    problem = synthetic.ring_scene().problem
    for budget in (0, 1000):
        cfg = SolveConfig(max_iterations=3, workspace_budget_bytes=budget)
        plan = adjuster.plan(problem, cfg)
        if budget:  # the tiny budget plans PCG, then refuses the workspace
            assert plan["linear_solver"] == A.SOLVER_ITERATIVE and plan["reason"] == A.PLAN_LOWEST_MEMORY
            with pytest.raises(ResourceExhausted):
                adjuster.solve(problem, cfg)
            continue
        sol = adjuster.solve(problem, cfg)
        assert sol.report.linear_solver_used == plan["linear_solver"] == A.SOLVER_DIRECT
        assert sol.report.plan_reason == plan["reason"] == A.PLAN_DIRECT_FITS_BUDGET


@pytest.mark.parametrize("loss", [A.LOSS_SQUARED, A.LOSS_HUBER, A.LOSS_CAUCHY])
def test_inlier_statistics_match_oracle(adjuster, oracle, loss):
    scene = synthetic.ring_scene()
    problem = scene.problem
    if loss != A.LOSS_SQUARED:
        problem.robust_loss = loss
        problem.loss_scale = 0.7
    cfg = SolveConfig(max_iterations=20, linear_solver=A.SOLVER_DIRECT)
    got = adjuster.solve(problem, cfg)
    want = solve_with(oracle, "bao_solve", problem, cfg)
    assert got.report.observation_count == want.report.observation_count
    assert got.report.inlier_count == want.report.inlier_count
    if loss == A.LOSS_SQUARED:
        assert got.report.inlier_count == got.report.observation_count
        # equals final_rmse (per-axis RMSE x sqrt 2) under the squared loss
        assert abs(got.report.final_inlier_rmse - got.report.final_rmse) <= 1e-12 * got.report.final_rmse
    else:
        assert 0 < got.report.inlier_count < got.report.observation_count
        assert got.report.final_inlier_rmse <= 0.7
    assert abs(got.report.final_inlier_rmse - want.report.final_inlier_rmse) <= 1e-9 * want.report.final_inlier_rmse
