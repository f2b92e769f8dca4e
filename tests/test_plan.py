"""LinearSolver::Auto's camera-step choice (ba_hip_plan) against the oracle's restatement of
kernel/multi_view/sparse/execution_policy.rs:190-227 (no GPU needed).

Self-calibration has no workload rule (select_self_calibrating takes the direct step whenever its
workspace estimate is at most 4/5 of the planning budget: execution_policy.rs:190-227, 280-286);
the 250,000-observation / D >= 768 rule of execution_policy.rs:21-22,138-139 belongs to
select_calibrated only.  So the 200-camera scene (D = 1,793) takes the direct step under the
automatic 1 GiB budget, and the 2,000-camera scene (D = 17,993) takes PCG.
"""

from __future__ import annotations

import ctypes as C

import pytest

from openmvg_ba import SolveConfig, _abi as A, _lib, synthetic
from openmvg_ba.problem import plan_with


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


@pytest.fixture(scope="module")
def oracle_plan(oracle):
    oracle.bao_plan.argtypes = [C.c_void_p] * 4 + [C.c_size_t]
    return oracle


def both(lib, oracle, problem, **cfg):
    config = SolveConfig(**cfg)
    return plan_with(lib, "ba_hip_plan", problem, config), plan_with(oracle, "bao_plan", problem, config)


def agree(got, want):
    assert got["linear_solver"] == want["linear_solver"]
    assert got["reason"] == want["reason"]
    assert got["budget_bytes"] == want["budget_bytes"]
    assert got["estimated_retained_bytes"] == want["estimated_retained_bytes"]
    assert got["reduced_dimension"] == want["reduced_dimension"]
    if got["reduced_nonzeros"]:  # counted exactly
        assert got["reduced_nonzeros"] == want["reduced_nonzeros"]
        assert got["estimated_direct_bytes"] == want["estimated_direct_bytes"]
    elif got["linear_solver"] == A.SOLVER_DIRECT:  # decided by the dense upper bound
        assert got["estimated_direct_bytes"] >= want["estimated_direct_bytes"]
    else:  # decided by the diagonal-block lower bound
        assert got["estimated_direct_bytes"] <= want["estimated_direct_bytes"]


def test_auto_small_and_medium_scenes_take_the_direct_step(lib, oracle_plan):
    for scene in (synthetic.ring_scene(), synthetic.medium_scene()):
        got, want = both(lib, oracle_plan, scene.problem)
        agree(got, want)
        assert got["linear_solver"] == A.SOLVER_DIRECT and got["reason"] == A.PLAN_DIRECT_FITS_BUDGET
        assert got["budget_bytes"] == 1 << 30


def test_exact_nonzero_count_between_the_bounds(lib, oracle_plan):
    """A budget whose 4/5 lies between the bounds makes the product count the reduced pattern
    (a bitmap of block pairs); the oracle counts the element pattern it builds."""
    problem = synthetic.medium_scene().problem
    _, want = both(lib, oracle_plan, problem)
    exact = want["estimated_direct_bytes"]
    for budget, solver in ((exact * 5 // 4 + 5, A.SOLVER_DIRECT), (exact * 5 // 4 - 5, A.SOLVER_ITERATIVE)):
        got, want = both(lib, oracle_plan, problem, workspace_budget_bytes=budget)
        assert got["reduced_nonzeros"] == want["reduced_nonzeros"] > 0
        agree(got, want)
        assert got["linear_solver"] == solver


def test_tiny_budget_and_explicit_solvers(lib, oracle_plan):
    problem = synthetic.ring_scene().problem
    got, want = both(lib, oracle_plan, problem, workspace_budget_bytes=1000)
    agree(got, want)
    assert got["linear_solver"] == A.SOLVER_ITERATIVE and got["reason"] == A.PLAN_LOWEST_MEMORY
    for solver in (A.SOLVER_DIRECT, A.SOLVER_ITERATIVE):
        got, want = both(lib, oracle_plan, problem, linear_solver=solver, workspace_budget_bytes=1000)
        assert got == want
        assert got["linear_solver"] == solver and got["reason"] == A.PLAN_EXPLICIT_SOLVER


@pytest.mark.slow
def test_auto_large_scene_takes_pcg(lib):
    """The 2,000-camera / 1M-point scene: the retained point cache alone exceeds 4/5 of 1 GiB."""
    problem = synthetic.large_scene().problem
    got = plan_with(lib, "ba_hip_plan", problem, SolveConfig())
    assert got["linear_solver"] == A.SOLVER_ITERATIVE
    assert got["reason"] == A.PLAN_LOWEST_MEMORY
    assert got["reduced_dimension"] == 17993
