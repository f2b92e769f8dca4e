"""The candidate-rule scene (synthetic.candidate_rule_scene) drives the oracle through the held and
reject-step branches of crates/openmvg-ba/src/kernel/point_geometry.rs:153-168 and
sparse/self_calibration/candidate.rs:25-233, for the direct and the PCG step, squared and Huber
loss, in the LM loop alone (the post-loop certification then holds both added points out: see
tests/test_certify.py).  The GPU suite (tests/test_gpu_candidate.py) compares the HIP path on the
same scene."""

import numpy as np
import pytest

from openmvg_ba import SolveConfig, _abi as A, synthetic
from openmvg_ba.problem import solve_with


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
@pytest.mark.parametrize("loss", [A.LOSS_SQUARED, A.LOSS_HUBER])
def test_scene_rejects_steps_and_holds_a_point(oracle, solver, loss):
    prob, q_near, _ = synthetic.candidate_rule_scene()
    prob.robust_loss, prob.loss_scale = loss, (1.0 if loss == A.LOSS_HUBER else 0.0)
    cfg = SolveConfig(max_iterations=30, linear_solver=solver, certification=A.CERTIFY_LOOP_ONLY)
    sol = solve_with(oracle, "bao_solve", prob, cfg)
    rep = sol.report
    assert rep.rejected_candidate >= 2  # the candidate rule rejected steps
    assert rep.accepted_steps >= 8
    # the near point is held at its previous position by every accepted step
    assert np.abs(sol.point_xyz[q_near] - prob.point_xyz[q_near]).max() <= 1e-12 * 10.0
    assert sol.point_quarantined[q_near] == 0
