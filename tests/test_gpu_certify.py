"""The post-loop certification on an MI355X against the oracle (tests/test_certify.py pins the
oracle's certificate): the rank verdict, the intrinsic marginal covariances (within 1e-8 of each
group's largest entry), the groups held by the re-solve, the weak-point re-solves and the
quarantined points, with the trajectories of every solve at the plain form's bars.

Scenes: the ring (full rank, nothing held), the narrow-field scene (every group held and
re-solved from the input), the rank-deficient scene (degenerate, every group held), the
candidate-rule scene (one re-solve without its near and far points; its ill-conditioned steps
are held to the trajectory bar of tests/test_gpu_candidate.py), and the 200-camera scene
(D = 1,793: 200 groups' 3x3 marginals from the device factor)."""

from __future__ import annotations

import time

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


def _both(adjuster, oracle, prob, cfg):
    want = solve_with(oracle, "bao_solve", prob, cfg, trace_capacity=400)
    got = adjuster.solve(prob, cfg, trace_capacity=400)
    return got, want


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
@pytest.mark.parametrize("scene", ["ring", "narrow", "rank_deficient"])
def test_certificate_matches_oracle(adjuster, oracle, scene, solver):
    prob = {"ring": lambda: synthetic.ring_scene().problem,
            "narrow": lambda: synthetic.narrow_field_scene().problem,
            "rank_deficient": synthetic.rank_deficient_scene}[scene]()
    got, want = _both(adjuster, oracle, prob, SolveConfig(max_iterations=50, linear_solver=solver))
    r = got.report
    print(f"{scene} solver {solver}: status {r.status}, observability {r.observability}, rank {r.pose_rank}/"
          f"{r.pose_dimension}, held {r.unconstrained_intrinsics}, iterations {r.iterations}")
    _compare_solutions(got, want)


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_weak_point_resolve_matches_oracle(adjuster, oracle, solver):
    prob, q_near, q_far = synthetic.candidate_rule_scene()
    got, want = _both(adjuster, oracle, prob, SolveConfig(max_iterations=30, linear_solver=solver))
    assert want.report.weak_point_resolves == 1
    assert np.nonzero(got.point_quarantined)[0].tolist() == [q_near, q_far]
    _compare_solutions(got, want, tol_cost=2e-9, point_norm=True)


def test_medium_scene_marginals_match_oracle(adjuster, oracle):
    prob = synthetic.medium_scene().problem
    cfg = SolveConfig(max_iterations=10, linear_solver=A.SOLVER_DIRECT)
    want = solve_with(oracle, "bao_solve", prob, cfg)
    t0 = time.perf_counter()
    got = adjuster.solve(prob, cfg)
    call = time.perf_counter() - t0
    loop_only = adjuster.solve(prob, cfg.with_(certification=A.CERTIFY_LOOP_ONLY))
    r = got.report
    print(f"medium: observability {r.observability}, rank {r.pose_rank}/{r.pose_dimension}, held "
          f"{r.unconstrained_intrinsics}; call {call:.3f} s (loop {r.seconds_iterations:.3f} s, after the loop "
          f"{call - r.seconds_setup - r.seconds_iterations:.3f} s)")
    assert r.observability == A.OBSERVABILITY_FULL_RANK
    assert np.isfinite(got.intrinsic_covariance).all()
    assert loop_only.report.final_cost == r.final_cost  # the certificate changes nothing here
    _compare_solutions(got, want)
