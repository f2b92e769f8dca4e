"""Direct and PCG solves of scenes whose cameras share one intrinsic group, on an MI355X against
the oracle.

With one camera per intrinsic group (and no camera observing a point twice) the direct step's
dense S is scattered from the explicit system's camera rows; a shared group couples every camera
through its intrinsic block, so these scenes take the block-pair path (U and W V^-1 W^T summed per
block pair over its terms) in the LM loop itself.  Bars: the same accept/reject sequence,
per-iteration cost within 1e-10 relative, parameters within 1e-8, the same certificate.
"""

from __future__ import annotations

from dataclasses import replace

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


def shared(problem):
    """Every observation through the first camera's intrinsics."""
    return replace(problem, intrinsics=[problem.intrinsics[0]],
                   obs_intrinsic=np.zeros_like(problem.obs_intrinsic))


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
@pytest.mark.parametrize("loss", [A.LOSS_SQUARED, A.LOSS_HUBER])
def test_shared_group_matches_oracle(adjuster, oracle, solver, loss):
    prob = replace(shared(synthetic.ring_scene().problem), robust_loss=loss,
                   loss_scale=1.0 if loss != A.LOSS_SQUARED else 0.0)
    cfg = SolveConfig(max_iterations=50, linear_solver=solver)
    want = solve_with(oracle, "bao_solve", prob, cfg, trace_capacity=200)
    got = adjuster.solve(prob, cfg, trace_capacity=200)
    r = got.report
    print(f"shared group, loss {loss}, solver {solver}: D {r.reduced_dimension}, {r.iterations} iterations, "
          f"cost {r.initial_cost:.6g} -> {r.final_cost:.10g}, observability {r.observability}")
    assert r.reduced_dimension == want.report.reduced_dimension
    _compare_solutions(got, want)
