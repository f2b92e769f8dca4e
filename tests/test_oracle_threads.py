"""The oracle's OpenMP passes run over point shards fixed by the problem (the rule of
crates/openmvg-ba/src/kernel/multi_view/sparse/retained/parallel.rs:94-122), so the thread count
changes the time, never the bits: a solve of the 200-camera / 50,000-point scene (297,949
observations: 32 shards) with 1 thread and with 8 threads gives identical traces and parameters,
for the direct and the PCG step."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from openmvg_ba import SolveConfig, _abi as A, synthetic
from openmvg_ba.problem import solve_with


def _threads(n: int | None = None) -> int:
    """Set (when n is given) and return the OpenMP thread count the oracle uses."""
    gomp = ctypes.CDLL("libgomp.so.1")
    if n is not None:
        gomp.omp_set_num_threads(n)
    return int(gomp.omp_get_max_threads())


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_thread_count_does_not_change_the_result(oracle, solver):
    prob = synthetic.medium_scene().problem
    cfg = SolveConfig(max_iterations=3, linear_solver=solver)
    runs, before = [], _threads()
    try:
        for n in (1, 8):
            _threads(n)
            runs.append(solve_with(oracle, "bao_solve", prob, cfg))
    finally:
        _threads(before)
    a, b = runs
    for k in ("cost", "candidate_cost", "damping", "predicted_reduction", "gain_ratio"):
        assert np.array_equal(a.trace[k], b.trace[k], equal_nan=True), k
    for k in ("accepted", "pcg_iterations"):
        assert np.array_equal(a.trace[k], b.trace[k]), k
    assert a.report.final_cost == b.report.final_cost
    for name in ("pose_rotation", "pose_center", "intrinsic_focal", "intrinsic_k1", "intrinsic_k2", "point_xyz"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
