"""The BAL importer (openmvg_ba/bal.py) on files written with BAL's own conventions
(synthetic.bal_file_text: P = R(w) X + t looking down -z, measurement f r(|p|^2) p with
p = -P.xy / P.z):

* convention: exact measurements of the unperturbed state import as a problem whose cost, by
  the oracle, is zero (1.4e-18 at 3,248 observations): the frame turn, the centre, the image
  flip and the principal-point shift agree with BAL's projection;
* round trip: bal_text(read_bal(text)) parses back to the same cameras, points and measurements
  (to 1e-12: rotations pass through their Rodrigues vectors), plain, .bz2 and .gz;
* malformed files (a missing value, an index out of range) raise ValueError;
* an imported noisy problem converges under the oracle, whose FP64 solve the GPU suite matches
  (tests/test_gpu_bal.py).
"""

from __future__ import annotations

import numpy as np
import pytest

from openmvg_ba import SolveConfig, _abi as A, bal, synthetic
from openmvg_ba.problem import solve_with


def test_exact_measurements_have_zero_cost(oracle):
    prob = bal.read_bal(synthetic.bal_file_text(noise=0.0))
    sol = solve_with(oracle, "bao_solve", prob, SolveConfig(max_iterations=1, certification=A.CERTIFY_LOOP_ONLY))
    assert sol.report.initial_cost <= 1e-15 * prob.n_observations
    # the intrinsics carry BAL's focal and radial terms, the principal point centres the image
    intr = prob.intrinsics[0]
    assert intr.model == A.CAMERA_BROWN_K1K2 and intr.var_focal and intr.var_radial
    assert intr.cx == intr.width / 2 and intr.cy == intr.height / 2
    assert np.all(np.abs(prob.obs_pixel - [intr.cx, intr.cy]) < [intr.cx, intr.cy])


@pytest.mark.parametrize("suffix", ["", ".bz2", ".gz"])
def test_round_trip(tmp_path, suffix):
    text = synthetic.bal_file_text(seed=7)
    cams, pts, oc, op, xy = bal.parse_bal(text)
    prob = bal.read_bal(text)
    path = tmp_path / f"problem.txt{suffix}"
    bal.write_bal(prob, path)
    cams2, pts2, oc2, op2, xy2 = bal.parse_bal(path)
    assert np.array_equal(oc, oc2) and np.array_equal(op, op2)
    assert np.abs(xy2 - xy).max() <= 1e-12 * np.abs(xy).max()
    assert np.abs(cams2 - cams).max() <= 1e-12 * np.abs(cams).max()
    assert np.array_equal(pts2, pts)
    back = bal.read_bal(path)
    assert np.abs(back.pose_rotation - prob.pose_rotation).max() <= 1e-12
    assert np.abs(back.pose_center - prob.pose_center).max() <= 1e-12 * np.abs(prob.pose_center).max()


def test_malformed_files_raise():
    text = synthetic.bal_file_text(seed=3, n_cameras=3, n_points=20)
    with pytest.raises(ValueError, match="expected"):
        bal.parse_bal(text.rsplit("\n", 2)[0])
    lines = text.split("\n")
    c, p, u, v = lines[1].split()
    lines[1] = f"{c} 999 {u} {v}"
    with pytest.raises(ValueError, match="out of range"):
        bal.parse_bal("\n".join(lines))


def test_imported_problem_converges(oracle):
    prob = bal.read_bal(synthetic.bal_file_text(), certification_waived=True)
    sol = solve_with(oracle, "bao_solve", prob, SolveConfig(max_iterations=50))
    r = sol.report
    assert r.status.startswith("converged") and r.final_cost < 0.1 * r.initial_cost
    assert r.observability == A.OBSERVABILITY_FULL_RANK and r.unconstrained_intrinsics == 0
    # per-axis residual RMS below the 0.5 px measurement noise (the fit absorbs its parameters'
    # share), not far below it
    assert 0.3 < r.final_rmse / np.sqrt(2.0) < 0.5
