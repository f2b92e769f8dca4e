"""Mixed-precision form against the FP64 oracle, through the C ABI, on an MI355X.

The mixed form stores and multiplies the per-observation Jacobians and per-point blocks in FP32,
sums costs, gradients and reduced-system blocks in FP64, factors the dense reduced system in FP32
on the matrix cores and refines each Newton step in FP64; the LM control (gain ratio, damping,
acceptance on the FP64 candidate cost) is FP64.  Its error against the FP64 oracle is stated,
not bit-matched:
  * dense solve: refined relative residual <= 1e-10, solution within 1e-8 of LAPACK;
  * LM solves: per-iteration cost within 1e-6 relative (5e-6 for the PCG solve, whose explicit
    reduced system is stored in FP32), final cost within 1e-9 relative, final parameters within
    1e-6 relative (scale max(|x|, 1)) of the FP64 oracle on the ring scene, and within 1e-5 of
    the FP64 form on the medium scene solved to tight tolerances.
The measured errors are printed with -s.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from openmvg_ba import SolveConfig, synthetic
from openmvg_ba import _abi as A
from openmvg_ba.problem import HipBundleAdjuster, solve_with

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def adjuster(torch_cuda):
    return HipBundleAdjuster(0)


def _spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.geomspace(1.0, 1.0 / cond, n)
    return (Q * ev) @ Q.T


@pytest.mark.parametrize("n,cond", [(5, 1e2), (64, 1e3), (130, 1e4), (257, 1e5), (1000, 1e5)])
def test_mixed_cholesky_refines_to_fp64(torch_cuda, adjuster, n, cond):
    torch = torch_cuda
    M = _spd(n, cond, n)
    b = np.random.default_rng(n + 1).normal(size=n)
    x_ref = np.linalg.solve(M, b)
    a = torch.from_numpy(np.ascontiguousarray(M)).cuda()
    bt = torch.from_numpy(b).cuda()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    code = adjuster.handle.lib.ba_hip_cholesky_solve(adjuster.handle.ptr, n, a.data_ptr(), bt.data_ptr(),
                                                     x.data_ptr(), A.PRECISION_MIXED)
    assert code == A.BA_OK
    xs = x.cpu().numpy()
    err = np.abs(xs - x_ref).max() / np.abs(x_ref).max()
    res = np.linalg.norm(b - M @ xs) / np.linalg.norm(b)
    print(f"n={n} cond={cond:g}: solution error {err:.2e}, residual {res:.2e}")
    assert res <= 1e-10 and err <= 1e-8 * max(1.0, cond * 1e-5)


def test_mixed_cholesky_reports_breakdown(torch_cuda, adjuster):
    """A matrix too ill-conditioned for an FP32 factor is refused, not silently wrong."""
    torch = torch_cuda
    n = 96
    M = _spd(n, 1e13, 7)
    b = np.ones(n)
    a = torch.from_numpy(np.ascontiguousarray(M)).cuda()
    bt = torch.from_numpy(b).cuda()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    code = adjuster.handle.lib.ba_hip_cholesky_solve(adjuster.handle.ptr, n, a.data_ptr(), bt.data_ptr(),
                                                     x.data_ptr(), A.PRECISION_MIXED)
    if code == A.BA_OK:  # refinement converged after all: the answer must then be right
        xs = x.cpu().numpy()
        assert np.linalg.norm(b - M @ xs) <= 1e-10 * np.linalg.norm(b)
    else:
        assert code == A.BA_ERR_NUMERICAL


def _errors(got, want):
    tg, tw = got.trace["cost"], want.trace["cost"]
    n = min(len(tg), len(tw))
    fin = np.isfinite(tw[:n]) & np.isfinite(tg[:n])
    it_err = float(np.max(np.abs(tg[:n][fin] - tw[:n][fin]) / np.abs(tw[:n][fin]))) if fin.any() else 0.0
    final = abs(got.report.final_cost - want.report.final_cost) / abs(want.report.final_cost)
    par = 0.0
    for name in ("pose_center", "point_xyz", "intrinsic_focal", "intrinsic_k1", "intrinsic_k2", "pose_rotation"):
        x, y = getattr(got, name), getattr(want, name)
        par = max(par, float(np.max(np.abs(x - y) / np.maximum(np.abs(y), 1.0))))
    return it_err, final, par


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_mixed_ring_solve_vs_fp64_oracle(adjuster, oracle, solver):
    sc = synthetic.ring_scene()
    got = adjuster.solve(sc.problem, SolveConfig(max_iterations=50, linear_solver=solver,
                                                 precision=A.PRECISION_MIXED))
    want = solve_with(oracle, "bao_solve", sc.problem, SolveConfig(max_iterations=50, linear_solver=solver))
    it_err, final, par = _errors(got, want)
    print(f"ring solver={solver}: per-iteration cost {it_err:.2e}, final cost {final:.2e}, parameters {par:.2e}, "
          f"fallbacks {got.report.precision_fallbacks}, refinement residual {got.report.refinement_residual:.2e}")
    it_bound = 1e-6 if solver == A.SOLVER_DIRECT else 5e-6
    assert it_err <= it_bound and final <= 1e-9 and par <= 1e-6


def test_mixed_medium_direct_vs_fp64(adjuster):
    """200 cameras, 50,000 points, ~300,000 observations, free f, k1, k2 (D = 1,793), direct:
    mixed against the FP64 GPU form (whose per-iteration parity with the oracle is tested on the
    ring and fixture scenes)."""
    sc = synthetic.medium_scene()
    cfg = SolveConfig(max_iterations=40, function_tolerance=1e-12, parameter_tolerance=1e-12,
                      gradient_tolerance=1e-14, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY)
    want = adjuster.solve(sc.problem, cfg)
    got = adjuster.solve(sc.problem, cfg.with_(precision=A.PRECISION_MIXED))
    it_err, final, par = _errors(got, want)
    print(f"medium: per-iteration cost {it_err:.2e}, final cost {final:.2e}, parameters {par:.2e}, "
          f"iterations {got.report.iterations}/{want.report.iterations}, fallbacks {got.report.precision_fallbacks}, "
          f"refinement residual {got.report.refinement_residual:.2e}, loop {got.report.seconds_iterations:.3f} s vs "
          f"{want.report.seconds_iterations:.3f} s")
    assert got.report.precision_fallbacks == 0
    assert final <= 1e-9 and par <= 1e-5
