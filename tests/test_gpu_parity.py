"""HIP path against the CPU oracle, through the C ABI, on an MI355X.

Tolerances: index maps and structure are bit-exact (tests/test_index_maps.py); FP64 quantities
agree to rounding, which differs only by summation order and fused multiply-adds:
per-observation values 1e-12 relative, Schur system 1e-10 relative to its largest entry,
per-iteration LM cost 1e-10 relative, final parameters 1e-8 relative (the plain form's bar).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from helpers import device_view, host_view, native_host, oracle_assemble, rel_err, to_device
from openmvg_ba import SolveConfig, synthetic
from openmvg_ba import _abi as A
from openmvg_ba.problem import HipBundleAdjuster, solve_with
from oracle_lib import load_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def oracle():
    return load_oracle()


@pytest.fixture(scope="module")
def adjuster(torch_cuda):
    return HipBundleAdjuster(0)


@pytest.fixture(scope="module")
def ring(oracle):
    sc = synthetic.ring_scene()
    nat, arrays = native_host(oracle, sc.problem)
    return sc, nat, arrays


def test_name_matches_contract(adjuster):
    assert adjuster.name() == "openmvg-ba-hip"


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 257, 1000])
def test_cholesky_solve_matches_lapack_and_oracle(torch_cuda, adjuster, oracle, n):
    torch = torch_cuda
    rng = np.random.default_rng(n)
    M = rng.normal(size=(n, n))
    A_ = M @ M.T + n * np.eye(n)
    b = rng.normal(size=n)
    x_ref = np.linalg.solve(A_, b)
    dA = torch.from_numpy(A_).cuda()
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros(n, dtype=torch.float64, device="cuda")
    lib = adjuster.handle.lib
    code = lib.ba_hip_cholesky_solve(adjuster.handle.ptr, n, dA.data_ptr(), db.data_ptr(), dx.data_ptr(),
                                     A.PRECISION_FP64)
    assert code == A.BA_OK
    x = dx.cpu().numpy()
    assert rel_err(x, x_ref) < 1e-11
    xo = np.zeros(n)
    assert oracle.bao_cholesky_solve(n, A_.ctypes.data, b.ctypes.data, xo.ctypes.data) == 1
    assert rel_err(x, xo) < 1e-11


@pytest.mark.parametrize("n", [5, 52, 64, 65, 130])
@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
def test_cholesky_badly_scaled(torch_cuda, adjuster, oracle, n, precision):
    """Diagonal scales from 1e7 to 1e15 (a camera system whose translation columns carry a large
    scene scale): the pivot floor 64 eps max|diag| then exceeds 1, which the unit pivots of a
    partial tile's identity padding must not be held to."""
    torch = torch_cuda
    rng = np.random.default_rng(100 + n)
    M = rng.normal(size=(n, n))
    d = 10.0 ** rng.uniform(7.0, 15.0, n)
    d[0] = 1e15
    A_ = (M @ M.T + n * np.eye(n)) * np.sqrt(np.outer(d, d))
    b = rng.normal(size=n) * np.sqrt(d)
    x_ref = np.linalg.solve(A_, b)
    dA = torch.from_numpy(A_).cuda()
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros(n, dtype=torch.float64, device="cuda")
    code = adjuster.handle.lib.ba_hip_cholesky_solve(adjuster.handle.ptr, n, dA.data_ptr(), db.data_ptr(),
                                                     dx.data_ptr(), precision)
    assert code == A.BA_OK
    xo = np.zeros(n)
    assert oracle.bao_cholesky_solve(n, A_.ctypes.data, b.ctypes.data, xo.ctypes.data) == 1
    x = dx.cpu().numpy()
    # componentwise in the scaled unknowns sqrt(d) x (the system is well conditioned after scaling)
    assert np.abs((x - x_ref) * np.sqrt(d)).max() <= 1e-9 * np.abs(x_ref * np.sqrt(d)).max()


def test_cholesky_rejects_indefinite(torch_cuda, adjuster):
    torch = torch_cuda
    n = 70
    A_ = np.eye(n)
    A_[40, 40] = -1.0
    dA = torch.from_numpy(A_).cuda()
    db = torch.ones(n, dtype=torch.float64, device="cuda")
    dx = torch.zeros(n, dtype=torch.float64, device="cuda")
    code = adjuster.handle.lib.ba_hip_cholesky_solve(adjuster.handle.ptr, n, dA.data_ptr(), db.data_ptr(),
                                                     dx.data_ptr(), A.PRECISION_FP64)
    assert code == A.BA_ERR_NUMERICAL


@pytest.mark.parametrize("loss,scale", [(A.LOSS_SQUARED, 0.0), (A.LOSS_HUBER, 1.0), (A.LOSS_CAUCHY, 2.0)])
def test_linearize_matches_oracle(torch_cuda, adjuster, oracle, ring, loss, scale):
    torch = torch_cuda
    sc, nat, arrays = ring
    no = nat.sizes["n_observations"]
    hv = host_view(nat, arrays, loss, scale)
    r0, j0, c0, k0, v0 = np.zeros(2 * no), np.zeros(6 * no), np.zeros(18 * no), np.zeros(no), np.zeros(no, np.uint8)
    oracle.bao_linearize(C.byref(hv), *(x.ctypes.data_as(C.c_void_p) for x in (r0, j0, c0, k0, v0)))
    dev = to_device(arrays)
    dv = device_view(nat, dev, arrays, loss, scale)
    out = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in (2 * no, 6 * no, 18 * no, no)]
    valid = torch.zeros(no, dtype=torch.uint8, device="cuda")
    code = adjuster.handle.lib.ba_hip_linearize(adjuster.handle.ptr, C.byref(dv), *(t.data_ptr() for t in out),
                                               valid.data_ptr())
    assert code == A.BA_OK
    assert np.array_equal(valid.cpu().numpy(), v0)
    for got, want in zip(out, (r0, j0, c0, k0)):
        assert rel_err(got.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("lam", [1e-4, 0.0])
def test_schur_matches_oracle(torch_cuda, adjuster, oracle, ring, lam):
    torch = torch_cuda
    sc, nat, arrays = ring
    D = nat.sizes["dimension"]
    hv = host_view(nat, arrays, A.LOSS_SQUARED, 0.0)
    S0, g0, fg0, dg0, s0 = oracle_assemble(oracle, hv, D, nat.sizes["n_points"], nat.sizes["n_point_parameters"], lam)
    dev = to_device(arrays)
    dv = device_view(nat, dev, arrays, A.LOSS_SQUARED, 0.0)
    S = torch.zeros(D * D, dtype=torch.float64, device="cuda")
    g, fg, dg = (torch.zeros(D, dtype=torch.float64, device="cuda") for _ in range(3))
    s1 = A.SchurScalars()
    code = adjuster.handle.lib.ba_hip_schur(adjuster.handle.ptr, C.byref(dv), lam, S.data_ptr(), g.data_ptr(),
                                           fg.data_ptr(), dg.data_ptr(), C.byref(s1))
    assert code == A.BA_OK
    assert s1.valid == s0.valid == 1
    assert rel_err(S.cpu().numpy().reshape(D, D), S0) < 1e-10
    assert rel_err(g.cpu().numpy(), g0) < 1e-10
    assert rel_err(fg.cpu().numpy(), fg0) < 1e-10
    assert rel_err(dg.cpu().numpy(), dg0) < 1e-12
    assert abs(s1.cost - s0.cost) <= 1e-12 * abs(s0.cost)
    assert abs(s1.point_model_reduction - s0.point_model_reduction) <= 1e-10 * abs(s0.point_model_reduction)
    assert abs(s1.scaled_gradient_inf_norm - s0.scaled_gradient_inf_norm) <= 1e-9 * s0.scaled_gradient_inf_norm
    assert s1.weak_point_count == s0.weak_point_count


def test_pcg_step_matches_direct_solve(torch_cuda, adjuster, oracle, ring):
    torch = torch_cuda
    sc, nat, arrays = ring
    D = nat.sizes["dimension"]
    lam = 0.1
    hv = host_view(nat, arrays, A.LOSS_SQUARED, 0.0)
    S0, g0, *_ = oracle_assemble(oracle, hv, D, nat.sizes["n_points"], nat.sizes["n_point_parameters"], lam)
    x_ref = np.linalg.solve(S0, -g0)
    dev = to_device(arrays)
    dv = device_view(nat, dev, arrays, A.LOSS_SQUARED, 0.0)
    step = torch.zeros(D, dtype=torch.float64, device="cuda")
    it, rel, acc = C.c_uint32(), C.c_double(), C.c_int32()
    code = adjuster.handle.lib.ba_hip_pcg(adjuster.handle.ptr, C.byref(dv), lam, 1e-12, 500, step.data_ptr(),
                                         C.byref(it), C.byref(rel), C.byref(acc))
    assert code == A.BA_OK and acc.value == 1 and rel.value <= 1e-12
    assert rel_err(step.cpu().numpy(), x_ref) < 1e-8
    # same PCG recurrence as the oracle at the native tolerance: same iteration count
    so = np.zeros(D)
    it0, rel0, acc0 = C.c_uint32(), C.c_double(), C.c_int32()
    oracle.bao_pcg(C.byref(hv), lam, 1e-1, 63, so.ctypes.data, C.byref(it0), C.byref(rel0), C.byref(acc0))
    code = adjuster.handle.lib.ba_hip_pcg(adjuster.handle.ptr, C.byref(dv), lam, 1e-1, 63, step.data_ptr(),
                                         C.byref(it), C.byref(rel), C.byref(acc))
    assert (it.value, acc.value) == (it0.value, acc0.value)
    assert rel_err(step.cpu().numpy(), so) < 1e-9


def _compare_solutions(a, b, tol_cost=1e-10, tol_param=1e-8, cost_floor=0.0, tol_final=None, point_norm=False,
                       tol_cov=1e-8):
    """Same LM trajectory (accept/reject sequence, per-iteration cost within tol_cost relative
    to max(|cost|, cost_floor)), same status, final cost within tol_final (default tol_cost) and
    parameters within tol_param (relative to max(|value|, 1); with point_norm, a point's three
    coordinates relative to max(|X|, 1)), and the same certificate (marginals within tol_cov)."""
    tol_final = tol_cost if tol_final is None else tol_final
    ta, tb = a.trace, b.trace
    n = min(len(ta["cost"]), len(tb["cost"]))
    assert len(ta["cost"]) == len(tb["cost"]), (ta["cost"], tb["cost"])
    assert np.array_equal(ta["accepted"], tb["accepted"])
    fin = np.isfinite(tb["cost"][:n])
    scale = np.maximum(np.abs(tb["cost"][:n][fin]), cost_floor)
    assert np.all(np.abs(ta["cost"][:n][fin] - tb["cost"][:n][fin]) <= tol_cost * scale)
    ra, rb = a.report, b.report
    assert ra.status == rb.status
    assert abs(ra.final_cost - rb.final_cost) <= tol_final * max(abs(rb.final_cost), cost_floor)
    for name in ("pose_center", "point_xyz", "intrinsic_focal", "intrinsic_k1", "intrinsic_k2", "pose_rotation"):
        x, y = getattr(a, name), getattr(b, name)
        scale = np.maximum(np.abs(y), 1.0)
        if point_norm and name == "point_xyz":
            scale = np.maximum(np.linalg.norm(y, axis=1, keepdims=True), 1.0)
        assert np.all(np.abs(x - y) <= tol_param * scale), name
    _compare_certificates(a, b, tol_cov)


def _compare_certificates(a, b, tol_cov=1e-8):
    """The post-loop certificate: the same observability verdict and pose rank, the same held
    groups, quarantined points and re-solve count, and the marginal covariances within tol_cov of
    each group's largest entry (NaN where neither formed one)."""
    ra, rb = a.report, b.report
    for name in ("observability", "pose_rank", "pose_dimension", "unconstrained_intrinsics", "weak_point_resolves",
                 "quarantined_points"):
        assert getattr(ra, name) == getattr(rb, name), (name, getattr(ra, name), getattr(rb, name))
    assert np.array_equal(a.intrinsic_unconstrained, b.intrinsic_unconstrained)
    assert np.array_equal(a.point_quarantined, b.point_quarantined)
    ca, cb = a.intrinsic_covariance, b.intrinsic_covariance
    assert np.array_equal(np.isnan(ca), np.isnan(cb))
    fin = ~np.isnan(cb).any(axis=(1, 2))
    if fin.any():
        scale = np.max(np.abs(cb[fin]), axis=(1, 2), keepdims=True)
        assert np.all(np.abs(ca[fin] - cb[fin]) <= tol_cov * scale)


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_ring_solve_matches_oracle(adjuster, oracle, ring, solver):
    sc, _, _ = ring
    cfg = SolveConfig(max_iterations=50, linear_solver=solver)
    got = adjuster.solve(sc.problem, cfg)
    want = solve_with(oracle, "bao_solve", sc.problem, cfg)
    _compare_solutions(got, want)


def test_reference_fixture_recovers_truth_direct(adjuster, oracle):
    """tests/sparse_self_calibration/dense_parity.rs:104-222: noise-free 4-camera fixture; the LM
    recovers truth (final kernel rmse < 1e-6) and matches the oracle's trajectory; objectives
    compared as the reference's assert_objective_close, |a - b| <= 5e-8 (1 + |a|)."""
    sc = synthetic.reference_fixture_scene()
    cfg = SolveConfig(max_iterations=80, function_tolerance=1e-12, parameter_tolerance=1e-10,
                      linear_solver=A.SOLVER_DIRECT)
    got = adjuster.solve(sc.problem, cfg)
    want = solve_with(oracle, "bao_solve", sc.problem, cfg)
    _compare_solutions(got, want, tol_cost=5e-8, tol_param=2e-5, cost_floor=1.0)
    assert got.report.final_rmse / np.sqrt(2.0) < 1e-6
    assert np.abs(got.point_xyz - sc.truth_points).max() < 3e-5
    assert np.abs(got.intrinsic_focal / sc.truth_focal - 1.0).max() < 2e-5


def test_reference_fixture_pcg_trajectory(adjuster, oracle):
    """The same fixture on the retained PCG (no direct fallback, native forcing tolerance):
    with the PCG cap min(600, D) = 29, rounding decides whether late solves reach the tolerance,
    so only the opening of the trajectory is compared; both must reduce the cost by 1e6."""
    sc = synthetic.reference_fixture_scene()
    cfg = SolveConfig(max_iterations=12, linear_solver=A.SOLVER_ITERATIVE)
    got = adjuster.solve(sc.problem, cfg)
    want = solve_with(oracle, "bao_solve", sc.problem, cfg)
    k = 6
    assert np.array_equal(got.trace["accepted"][:k], want.trace["accepted"][:k])
    assert np.allclose(got.trace["cost"][:k], want.trace["cost"][:k], rtol=1e-8, atol=0)
    assert got.report.final_cost < 1e-6 * got.report.initial_cost
