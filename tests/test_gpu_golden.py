"""The product on an MI355X against the committed golden fixtures (tests/golden/, oracle-generated
on the seeded ring scene), through the C ABI, to the plain form's bars: per-observation values
1e-12 relative, the reduced system 1e-10 relative to its largest entry, per-iteration cost 1e-10
relative with the same accept/reject sequence, final parameters 1e-8 relative."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from golden_lib import arr, load
from helpers import device_view, native_host, rel_err, to_device
from openmvg_ba import SolveConfig, synthetic
from openmvg_ba import _abi as A
from openmvg_ba.problem import HipBundleAdjuster

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


@pytest.fixture(scope="module")
def ring(oracle):
    sc = synthetic.ring_scene()
    nat, arrays = native_host(oracle, sc.problem)
    return sc, nat, arrays


@pytest.mark.parametrize("loss,scale,key", [(A.LOSS_SQUARED, 0.0, "squared"), (A.LOSS_HUBER, 1.0, "huber")])
def test_linearize_matches_golden(torch_cuda, adjuster, ring, loss, scale, key):
    torch = torch_cuda
    _, nat, arrays = ring
    g = load("ring_linearize")
    n, no = g["n_observations"], nat.sizes["n_observations"]
    dev = to_device(arrays)  # the view holds raw pointers: keep the tensors alive
    dv = device_view(nat, dev, arrays, loss, scale)
    out = [torch.zeros(m, dtype=torch.float64, device="cuda") for m in (2 * no, 6 * no, 18 * no, no)]
    valid = torch.zeros(no, dtype=torch.uint8, device="cuda")
    code = adjuster.handle.lib.ba_hip_linearize(adjuster.handle.ptr, C.byref(dv), *(t.data_ptr() for t in out),
                                               valid.data_ptr())
    assert code == A.BA_OK
    want = g[key]
    assert np.array_equal(valid.cpu().numpy()[:n], np.asarray(want["valid"]))
    for got, name, m in zip(out, ("residual", "point_jacobian", "camera_jacobian", "cost"), (2, 6, 18, 1)):
        assert rel_err(got.cpu().numpy()[:m * n], arr(want[name])) < 1e-12, name


def test_schur_matches_golden(torch_cuda, adjuster, ring):
    torch = torch_cuda
    _, nat, arrays = ring
    g = load("ring_schur")
    D = g["dimension"]
    dev = to_device(arrays)
    dv = device_view(nat, dev, arrays, A.LOSS_SQUARED, 0.0)
    S = torch.zeros(D * D, dtype=torch.float64, device="cuda")
    gr, fg, dg = (torch.zeros(D, dtype=torch.float64, device="cuda") for _ in range(3))
    s1 = A.SchurScalars()
    code = adjuster.handle.lib.ba_hip_schur(adjuster.handle.ptr, C.byref(dv), g["damping"], S.data_ptr(),
                                           gr.data_ptr(), fg.data_ptr(), dg.data_ptr(), C.byref(s1))
    assert code == A.BA_OK
    assert rel_err(S.cpu().numpy(), arr(g["S"])) < 1e-10
    for got, name in ((gr, "gradient"), (fg, "full_gradient"), (dg, "diagonal")):
        assert rel_err(got.cpu().numpy(), arr(g[name])) < 1e-10, name
    assert abs(s1.cost - g["cost"]) <= 1e-12 * abs(g["cost"])
    assert s1.weak_point_count == g["weak_point_count"]


@pytest.mark.parametrize("solver,key", [(A.SOLVER_DIRECT, "direct"), (A.SOLVER_ITERATIVE, "iterative")])
def test_solve_matches_golden(adjuster, ring, solver, key):
    sc, _, _ = ring
    g = load("ring_solve")
    want = g[key]
    sol = adjuster.solve(sc.problem, SolveConfig(max_iterations=g["max_iterations"], linear_solver=solver))
    assert np.array_equal(sol.trace["accepted"], np.asarray(want["accepted"]))
    cost = arr(want["cost"])
    assert np.all(np.abs(sol.trace["cost"] - cost) <= 1e-10 * np.abs(cost))
    assert abs(sol.report.final_cost - want["final_cost"]) <= 1e-10 * abs(want["final_cost"])
    assert sol.report.status == want["status"]
    for name in ("pose_rotation", "pose_center", "intrinsic_focal", "point_xyz"):
        x, y = getattr(sol, name).ravel(), arr(want[name])
        assert np.all(np.abs(x - y) <= 1e-8 * np.maximum(np.abs(y), 1.0)), name
