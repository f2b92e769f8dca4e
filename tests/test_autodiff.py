"""Differentiation form: robust reprojection cost and its hand-written gradient.

CPU tests pin the oracle's objective (bao_cost_grad): its gradient against FP64 central finite
differences (h = 1e-6) of its own cost, its cost against the per-observation radial projection of
the reference model (bao_radial_projection, itself pinned by test_oracle_pins), and the noise-free
truth as a stationary point with zero cost.  GPU tests compare ba_hip_cost_grad with the oracle
(FP64 and FP32), check the GPU gradient against central differences at the full 1,000-camera /
200,000-point / 1,000,000-observation size, and run the torch.autograd.Function.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from openmvg_ba import _abi as A
from openmvg_ba import synthetic


def oracle_cost_grad(oracle, cams, pts, oc, op, uv, loss, scale, grad=True):
    cams = np.ascontiguousarray(cams, dtype=np.float64)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    uv = np.ascontiguousarray(uv, dtype=np.float64)
    oc = np.ascontiguousarray(oc, dtype=np.uint32)
    op = np.ascontiguousarray(op, dtype=np.uint32)
    cost = C.c_double()
    invalid = C.c_uint32()
    gc = np.zeros_like(cams) if grad else None
    gp = np.zeros_like(pts) if grad else None
    oracle.bao_cost_grad(len(cams), len(pts), len(oc), cams.ctypes.data, pts.ctypes.data, oc.ctypes.data,
                         op.ctypes.data, uv.ctypes.data, loss, scale, C.byref(cost), C.byref(invalid),
                         gc.ctypes.data if grad else None, gp.ctypes.data if grad else None)
    return cost.value, invalid.value, gc, gp


def restricted_fd(oracle, cams, pts, oc, op, uv, loss, scale, kind, index, h=1e-6):
    """Central difference of the cost of the observations that touch one parameter (the rest of
    the sum does not depend on it), so the difference is not swamped by the rounding of ~1e6 terms."""
    if kind == "camera":
        c, k = divmod(index, 9)
        sel = np.nonzero(oc == c)[0]
        base, arr = cams, "cams"
    else:
        p, k = divmod(index, 3)
        sel = np.nonzero(op == p)[0]
        base, arr = pts, "pts"
    vals = []
    for s in (+1, -1):
        mod = base.copy()
        flat = mod.reshape(-1)
        flat[index] += s * h
        args = dict(cams=cams, pts=pts)
        args[arr] = mod
        vals.append(oracle_cost_grad(oracle, args["cams"], args["pts"], oc[sel], op[sel], uv[sel], loss, scale,
                                     grad=False)[0])
    return (vals[0] - vals[1]) / (2 * h)


def small_scene(seed=0xBA0005, n_cameras=12, n_points=300, n_obs=1500):
    return synthetic.bal_scene(seed, n_cameras, n_points, n_obs)


@pytest.mark.parametrize("loss,scale", [(A.LOSS_SQUARED, 1.0), (A.LOSS_HUBER, 1.0), (A.LOSS_CAUCHY, 2.0)])
def test_oracle_gradient_matches_central_differences(oracle, loss, scale):
    cams, pts, oc, op, uv = small_scene()
    cost, invalid, gc, gp = oracle_cost_grad(oracle, cams, pts, oc, op, uv, loss, scale)
    assert invalid == 0 and np.isfinite(cost)
    rng = np.random.default_rng(7)
    for kind, n, g in (("camera", cams.size, gc.reshape(-1)), ("point", pts.size, gp.reshape(-1))):
        for index in rng.choice(n, 40, replace=False):
            fd = restricted_fd(oracle, cams, pts, oc, op, uv, loss, scale, kind, int(index))
            assert abs(fd - g[index]) <= 1e-6 * max(1.0, abs(g[index])), (kind, index, fd, g[index])


def test_oracle_rotation_gradient_far_from_identity(oracle):
    """Large angle-axis vectors exercise the SO(3) left-Jacobian factor of the rotation gradient."""
    cams, pts, oc, op, uv = small_scene()
    cams = cams.copy()
    rng = np.random.default_rng(3)
    # rotate each camera by a large angle about its optical axis: points stay in front
    for c in range(len(cams)):
        axis = np.array([0.0, 0.0, 1.0]) + rng.normal(0, 0.02, 3)
        R = synthetic.exp_so3(axis / np.linalg.norm(axis) * rng.uniform(0.5, 2.5)) @ synthetic.exp_so3(cams[c, :3])
        cams[c, :3] = synthetic.log_so3(R)
    _, invalid, gc, _ = oracle_cost_grad(oracle, cams, pts, oc, op, uv, A.LOSS_HUBER, 1.0)
    assert invalid == 0
    for c in range(len(cams)):
        for k in range(3):
            fd = restricted_fd(oracle, cams, pts, oc, op, uv, A.LOSS_HUBER, 1.0, "camera", 9 * c + k)
            assert abs(fd - gc[c, k]) <= 1e-6 * max(1.0, abs(gc[c, k])), (c, k, fd, gc[c, k])


def test_oracle_cost_is_sum_of_radial_projections(oracle):
    cams, pts, oc, op, uv = small_scene(n_obs=700)
    cost, _, _, _ = oracle_cost_grad(oracle, cams, pts, oc, op, uv, A.LOSS_SQUARED, 1.0, grad=False)
    total = 0.0
    for i in range(len(oc)):
        c = cams[oc[i]]
        R = np.zeros(9)
        oracle.bao_exp_so3(np.ascontiguousarray(c[:3]).ctypes.data, R.ctypes.data)
        Xc = R.reshape(3, 3) @ pts[op[i]] + c[3:6]
        res = np.zeros(2)
        J = np.zeros(6)
        ok = oracle.bao_radial_projection(np.ascontiguousarray(c[6:9]).ctypes.data,
                                          np.ascontiguousarray(uv[i]).ctypes.data,
                                          np.ascontiguousarray(Xc).ctypes.data, res.ctypes.data, J.ctypes.data)
        assert ok
        total += 0.5 * float(res @ res)
    assert abs(cost - total) <= 1e-12 * total


def test_oracle_truth_is_stationary(oracle):
    """Noise-free observations at the true parameters: zero cost and zero gradient."""
    rng = np.random.default_rng(11)
    nc, npt = 6, 80
    cams = np.concatenate([rng.normal(0, 0.1, (nc, 3)), rng.uniform(-0.5, 0.5, (nc, 3)),
                           rng.uniform(800, 1200, (nc, 1)), rng.uniform(-0.1, 0.1, (nc, 1)),
                           rng.uniform(-0.05, 0.05, (nc, 1))], -1)
    pts = np.stack([rng.uniform(-2, 2, npt), rng.uniform(-2, 2, npt), rng.uniform(8, 12, npt)], -1)
    oc = np.repeat(np.arange(nc), npt).astype(np.uint32)
    op = np.tile(np.arange(npt), nc).astype(np.uint32)
    R = synthetic.exp_so3(cams[:, :3])
    pc = np.einsum("nij,nj->ni", R[oc], pts[op]) + cams[oc, 3:6]
    n = pc[:, :2] / pc[:, 2:3]
    r2 = np.sum(n * n, -1)
    uv = (cams[oc, 6] * (1 + cams[oc, 7] * r2 + cams[oc, 8] * r2 * r2))[:, None] * n
    cost, invalid, gc, gp = oracle_cost_grad(oracle, cams, pts, oc, op, uv, A.LOSS_HUBER, 1.0)
    assert invalid == 0
    assert cost < 1e-18
    assert np.abs(gc).max() < 1e-6 and np.abs(gp).max() < 1e-6


# ---------------------------------------------------------------------------------------------
# GPU


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch


def _gpu_eval(torch, scene, dtype, loss, scale, grad=True):
    from openmvg_ba.autodiff import ReprojectionTopology, cost_grad

    cams, pts, oc, op, uv = scene
    topo = ReprojectionTopology(oc, op, len(cams), len(pts))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)
    cost, invalid, gc, gp = cost_grad(topo, t(cams), t(pts), t(uv), loss, scale, with_grad=grad)
    torch.cuda.synchronize()
    out = (float(cost.item()), int(invalid.item()),
           gc.double().cpu().numpy() if grad else None, gp.double().cpu().numpy() if grad else None)
    topo.close()
    return out


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.gpu
@pytest.mark.parametrize("loss,scale", [(A.LOSS_SQUARED, 1.0), (A.LOSS_HUBER, 1.0), (A.LOSS_CAUCHY, 2.0)])
def test_gpu_cost_grad_fp64_matches_oracle(torch_cuda, oracle, loss, scale):
    scene = synthetic.bal_scene(0xBA0005, 50, 2000, 12000)
    cost, invalid, gc, gp = _gpu_eval(torch_cuda, scene, torch_cuda.float64, loss, scale)
    want = oracle_cost_grad(oracle, *scene, loss, scale)
    assert invalid == want[1] == 0
    assert abs(cost - want[0]) <= 1e-12 * abs(want[0])
    assert _rel(gc, want[2]) <= 1e-11
    assert _rel(gp, want[3]) <= 1e-11


@pytest.mark.gpu
def test_gpu_cost_grad_fp32_error_vs_fp64_oracle(torch_cuda, oracle):
    """FP32 form: per-observation arithmetic in float, sums in double; the stated error bound
    against the FP64 oracle is 1e-5 relative for the cost and 2e-3 (max-norm) for gradients."""
    scene = synthetic.bal_scene(0xBA0005, 50, 2000, 12000)
    cost, invalid, gc, gp = _gpu_eval(torch_cuda, scene, torch_cuda.float32, A.LOSS_HUBER, 1.0)
    want = oracle_cost_grad(oracle, *scene, A.LOSS_HUBER, 1.0)
    assert invalid == 0
    assert abs(cost - want[0]) <= 1e-5 * abs(want[0])
    assert _rel(gc, want[2]) <= 2e-3
    assert _rel(gp, want[3]) <= 2e-3


@pytest.mark.gpu
def test_gpu_gradient_full_size_central_differences(torch_cuda, oracle):
    """1,000 cameras, 200,000 points, 1,000,000 observations, Huber: the GPU gradient against the
    oracle's analytic gradient (whole vector) and against central differences (h = 1e-6) on a
    seeded subset of 1,000 parameters."""
    scene = synthetic.bal_scene()
    cams, pts, oc, op, uv = scene
    cost, invalid, gc, gp = _gpu_eval(torch_cuda, scene, torch_cuda.float64, A.LOSS_HUBER, 1.0)
    want = oracle_cost_grad(oracle, *scene, A.LOSS_HUBER, 1.0)
    assert invalid == want[1]
    assert abs(cost - want[0]) <= 1e-12 * abs(want[0])
    assert _rel(gc, want[2]) <= 1e-10 and _rel(gp, want[3]) <= 1e-10
    rng = np.random.default_rng(0xFD)
    n_cam = 9 * len(cams)
    picks = rng.choice(n_cam + 3 * len(pts), 1000, replace=False)
    worst = 0.0
    for idx in picks:
        if idx < n_cam:
            g = gc.reshape(-1)[idx]
            fd = restricted_fd(oracle, cams, pts, oc, op, uv, A.LOSS_HUBER, 1.0, "camera", int(idx))
        else:
            g = gp.reshape(-1)[idx - n_cam]
            fd = restricted_fd(oracle, cams, pts, oc, op, uv, A.LOSS_HUBER, 1.0, "point", int(idx - n_cam))
        worst = max(worst, abs(fd - g) / max(1.0, abs(g)))
    assert worst <= 1e-6, worst


@pytest.mark.gpu
def test_autograd_function(torch_cuda, oracle):
    torch = torch_cuda
    from openmvg_ba.autodiff import ReprojectionTopology, reprojection_cost

    cams, pts, oc, op, uv = synthetic.bal_scene(0xBA0005, 30, 800, 5000)
    topo = ReprojectionTopology(oc, op, len(cams), len(pts))
    tc = torch.tensor(cams, device="cuda", requires_grad=True)
    tp = torch.tensor(pts, device="cuda", requires_grad=True)
    tuv = torch.tensor(uv, device="cuda")
    cost = reprojection_cost(tc, tp, tuv, topo, A.LOSS_HUBER, 1.0)
    (3.0 * cost).backward()
    want = oracle_cost_grad(oracle, cams, pts, oc, op, uv, A.LOSS_HUBER, 1.0)
    assert abs(cost.item() - want[0]) <= 1e-12 * abs(want[0])
    assert _rel(tc.grad.cpu().numpy(), 3.0 * want[2]) <= 1e-11
    assert _rel(tp.grad.cpu().numpy(), 3.0 * want[3]) <= 1e-11
    # torch.autograd.gradcheck on a tiny scene (FP64 central differences inside torch)
    cams, pts, oc, op, uv = synthetic.bal_scene(0xBA0006, 3, 12, 40)
    topo2 = ReprojectionTopology(oc, op, len(cams), len(pts))
    tc = torch.tensor(cams, device="cuda", requires_grad=True)
    tp = torch.tensor(pts, device="cuda", requires_grad=True)
    tuv = torch.tensor(uv, device="cuda")
    assert torch.autograd.gradcheck(lambda a, b: reprojection_cost(a, b, tuv, topo2, A.LOSS_HUBER, 1.0), (tc, tp),
                                    eps=1e-6, atol=1e-5, rtol=1e-6)
    topo.close()
    topo2.close()


@pytest.mark.gpu
def test_cost_grad_deterministic(torch_cuda):
    scene = synthetic.bal_scene(0xBA0005, 100, 20000, 150000)
    a = _gpu_eval(torch_cuda, scene, torch_cuda.float64, A.LOSS_HUBER, 1.0)
    b = _gpu_eval(torch_cuda, scene, torch_cuda.float64, A.LOSS_HUBER, 1.0)
    assert a[0] == b[0] and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
