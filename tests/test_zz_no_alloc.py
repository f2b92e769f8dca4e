"""The allocation-free dense solve (ba_hip_cholesky_solve_noalloc) against the allocating one
(ba_hip_cholesky_solve), through the C ABI, on an MI355X: the same kernels on the same inputs, so
the solutions are bit-identical, on two calls in a row (the second smaller, so the shared scratch
holds the first call's contents), in FP64 and in the mixed form; sizes beyond
BA_NOALLOC_MAX_DIMENSION are refused; a non-SPD matrix fails the same way in both."""

from __future__ import annotations

import numpy as np
import pytest

from openmvg_ba import _abi as A
from openmvg_ba.problem import HipBundleAdjuster

pytestmark = pytest.mark.gpu

MAX_DIMENSION = 2048


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def adjuster(torch_cuda):
    return HipBundleAdjuster(0)


def spd(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.normal(size=(n, n))
    d = 10.0 ** rng.uniform(0.0, 6.0, n)  # a camera system's spread of scales
    return (M @ M.T + n * np.eye(n)) * np.sqrt(np.outer(d, d)), rng.normal(size=n) * np.sqrt(d)


def solve_both(torch, adjuster, A_, b, precision):
    lib, h = adjuster.handle.lib, adjuster.handle.ptr
    n = len(b)
    dA = torch.from_numpy(A_).cuda()
    db = torch.from_numpy(b).cuda()
    x_alloc = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    x_free = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    c1 = lib.ba_hip_cholesky_solve(h, n, dA.data_ptr(), db.data_ptr(), x_alloc.data_ptr(), precision)
    c2 = lib.ba_hip_cholesky_solve_noalloc(h, n, dA.data_ptr(), db.data_ptr(), x_free.data_ptr(), precision)
    return c1, c2, x_alloc.cpu().numpy(), x_free.cpu().numpy()


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
def test_noalloc_equals_allocating_solve_on_two_calls(torch_cuda, adjuster, precision):
    for n, seed in ((1500, 1), (197, 2)):  # the second call smaller than the first
        A_, b = spd(n, seed)
        c1, c2, xa, xf = solve_both(torch_cuda, adjuster, A_, b, precision)
        assert c1 == c2 == A.BA_OK
        assert np.array_equal(xa, xf), (n, np.abs(xa - xf).max())
        x_ref = np.linalg.solve(A_, b)
        assert np.abs(xf - x_ref).max() <= 1e-8 * np.abs(x_ref).max()


def test_noalloc_limits_and_failures(torch_cuda, adjuster):
    torch = torch_cuda
    lib, h = adjuster.handle.lib, adjuster.handle.ptr
    # above the maximum: refused before any device work
    n = MAX_DIMENSION + 1
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert lib.ba_hip_cholesky_solve_noalloc(h, n, dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(),
                                             A.PRECISION_FP64) == A.BA_ERR_RESOURCE_EXHAUSTED
    assert lib.ba_hip_cholesky_solve_noalloc(h, 4, dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(),
                                             7) == A.BA_ERR_INVALID_CONFIG
    # an indefinite matrix fails in both, after a successful call left its factor in the scratch
    A_, b = spd(70, 3)
    assert solve_both(torch, adjuster, A_, b, A.PRECISION_FP64)[:2] == (A.BA_OK, A.BA_OK)
    bad = np.eye(70)
    bad[40, 40] = -1.0
    c1, c2, _, _ = solve_both(torch, adjuster, bad, np.ones(70), A.PRECISION_FP64)
    assert c1 == c2 == A.BA_ERR_NUMERICAL
    # and the next good call succeeds again (the failure flag is cleared by the kernels)
    c1, c2, xa, xf = solve_both(torch, adjuster, A_, b, A.PRECISION_FP64)
    assert c1 == c2 == A.BA_OK and np.array_equal(xa, xf)
