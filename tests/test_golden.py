"""The committed golden fixtures (tests/golden/, oracle-generated on the seeded ring scene) against
the oracle and the product's host lowering, on the CPU.

The oracle is deterministic, so it must reproduce its fixtures to rounding (1e-13 relative, the
trajectories to 1e-12); the product's lowering must reproduce the index maps exactly.  The GPU
side of the same fixtures is tests/test_gpu_golden.py."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from golden_lib import arr, load, scene_digest
from helpers import host_view, native_host, oracle_assemble, rel_err
from openmvg_ba import SolveConfig, _lib, synthetic
from openmvg_ba import _abi as A
from openmvg_ba.problem import prepare_with, solve_with


@pytest.fixture(scope="module")
def ring(oracle):
    sc = synthetic.ring_scene()
    nat, arrays = native_host(oracle, sc.problem)
    return sc, nat, arrays


def test_fixtures_were_made_from_this_scene(ring):
    sc, _, _ = ring
    digest = scene_digest(sc.problem)
    for name in ("ring_index_maps", "ring_linearize", "ring_schur", "ring_solve"):
        assert load(name)["scene_sha256"] == digest, name


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_index_maps(ring, oracle, impl):
    sc, _, _ = ring
    g = load("ring_index_maps")
    nat = prepare_with(oracle, "bao_prepare", sc.problem) if impl == "oracle" else \
        prepare_with(_lib.load(), "ba_hip_prepare", sc.problem)
    assert {k: int(v) for k, v in nat.sizes.items()} == g["sizes"]
    assert nat.gauge_camera == g["gauge_camera"] and nat.scene_scale == g["scene_scale"]
    for k in ("camera_offset", "camera_dim", "group_offset", "group_dim", "camera_group", "point_offsets",
              "obs_camera", "point_fixed"):
        assert np.array_equal(np.asarray(nat.arrays[k]).ravel(), np.asarray(g[k])), k


@pytest.mark.parametrize("loss,scale,key", [(A.LOSS_SQUARED, 0.0, "squared"), (A.LOSS_HUBER, 1.0, "huber")])
def test_oracle_linearize(ring, oracle, loss, scale, key):
    _, nat, arrays = ring
    g = load("ring_linearize")
    n, no = g["n_observations"], nat.sizes["n_observations"]
    hv = host_view(nat, arrays, loss, scale)
    r, j, c, k, v = np.zeros(2 * no), np.zeros(6 * no), np.zeros(18 * no), np.zeros(no), np.zeros(no, np.uint8)
    oracle.bao_linearize(C.byref(hv), *(x.ctypes.data_as(C.c_void_p) for x in (r, j, c, k, v)))
    want = g[key]
    assert np.array_equal(v[:n], np.asarray(want["valid"]))
    for got, name, m in ((r, "residual", 2), (j, "point_jacobian", 6), (c, "camera_jacobian", 18), (k, "cost", 1)):
        assert rel_err(got[:m * n], arr(want[name])) <= 1e-13, name


def test_oracle_schur(ring, oracle):
    _, nat, arrays = ring
    g = load("ring_schur")
    D = nat.sizes["dimension"]
    S, gr, fg, dg, sc = oracle_assemble(oracle, host_view(nat, arrays, A.LOSS_SQUARED, 0.0), D,
                                        nat.sizes["n_points"], nat.sizes["n_point_parameters"], g["damping"])
    assert D == g["dimension"]
    assert rel_err(S.ravel(), arr(g["S"])) <= 1e-13
    for got, name in ((gr, "gradient"), (fg, "full_gradient"), (dg, "diagonal")):
        assert rel_err(got, arr(g[name])) <= 1e-13, name
    assert abs(sc.cost - g["cost"]) <= 1e-13 * abs(g["cost"])
    assert sc.weak_point_count == g["weak_point_count"]


@pytest.mark.parametrize("solver,key", [(A.SOLVER_DIRECT, "direct"), (A.SOLVER_ITERATIVE, "iterative")])
def test_oracle_solve(ring, oracle, solver, key):
    sc, _, _ = ring
    g = load("ring_solve")
    want = g[key]
    sol = solve_with(oracle, "bao_solve", sc.problem, SolveConfig(max_iterations=g["max_iterations"],
                                                                  linear_solver=solver))
    assert np.array_equal(sol.trace["accepted"], np.asarray(want["accepted"]))
    assert rel_err(sol.trace["cost"], arr(want["cost"])) <= 1e-12
    assert abs(sol.report.final_cost - want["final_cost"]) <= 1e-12 * abs(want["final_cost"])
    assert rel_err(sol.point_xyz.ravel(), arr(want["point_xyz"])) <= 1e-10
    assert rel_err(sol.pose_center.ravel(), arr(want["pose_center"])) <= 1e-10
