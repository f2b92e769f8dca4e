"""The oracle's post-loop certification, pinned independently of the oracle, and its solve flow.

Pins:
* rank and marginals (bao_certify): from the oracle's per-observation Jacobians (themselves
  pinned by central differences in tests/test_oracle_assembly.py) the test forms the undamped
  reduced system S in numpy by dense normal equations and point elimination; the oracle's verdict
  is the reference's spectral rank rule (direct.rs:342-457) computed in numpy: after diagonal
  equilibration, FULL_RANK when every eigenvalue exceeds 64 eps D lambda_max (clearly so here), and its
  marginal covariance of each free intrinsic group equals the group's diagonal block of numpy's
  inverse of S (crates/openmvg-ba/src/kernel/multi_view/sparse/covariance.rs:562-587), to 1e-8;
  a camera seen through one point makes S singular (numpy: a null eigenvalue) and the oracle
  reports RANK_DEFICIENT.
* weak points (bao_weak_points, solver.rs:1149-1219): each point's 3x3 information at a solution
  rebuilt in numpy from central-difference Jacobians of the pinhole/Brown projection in the
  caller's frame, weighted by the loss, with numpy's eigenvalues and the 1e-5 ratio; the
  degenerate-geometry verdict is the track-geometry rule pinned in tests/test_oracle_pins.py.

Flow (native/mod.rs:252-350, native/self_calibration.rs:223-303): the ring scene certifies full
rank with nothing held; the narrow-field scene holds every group (k1, k2 unconstrained at the
image corner) and re-solves from the input; the rank-deficient scene ends degenerate with every
group held (no marginals); the candidate-rule scene re-solves once without its near and far
points, which keep their input positions.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from helpers import host_view, native_host
from openmvg_ba import SolveConfig, _abi as A, synthetic
from openmvg_ba.problem import solve_with
from test_oracle_assembly import dense_jacobian, linearize


@pytest.fixture(scope="module")
def cert(oracle):
    oracle.bao_certify.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    oracle.bao_weak_points.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    oracle.bao_linearize.argtypes = [C.c_void_p] * 6
    return oracle


def undamped_schur(nat, arrays, J, r):
    """S = U - W V^-1 W^T at zero damping, from the dense Jacobian."""
    D, np_ = nat.sizes["dimension"], nat.sizes["n_points"]
    H = J.T @ J
    S = H[:D, :D].copy()
    fixed = arrays["point_fixed"].astype(bool)
    for q in range(np_):
        if fixed[q]:
            continue
        sl = slice(D + 3 * q, D + 3 * q + 3)
        W = H[:D, sl]
        S -= W @ np.linalg.inv(H[sl, sl]) @ W.T
    return S


def spectral_rank(S):
    """The reference's rank of a reduced system (direct.rs:402-457): diagonal equilibration with
    the floor 64 eps max diag, then the eigenvalues above 64 eps D lambda_max."""
    d = np.diag(S)
    floor = 64.0 * np.finfo(float).eps * np.abs(d).max()
    s = 1.0 / np.sqrt(np.maximum(d, floor))
    ev = np.linalg.eigvalsh(S * s[:, None] * s[None, :])
    tol = 64.0 * np.finfo(float).eps * len(S) * np.abs(ev).max()
    return int((ev > tol).sum())


def certify(oracle, nat, arrays, loss=A.LOSS_SQUARED, scale=0.0):
    G = nat.sizes["n_groups"]
    obs = C.c_int32()
    cov = np.zeros(9 * G)
    view = host_view(nat, arrays, loss, scale)
    code = oracle.bao_certify(C.byref(view), C.byref(obs), cov.ctypes.data)
    return code, obs.value, cov.reshape(G, 3, 3)


@pytest.mark.parametrize("scene,loss", [("ring", A.LOSS_SQUARED), ("ring", A.LOSS_HUBER),
                                        ("fixture", A.LOSS_SQUARED)])
def test_marginals_are_blocks_of_the_inverse_reduced_system(cert, scene, loss):
    oracle = cert
    prob = {"ring": lambda: synthetic.ring_scene(n_points=200).problem,
            "fixture": lambda: synthetic.reference_fixture_scene().problem}[scene]()
    scale = 1.0 if loss == A.LOSS_HUBER else 0.0
    nat, arrays = native_host(oracle, prob)
    r, jx, jc = linearize(oracle, nat, arrays, loss, scale)
    S = undamped_schur(nat, arrays, dense_jacobian(nat, arrays, jx, jc), r)
    ev = np.linalg.eigvalsh(S)
    assert ev[0] > 1e-9 * ev[-1]  # clearly positive definite
    assert spectral_rank(S) == len(S)
    code, obs, cov = certify(oracle, nat, arrays, loss, scale)
    assert code == A.BA_OK and obs == A.OBSERVABILITY_FULL_RANK
    inv = np.linalg.inv(S)
    worst = 0.0
    for g, (off, dim) in enumerate(zip(arrays["group_offset"], arrays["group_dim"])):
        if dim == 0:
            assert np.isnan(cov[g]).all()
            continue
        want = np.zeros((3, 3))
        want[:dim, :dim] = inv[off:off + dim, off:off + dim]
        err = np.abs(cov[g] - want).max() / np.abs(want).max()
        worst = max(worst, err)
    print(f"{scene}: marginal error {worst:.2e}")
    assert worst <= 1e-8


def test_singular_reduced_system_is_rank_deficient(cert):
    oracle = cert
    nat, arrays = native_host(oracle, synthetic.rank_deficient_scene())
    r, jx, jc = linearize(oracle, nat, arrays, A.LOSS_SQUARED, 0.0)
    S = undamped_schur(nat, arrays, dense_jacobian(nat, arrays, jx, jc), r)
    ev = np.linalg.eigvalsh(S)
    assert ev[0] <= 1e-9 * ev[-1]  # a null direction (the lone camera's unobserved motion)
    assert spectral_rank(S) < len(S)
    code, obs, cov = certify(oracle, nat, arrays)
    assert obs == A.OBSERVABILITY_RANK_DEFICIENT and code == A.BA_ERR_NUMERICAL
    assert np.isnan(cov).all()


def _project(R, Cc, f, k1, k2, cx, cy, X):
    pc = R @ (X - Cc)
    x, y = pc[0] / pc[2], pc[1] / pc[2]
    q = x * x + y * y
    d = 1.0 + k1 * q + k2 * q * q
    return np.array([f * x * d + cx, f * y * d + cy])


def numpy_weak_points(prob, sol, track_degenerate):
    """Solution::weak_points restated with central-difference Jacobians and numpy eigenvalues."""
    N = prob.n_points
    info = np.zeros((N, 3, 3))
    scale = prob.loss_scale / np.sqrt(2.0) if prob.robust_loss == A.LOSS_CAUCHY else prob.loss_scale
    for o in range(prob.n_observations):
        q, c, g = int(prob.obs_point[o]), int(prob.obs_pose[o]), int(prob.obs_intrinsic[o])
        args = (sol.pose_rotation[c], sol.pose_center[c], sol.intrinsic_focal[g], sol.intrinsic_k1[g],
                sol.intrinsic_k2[g], prob.intrinsics[g].cx, prob.intrinsics[g].cy)
        X = sol.point_xyz[q]
        res = np.linalg.norm(_project(*args, X) - prob.obs_pixel[o])
        h = 1e-6 * max(1.0, np.abs(X).max())
        J = np.stack([(_project(*args, X + h * e) - _project(*args, X - h * e)) / (2 * h) for e in np.eye(3)], 1)
        if prob.robust_loss == A.LOSS_HUBER:
            w2 = 1.0 if res <= scale else scale / res
        elif prob.robust_loss == A.LOSS_CAUCHY:
            w2 = 1.0 / (1.0 + 0.5 * (res / scale) ** 2)
        else:
            w2 = 1.0
        info[q] += w2 * J.T @ J
    weak = np.zeros(N, dtype=bool)
    for q in range(N):
        ev = np.linalg.eigvalsh(info[q])
        weak[q] = track_degenerate[q] or not (ev[0] > 1e-5 * ev[-1])
    return weak


def test_weak_points_match_a_numpy_restatement(cert):
    """On the candidate-rule scene's loop-only solution the near point (next to camera 3's centre)
    and the far point (parallax just above the track threshold) are weak; every other verdict
    agrees with the numpy restatement."""
    oracle = cert
    prob, q_near, q_far = synthetic.candidate_rule_scene()
    sol = solve_with(oracle, "bao_solve", prob,
                     SolveConfig(max_iterations=30, linear_solver=A.SOLVER_DIRECT, certification=A.CERTIFY_LOOP_ONLY))
    s = A.SolutionC()
    keep = [np.ascontiguousarray(a) for a in (sol.pose_rotation, sol.pose_center, sol.intrinsic_focal,
                                             sol.intrinsic_k1, sol.intrinsic_k2, sol.point_xyz)]
    s.pose_rotation, s.pose_center, s.intrinsic_focal, s.intrinsic_k1, s.intrinsic_k2, s.point_xyz = [
        A.ptr(a, C.c_double) for a in keep]
    cp, keep_p = prob.to_c()
    weak = np.zeros(prob.n_points, dtype=np.uint8)
    oracle.bao_weak_points(C.byref(cp), C.byref(s), weak.ctypes.data)
    # track geometry at the solution, by the oracle's pinned TrackGeometry
    offs = sol.point_xyz[prob.obs_point] - sol.pose_center[prob.obs_pose]
    scene_len = np.partition(np.linalg.norm(offs, axis=1), len(offs) // 2)[len(offs) // 2]
    oracle.bao_track_geometry.restype = C.c_int
    degenerate = np.zeros(prob.n_points, dtype=bool)
    for q in range(prob.n_points):
        o = np.ascontiguousarray(offs[prob.obs_point == q])
        degenerate[q] = bool(oracle.bao_track_geometry(len(o), o.ctypes.data, scene_len))
    want = numpy_weak_points(prob, sol, degenerate)
    assert weak[q_near] and weak[q_far]
    assert np.array_equal(weak.astype(bool), want), np.nonzero(weak.astype(bool) != want)


def _solve(oracle, prob, solver, **kw):
    return solve_with(oracle, "bao_solve", prob, SolveConfig(max_iterations=50, linear_solver=solver, **kw),
                      trace_capacity=400)


@pytest.mark.parametrize("solver", [A.SOLVER_DIRECT, A.SOLVER_ITERATIVE])
def test_certified_flows(cert, solver):
    oracle = cert
    # ring: full rank, constrained, nothing held or re-solved
    ring = _solve(oracle, synthetic.ring_scene().problem, solver)
    r = ring.report
    assert r.observability == A.OBSERVABILITY_FULL_RANK and r.pose_rank == r.pose_dimension > 0
    assert r.unconstrained_intrinsics == 0 and r.weak_point_resolves == 0 and r.status.startswith("converged")
    assert np.isfinite(ring.intrinsic_covariance[:, 0, 0]).all()
    # narrow field: every group held at its input values, the second solve's report
    sc = synthetic.narrow_field_scene()
    narrow = _solve(oracle, sc.problem, solver)
    r = narrow.report
    assert r.unconstrained_intrinsics == len(sc.problem.intrinsics) and narrow.intrinsic_unconstrained.all()
    assert np.array_equal(narrow.intrinsic_focal, [i.focal for i in sc.problem.intrinsics])
    assert np.array_equal(narrow.intrinsic_k1, [i.k1 for i in sc.problem.intrinsics])
    assert r.observability == A.OBSERVABILITY_FULL_RANK and len(narrow.trace["cost"]) == r.iterations
    sd_k1 = np.sqrt(narrow.intrinsic_covariance[:, 1, 1])
    assert (sd_k1 * 0.41 > 0.01).all()  # corner radius^2 = 0.5^2 + 0.375^2 at f ~ 1000
    # rank deficient: degenerate, every group held (no marginals), nothing to quarantine
    rd = _solve(oracle, synthetic.rank_deficient_scene(), solver)
    r = rd.report
    assert r.observability == A.OBSERVABILITY_RANK_DEFICIENT and r.pose_rank == 0 and r.status == "degenerate"
    assert rd.intrinsic_unconstrained.all() and np.isnan(rd.intrinsic_covariance).all()
    # candidate rule: one weak-point re-solve without the near and far points
    prob, q_near, q_far = synthetic.candidate_rule_scene()
    cand = _solve(oracle, prob, solver)
    r = cand.report
    assert r.weak_point_resolves == 1 and r.quarantined_points == 2
    assert np.nonzero(cand.point_quarantined)[0].tolist() == [q_near, q_far]
    assert np.array_equal(cand.point_xyz[[q_near, q_far]], prob.point_xyz[[q_near, q_far]])
    assert r.observability == A.OBSERVABILITY_FULL_RANK and r.weak_point_count == 0
    assert len(cand.trace["cost"]) == r.iterations  # both solves' iterations
    # loop only: no certificate
    lo = _solve(oracle, synthetic.ring_scene().problem, solver, certification=A.CERTIFY_LOOP_ONLY)
    assert lo.report.observability == A.OBSERVABILITY_NOT_ASSESSED and np.isnan(lo.intrinsic_covariance).all()
    assert lo.report.final_cost == ring.report.final_cost
