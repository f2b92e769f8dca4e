"""Pins of the oracle's reduced camera system and PCG that do not use the oracle's own assembly.

* Jacobian vs the retraction: the oracle's per-observation Jacobians (bao_linearize) equal central
  differences (h = 1e-6) of its residuals along the reference's retraction coordinates: rotation
  exp(w) R (optimizer.rs:386-437), translation t + s dt, the gauge camera's translation
  b * retract_s2 (the FixedBaseline chart, manifold.rs:144-158), intrinsics f e^d0, k1 + d1,
  k2 + d2 (self_calibration/optimizer.rs:565-586), points X + s dX.  A wrong chart column, sign or
  scale fails it.
* Dense normal equations: from those Jacobians the test forms J (all observations x all
  parameters) in numpy, H = J^T J and g = J^T r~, applies the reference's damping (camera diagonal
  clamped per block to >= 64 eps blockmax, unit for an empty block: schur/assembly.rs:691-766;
  point blocks V^ = V + lambda max(diag V, 64 eps maxdiag V): point_block.rs:85-110) and
  eliminates every point with numpy's inverse: S = U + lambda diag - W V^-1 W^T,
  g^ = g_c - W V^-1 g_p.  The oracle's bao_assemble must match to 1e-10 of the largest entry.
* PCG: at relative tolerance 1e-10 the oracle's PCG (bao_pcg) equals numpy's solve of that S, on
  the scene scaled by 1e-6, 1 and 1e6 (the reference's gate,
  tests/sparse_scale_backend_differential.rs:60-80).

Scenes: the 10-camera ring (free gauge: the native FixedBaseline chart on the farthest camera,
5 pose columns) and the reference's 4-camera fixture with Gauge::FixedBaseline(0, 1), squared and
Huber loss, fixed points and a fixed pose.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from helpers import host_view, native_host
from openmvg_ba import _abi as A, synthetic

EPS = np.finfo(np.float64).eps


def exp_so3(oracle, w):
    R = np.zeros(9)
    oracle.bao_exp_so3(np.ascontiguousarray(w, dtype=np.float64).ctypes.data, R.ctypes.data)
    return R.reshape(3, 3)


def linearize(oracle, nat, arrays, loss, scale, state=None):
    no = nat.sizes["n_observations"]
    view = host_view(nat, arrays, loss, scale, state=state)
    r = np.zeros(2 * no)
    jx = np.zeros(6 * no)
    jc = np.zeros(18 * no)
    cost = np.zeros(no)
    valid = np.zeros(no, dtype=np.uint8)
    oracle.bao_linearize(C.byref(view), r.ctypes.data, jx.ctypes.data, jc.ctypes.data, cost.ctypes.data,
                         valid.ctypes.data)
    assert valid.all()
    return r.reshape(no, 2), jx.reshape(no, 2, 3), jc.reshape(no, 2, 9)


def columns(nat, arrays):
    """Global column of every (observation, camera-Jacobian column) and the point columns."""
    D = nat.sizes["dimension"]
    cam = arrays["obs_camera"].astype(np.int64)
    grp = arrays["camera_group"][cam].astype(np.int64)
    return D, cam, grp


def dense_jacobian(nat, arrays, jx, jc):
    D, cam, grp = columns(nat, arrays)
    no, np_ = nat.sizes["n_observations"], nat.sizes["n_points"]
    J = np.zeros((2 * no, D + 3 * np_))
    pt = arrays["obs_point"].astype(np.int64)
    fixed = arrays["point_fixed"].astype(bool)
    for o in range(no):
        c, g = cam[o], grp[o]
        co, cd = arrays["camera_offset"][c], arrays["camera_dim"][c]
        go, gd = arrays["group_offset"][g], arrays["group_dim"][g]
        J[2 * o:2 * o + 2, co:co + cd] = jc[o][:, :cd]
        J[2 * o:2 * o + 2, go:go + gd] = jc[o][:, 6:6 + gd]
        if not fixed[pt[o]]:
            J[2 * o:2 * o + 2, D + 3 * pt[o]:D + 3 * pt[o] + 3] = jx[o]
    return J


def reference_reduction(nat, arrays, J, r, lam):
    """S, g^, g_c and the clamped damping diagonal from the dense J by numpy."""
    D, np_ = nat.sizes["dimension"], nat.sizes["n_points"]
    H = J.T @ J
    g = J.T @ r.reshape(-1)
    U, W, gc = H[:D, :D], H[:D, D:], g[:D]
    diag = np.diag(U).copy()
    for off, dim in list(zip(arrays["camera_offset"], arrays["camera_dim"])) + \
            list(zip(arrays["group_offset"], arrays["group_dim"])):
        if dim == 0:
            continue
        blk = diag[off:off + dim]
        scale = blk.max()
        floor = 64.0 * EPS * scale if scale > 0 else 1.0
        diag[off:off + dim] = np.maximum(blk, floor)
    S = U + lam * np.diag(diag)
    gr = gc.copy()
    fixed = arrays["point_fixed"].astype(bool)
    for q in range(np_):
        if fixed[q]:
            continue
        sl = slice(D + 3 * q, D + 3 * q + 3)
        V = H[sl, sl]
        dv = np.diag(V)
        damped = V + lam * np.diag(np.maximum(dv, 64.0 * EPS * dv.max()))
        Vi = np.linalg.inv(damped)
        Wq = W[:, 3 * q:3 * q + 3]
        S -= Wq @ Vi @ Wq.T
        gr -= Wq @ (Vi @ g[sl])
    return S, gr, gc, diag


def perturbed_state(oracle, nat, arrays, col, h):
    """The state retracted by h along native coordinate `col` (cameras' then points')."""
    D = nat.sizes["dimension"]
    s = nat.scene_scale
    R = arrays["camera_rotation"].reshape(-1, 3, 3).copy()
    t = arrays["camera_translation"].reshape(-1, 3).copy()
    gp = arrays["group_params"].reshape(-1, 3).copy()
    X = arrays["point_xyz"].reshape(-1, 3).copy()
    if col >= D:
        q, k = divmod(col - D, 3)
        X[q, k] += s * h
    else:
        hit = False
        for c, (off, dim) in enumerate(zip(arrays["camera_offset"], arrays["camera_dim"])):
            if dim and off <= col < off + dim:
                local = col - off
                d = np.zeros(dim)
                d[local] = h
                R[c] = exp_so3(oracle, d[:3]) @ R[c]
                if c == nat.gauge_camera:  # b * retract_s2(direction, basis, d[3:5])
                    b = np.linalg.norm(t[c])
                    from openmvg_ba.problem import s2_tangent_basis
                    basis = s2_tangent_basis(t[c])
                    out = np.zeros(3)
                    direction = np.ascontiguousarray(t[c] / b)
                    oracle.bao_retract_s2(direction.ctypes.data, basis.ctypes.data,
                                          np.ascontiguousarray(d[3:5]).ctypes.data, out.ctypes.data)
                    t[c] = b * out
                else:
                    t[c] = t[c] + s * d[3:6]
                hit = True
        for g, (off, dim) in enumerate(zip(arrays["group_offset"], arrays["group_dim"])):
            if dim and off <= col < off + dim:
                local = col - off
                if local == 0:
                    gp[g, 0] = gp[g, 0] * np.exp(h)
                else:
                    gp[g, local] += h
                hit = True
        assert hit
    return {"camera_rotation": R.reshape(-1), "camera_translation": t.reshape(-1), "group_params": gp.reshape(-1),
            "point_xyz": X.reshape(-1)}


def ring_fixed():
    """The ring scene with every seventh point fixed and pose 3 fixed (a fixed observed pose is
    required with fixed world points: native/self_calibration.rs:47-85)."""
    sc = synthetic.ring_scene(seed=0xBA0011)
    prob = sc.problem
    fixed = np.zeros(len(prob.point_xyz), dtype=np.uint8)
    fixed[::7] = 1
    prob.point_fixed = fixed
    var = np.ones(len(prob.pose_center), dtype=np.uint8)
    var[3] = 0
    prob.pose_var_rotation, prob.pose_var_translation = var, var.copy()
    return prob


SCENES = {
    "ring": lambda: synthetic.ring_scene(n_points=200).problem,
    "fixture": lambda: synthetic.reference_fixture_scene().problem,
    "ring_fixed": ring_fixed,
}


@pytest.fixture(scope="module")
def pins(oracle):
    oracle.bao_exp_so3.argtypes = [C.c_void_p, C.c_void_p]
    oracle.bao_retract_s2.argtypes = [C.c_void_p] * 4
    oracle.bao_linearize.argtypes = [C.c_void_p] * 6
    return oracle


@pytest.mark.parametrize("scene", ["ring", "fixture"])
def test_jacobians_are_derivatives_along_the_retraction(pins, scene):
    oracle = pins
    nat, arrays = native_host(oracle, SCENES[scene]())
    r0, jx, jc = linearize(oracle, nat, arrays, A.LOSS_SQUARED, 0.0)
    J = dense_jacobian(nat, arrays, jx, jc)
    D = nat.sizes["dimension"]
    ncol = J.shape[1]
    rng = np.random.default_rng(7)
    cols = list(range(D)) + sorted(rng.choice(np.arange(D, ncol), size=min(60, ncol - D), replace=False).tolist())
    h = 1e-6
    worst = 0.0
    for col in cols:
        rp, _, _ = linearize(oracle, nat, arrays, A.LOSS_SQUARED, 0.0, perturbed_state(oracle, nat, arrays, col, h))
        rm, _, _ = linearize(oracle, nat, arrays, A.LOSS_SQUARED, 0.0, perturbed_state(oracle, nat, arrays, col, -h))
        fd = (rp - rm).reshape(-1) / (2 * h)
        an = J[:, col]
        scale = max(np.abs(an).max(), 1e-12)
        worst = max(worst, np.abs(fd - an).max() / scale)
    # the gauge camera's two S^2 columns are among the checked ones
    assert nat.gauge_camera >= 0 and arrays["camera_dim"][nat.gauge_camera] == 5
    print(f"{scene}: worst column error {worst:.2e} of {len(cols)} columns")
    assert worst < 1e-8, worst


@pytest.mark.parametrize("scene,loss", [("ring", A.LOSS_SQUARED), ("ring", A.LOSS_HUBER), ("fixture", A.LOSS_SQUARED),
                                        ("ring_fixed", A.LOSS_HUBER)])
@pytest.mark.parametrize("lam", [0.0, 1e-4, 0.1])
def test_assembly_equals_dense_normal_equations(pins, scene, loss, lam):
    oracle = pins
    scale = 1.0 if loss == A.LOSS_HUBER else 0.0
    nat, arrays = native_host(oracle, SCENES[scene]())
    r, jx, jc = linearize(oracle, nat, arrays, loss, scale)  # robust-weighted r~ and J~
    J = dense_jacobian(nat, arrays, jx, jc)
    S_ref, g_ref, gc_ref, diag_ref = reference_reduction(nat, arrays, J, r, lam)
    D = nat.sizes["dimension"]
    S = np.zeros((D, D))
    g, fg, dg = np.zeros(D), np.zeros(D), np.zeros(D)
    sc = A.SchurScalars()
    view = host_view(nat, arrays, loss, scale)
    oracle.bao_assemble(C.byref(view), lam, S.ctypes.data, g.ctypes.data, fg.ctypes.data, dg.ctypes.data,
                        None, None, None, C.byref(sc))
    assert sc.valid
    top = np.abs(S_ref).max()
    assert np.abs(S - S_ref).max() <= 1e-10 * top, np.abs(S - S_ref).max() / top
    assert np.abs(g - g_ref).max() <= 1e-10 * np.abs(g_ref).max()
    assert np.abs(fg - gc_ref).max() <= 1e-10 * np.abs(gc_ref).max()
    assert np.abs(dg - diag_ref).max() <= 1e-12 * np.abs(diag_ref).max()
    # the cost is half the squared weighted residual under squared loss
    if loss == A.LOSS_SQUARED:
        assert abs(sc.cost - 0.5 * np.sum(r * r)) <= 1e-12 * sc.cost


@pytest.mark.parametrize("factor", [1e-6, 1.0, 1e6])
def test_pcg_at_tight_tolerance_equals_direct_solve(pins, factor):
    """bao_pcg at rtol 1e-10 against numpy's solve of the dense-normal-equation S (scene scaled by
    `factor`: camera centres and points)."""
    oracle = pins
    prob = synthetic.ring_scene(n_points=200).problem
    prob.pose_center = prob.pose_center * factor
    prob.point_xyz = prob.point_xyz * factor
    nat, arrays = native_host(oracle, prob)
    lam = 1e-3
    r, jx, jc = linearize(oracle, nat, arrays, A.LOSS_SQUARED, 0.0)
    S_ref, g_ref, _, _ = reference_reduction(nat, arrays, dense_jacobian(nat, arrays, jx, jc), r, lam)
    want = np.linalg.solve(S_ref, -g_ref)
    D = nat.sizes["dimension"]
    step = np.zeros(D)
    it, rel, acc = C.c_uint32(), C.c_double(), C.c_int32()
    view = host_view(nat, arrays, A.LOSS_SQUARED, 0.0)
    oracle.bao_pcg(C.byref(view), lam, 1e-10, 10 * D, step.ctypes.data, C.byref(it), C.byref(rel), C.byref(acc))
    assert acc.value == 1 and rel.value <= 1e-10
    assert np.abs(step - want).max() <= 1e-8 * np.abs(want).max()
