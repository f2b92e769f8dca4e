"""Pins of the CPU oracle against the reference's own known-answer tests and against
mathematics that does not go through the oracle's formulas.

Each test cites the reference test it restates (paths under crates/openmvg-ba/).
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from openmvg_ba import _abi as A
from oracle_lib import load_oracle

D3 = C.c_double * 3
D2 = C.c_double * 2
D6 = C.c_double * 6
D9 = C.c_double * 9


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def lin(orc, cam, obs, pt):
    res, J, Ji = D2(), D6(), D6()
    ok = orc.bao_radial_linearization(D3(*cam), D2(*obs), D3(*pt), res, J, Ji)
    return bool(ok), np.array(res[:]), np.array(J[:]).reshape(2, 3), np.array(Ji[:]).reshape(2, 3)


def proj(orc, cam, obs, pt):
    res, J = D2(), D6()
    ok = orc.bao_radial_projection(D3(*cam), D2(*obs), D3(*pt), res, J)
    return bool(ok), np.array(res[:]), np.array(J[:]).reshape(2, 3)


CAM = (730.0, 0.08, -0.025)
OBS = (17.0, -23.0)
PT = (0.71, -0.43, 2.4)


def close(a, e, tol):
    return abs(a - e) <= tol * (1.0 + abs(e))


# ---- radial model: src/kernel/multi_view/self_calibration/model.rs tests --------------------
def test_point_jacobian_matches_central_differences(orc):  # model.rs:396-424
    ok, _, J, _ = lin(orc, CAM, OBS, PT)
    assert ok
    eps = 1e-6
    for c in range(3):
        p, m = list(PT), list(PT)
        p[c] += eps
        m[c] -= eps
        rp = proj(orc, CAM, OBS, p)[1]
        rm = proj(orc, CAM, OBS, m)[1]
        for r in range(2):
            assert close(J[r, c], (rp[r] - rm[r]) / (2 * eps), 2e-8)


def test_intrinsic_jacobian_matches_log_focal_retraction(orc):  # model.rs:426-467
    ok, _, _, Ji = lin(orc, CAM, OBS, PT)
    eps = 1e-6
    for k in range(3):
        cp, cm = list(CAM), list(CAM)
        if k == 0:
            cp[0] *= math.exp(eps)
            cm[0] *= math.exp(-eps)
        else:
            cp[k] += eps
            cm[k] -= eps
        # prediction = residual + observation
        rp = proj(orc, cp, (0.0, 0.0), PT)[1]
        rm = proj(orc, cm, (0.0, 0.0), PT)[1]
        for r in range(2):
            assert close(Ji[r, k], (rp[r] - rm[r]) / (2 * eps), 2e-8)


class Dual:
    """Forward-mode dual number with a gradient vector (a restatement of numerics/jet.rs's idea)."""

    def __init__(self, v, d):
        self.v = float(v)
        self.d = np.asarray(d, dtype=np.float64)

    @staticmethod
    def var(v, i, n):
        d = np.zeros(n)
        d[i] = 1.0
        return Dual(v, d)

    def _w(self, o):
        return o if isinstance(o, Dual) else Dual(o, np.zeros_like(self.d))

    def __add__(self, o):
        o = self._w(o)
        return Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = self._w(o)
        return Dual(self.v - o.v, self.d - o.d)

    def __mul__(self, o):
        o = self._w(o)
        return Dual(self.v * o.v, self.d * o.v + o.d * self.v)

    __rmul__ = __mul__

    def recip(self):
        return Dual(1.0 / self.v, -self.d / (self.v * self.v))

    def exp(self):
        e = math.exp(self.v)
        return Dual(e, self.d * e)


def test_jacobians_match_dual_number_oracle(orc):  # model.rs:598-634
    n = 6
    X = [Dual.var(PT[i], i, n) for i in range(3)]
    logf = Dual.var(math.log(CAM[0]), 3, n)
    k1 = Dual.var(CAM[1], 4, n)
    k2 = Dual.var(CAM[2], 5, n)
    iz = X[2].recip()
    nx, ny = X[0] * iz, X[1] * iz
    r2 = nx * nx + ny * ny
    d = 1.0 + k1 * r2 + k2 * r2 * r2
    f = logf.exp()
    res = [f * d * nx - OBS[0], f * d * ny - OBS[1]]
    ok, r0, J, Ji = lin(orc, CAM, OBS, PT)
    for r in range(2):
        assert close(res[r].v, r0[r], 5e-12)
        for c in range(3):
            assert close(res[r].d[c], J[r, c], 5e-12)
            assert close(res[r].d[3 + c], Ji[r, c], 5e-12)


def test_projection_domain_is_the_increasing_branch(orc):  # model.rs:364-394
    cam = (500.0, -0.1, 0.0)
    assert not proj(orc, cam, (0, 0), (math.sqrt(10.0), 0.0, 1.0))[0]
    assert not lin(orc, cam, (0, 0), (math.sqrt(10.0), 0.0, 1.0))[0]
    pt = (0.5, 0.2, 1.0)
    pred = proj(orc, cam, (0, 0), pt)[1]
    b = D3()
    assert orc.bao_radial_bearing(D3(*cam), D2(*pred), b)
    nrm = math.sqrt(sum(v * v for v in pt))
    for i in range(3):
        assert abs(b[i] - pt[i] / nrm) < 1e-12
    extreme = (1.0, -1e308, 0.0)
    bound = orc.bao_first_positive_radial_turn(D3(*extreme))
    assert math.isfinite(bound) and bound > 0
    assert proj(orc, extreme, (0, 0), (math.sqrt(bound * 0.25), 0, 1))[0]
    assert not proj(orc, extreme, (0, 0), (math.sqrt(bound * 2.0), 0, 1))[0]
    ext2 = (1.0, 0.0, -1e308)
    b2 = orc.bao_first_positive_radial_turn(D3(*ext2))
    assert math.isfinite(b2) and b2 > 0
    assert proj(orc, ext2, (0, 0), (math.sqrt(b2 * 0.25), 0, 1))[0]


def test_first_turn_is_the_root_of_the_radial_slope(orc):
    """1 + 3 k1 r^2 + 5 k2 r^4 vanishes at the returned r^2 (quadratic formula, closed form)."""
    for k1, k2 in [(-0.1, 0.0), (0.08, -0.025), (-0.3, 0.02), (-1.0, 0.0), (0.1, -0.05)]:
        r2 = orc.bao_first_positive_radial_turn(D3(1.0, k1, k2))
        roots = np.roots([5 * k2, 3 * k1, 1.0]) if k2 != 0 else np.array([-1.0 / (3 * k1)])
        pos = sorted(r.real for r in np.atleast_1d(roots) if abs(r.imag) < 1e-12 and r.real > 0)
        assert pos and abs(r2 - pos[0]) <= 1e-12 * pos[0]
    assert math.isnan(orc.bao_first_positive_radial_turn(D3(1.0, 0.1, 0.05)))


def test_bearing_rejects_folded_inverse_and_round_trips(orc):  # model.rs:469-546
    b = D3()
    assert not orc.bao_radial_bearing(D3(1.0, -1.0, 0.0), D2(0.5, 0.0), b)
    for cam in [(500.0, 0.0, 0.0), (700.0, 0.12, 0.02)]:
        for radius in [0.0, 1e-8, 0.1, 0.5, 1.0, 2.0, 4.0]:
            for az in [0.0, math.pi / 6, math.pi / 2, 2.3]:
                for depth in [0.2, 1.0, 17.0]:
                    pt = (radius * math.cos(az) * depth, radius * math.sin(az) * depth, depth)
                    ok, m, _ = proj(orc, cam, (0, 0), pt)
                    assert ok
                    assert orc.bao_radial_bearing(D3(*cam), D2(*m), b)
                    nrm = math.sqrt(sum(v * v for v in pt))
                    for i in range(3):
                        assert close(b[i], pt[i] / nrm, 2e-12)


def test_derivative_overflow_fails_closed(orc):  # model.rs:636-647
    assert not proj(orc, (1.7976931348623157e308 / 2.0, 100.0, 0.0), (0, 0), (0.1, 0.0, 1.0))[0]
    assert not lin(orc, (1.7976931348623157e308 / 8.0, 0.1, 0.0), (0, 0), (2.0, 0.0, 1.0))[0]


# ---- robust losses: src/numerics/robust.rs:50-131 ------------------------------------------
def test_loss_weights_are_cost_derivatives(orc):
    cost = orc.bao_loss_cost
    w = orc.bao_loss_sqrt_weight
    assert w(A.LOSS_CAUCHY, 0.01, 0.5) < 0.1 * w(A.LOSS_CAUCHY, 0.01, 0.001)
    for e in range(-4, 5):
        r = 0.01 * 10.0 ** e
        eps = 1e-6 * r
        num = (cost(A.LOSS_CAUCHY, 0.01, r + eps) - cost(A.LOSS_CAUCHY, 0.01, r - eps)) / (2 * eps)
        ana = w(A.LOSS_CAUCHY, 0.01, r) ** 2 * r
        rounding = 4 * np.finfo(float).eps * cost(A.LOSS_CAUCHY, 0.01, r) / eps
        assert abs(num - ana) <= 1e-6 * abs(ana) + rounding
    assert w(A.LOSS_CAUCHY, 0.01, 1e18) ** 2 < 1e-14
    for r in [0.004, 0.04, -0.04]:
        num = (cost(A.LOSS_HUBER, 0.01, r + 1e-8) - cost(A.LOSS_HUBER, 0.01, r - 1e-8)) / 2e-8
        assert abs(num - w(A.LOSS_HUBER, 0.01, r) ** 2 * r) < 1e-10
    for kind in (A.LOSS_HUBER, A.LOSS_CAUCHY):
        assert cost(kind, 0.01, 0.0) == 0.0
        prev = 0.0
        for e in range(-8, 9):
            r = 0.01 * 10.0 ** e
            assert cost(kind, 0.01, r) == cost(kind, 0.01, -r)
            assert cost(kind, 0.01, r) >= prev
            prev = cost(kind, 0.01, r)
    for r in [1e-5, 1e-9, 1e-150]:
        assert abs(cost(A.LOSS_CAUCHY, 1.0, r) - 0.5 * r * r) <= 1e-9 * 0.5 * r * r
    d, e = 0.01, 1e-12
    assert abs(cost(A.LOSS_HUBER, d, d + e) - cost(A.LOSS_HUBER, d, d - e) - 2 * e * d) <= 1e-20
    assert w(A.LOSS_HUBER, d, d) == 1.0
    assert abs(w(A.LOSS_HUBER, d, d + e) - 1.0) <= 1e-9
    assert cost(A.LOSS_SQUARED, 0.0, 3.0) == 4.5 and w(A.LOSS_SQUARED, 0.0, 3.0) == 1.0


# ---- point block: src/kernel/point_block.rs:214-313 ----------------------------------------
def factor(orc, m, damping=0.0):
    inv = D6()
    ok = orc.bao_point_block_factor(D9(*np.asarray(m, dtype=float).ravel()), damping, inv)
    if not ok:
        return None
    i00, i10, i11, i20, i21, i22 = inv[:]
    return np.array([[i00, i10, i20], [i10, i11, i21], [i20, i21, i22]])


def test_point_block_factor_known_answers(orc):
    for m in [np.diag([1.0, 1e-8, 1e-8]), np.diag([1e12, 1.0, 1.0])]:
        inv = factor(orc, m)
        x = inv @ np.array([0.4, -0.3, 0.2])
        assert np.abs(m @ x - np.array([0.4, -0.3, 0.2])).max() < 1e-11
    q = np.array([[1.0, 2.0, 2.0], [2.0, 1.0, -2.0], [2.0, -2.0, 1.0]]) / 3.0
    d = np.array([1e6, 1.0, 1e-3])
    H = q @ np.diag(d) @ q.T
    Hinv = q @ np.diag(1.0 / d) @ q.T
    assert np.abs(factor(orc, H) - Hinv).max() <= 1e-6 * 1e3
    assert factor(orc, np.diag([1.0, 1.0, 0.0])) is None
    m = np.array([[4.0, 0.2, -0.1], [0.2, 2.0, 0.1], [-0.1, 0.1, 1.0]])
    assert np.abs(m @ factor(orc, m) - np.eye(3)).max() < 1e-12
    # damped block: V + lambda max(diag, floor) on the diagonal
    lam = 0.3
    mm = np.array([[4.0, 0.2, 0.0], [0.2, 2.0, 0.1], [0.0, 0.1, 1.0]])
    damped = mm + lam * np.diag(np.diag(mm))
    assert np.abs(factor(orc, mm, lam) - np.linalg.inv(damped)).max() < 1e-14
    rc = orc.bao_point_block_rcond(D9(*mm.ravel()))
    expect = 1.0 / (np.abs(mm).sum(1).max() * np.abs(np.linalg.inv(mm)).sum(1).max())
    assert abs(rc - expect) <= 1e-12 * expect


# ---- track geometry: src/kernel/point_geometry.rs:170-227 ----------------------------------
def geo(orc, offsets, scene=0.0):
    arr = np.asarray(offsets, dtype=np.float64).ravel()
    return orc.bao_track_geometry(len(offsets), arr.ctypes.data, scene)


def test_track_geometry_cases(orc):
    assert geo(orc, [[0, 0, 5], [1, 0, 5], [-1, 0, 6]]) & 1 == 0
    assert geo(orc, [[0, 0, 5], [1e-7, 0, 1e-7]]) & 1
    assert geo(orc, [[0, 0, 5], [0, 0, 0]]) & 1
    assert geo(orc, [[0, 0, 1e5], [1, 0, 1e5]]) & 1
    assert geo(orc, [[0, 0, 5], [0, 0, -5]]) & 1 == 0
    collapsed = [[0, 0, 1e-6], [1e-6, 0, 1e-6]]
    assert geo(orc, collapsed) & 1 == 0
    assert geo(orc, collapsed, 10.0) == 3
    assert geo(orc, [[0, 0, 1e5], [1, 0, 1e5]], 10.0) == 1
    assert geo(orc, [[0, 0, 5]]) == 0


# ---- LM controller, run state, stages ------------------------------------------------------
def accept(orc, state, actual, predicted):
    v, nu, g = C.c_double(state[0]), C.c_double(state[1]), C.c_double()
    ok = orc.bao_nielsen_accept(C.byref(v), C.byref(nu), actual, predicted, C.byref(g))
    state[0], state[1] = v.value, nu.value
    return bool(ok), g.value


def reject(orc, state):
    v, nu = C.c_double(state[0]), C.c_double(state[1])
    orc.bao_nielsen_reject(C.byref(v), C.byref(nu))
    state[0], state[1] = v.value, nu.value


def test_nielsen_tables(orc):  # controller.rs:97-114 and run_state.rs:214-243
    s = [1e-4, 2.0]
    ok, g = accept(orc, s, 2.0, 2.0)
    assert ok and g == 1.0 and abs(s[0] - 1e-4 / 3.0) < 1e-15
    reject(orc, s)
    assert abs(s[0] - 2e-4 / 3.0) < 1e-15
    reject(orc, s)
    assert abs(s[0] - 8e-4 / 3.0) < 1e-15
    ok, g = accept(orc, s, -1.0, 1.0)
    assert not ok and g == -1.0
    ok, g = accept(orc, s, 1.0, -1.0)
    assert not ok and math.isnan(g)
    s = [1e-4, 2.0]
    for expect in (2e-4, 8e-4):
        reject(orc, s)
        assert abs(s[0] - expect) <= 8 * np.finfo(float).eps * max(expect, 1.0)
    accept(orc, s, -1.0, 1.0)
    assert abs(s[0] - 6.4e-3) <= 8 * np.finfo(float).eps
    accept(orc, s, 1.0, -1.0)
    assert abs(s[0] - 0.1024) <= 8 * np.finfo(float).eps
    s = [1e12, 1e12]
    reject(orc, s)
    assert s == [1e12, 1e12]
    s = [1e-12, 2.0]
    accept(orc, s, 1.0, 1.0)
    assert s[0] == 1e-12


def test_stage_schedule(orc):  # continuation.rs:81-111
    lo = (C.c_int32 * 3)()
    it = (C.c_uint32 * 3)()
    for n in range(20):
        for loss in (A.LOSS_SQUARED, A.LOSS_HUBER, A.LOSS_CAUCHY):
            orc.bao_stages(loss, n, lo, it)
            assert sum(it) == n
            if n > 0:
                assert any(it[k] > 0 and lo[k] == loss for k in range(3))
            if loss == A.LOSS_CAUCHY:
                assert it[0] == min(n // 4, 5) and lo[0] == A.LOSS_SQUARED


# ---- manifold maps --------------------------------------------------------------------------
def test_exp_so3_is_rodrigues(orc):
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(3)
    for w in list(rng.normal(size=(8, 3))) + [np.array([1e-12, -2e-12, 0.0]), np.zeros(3)]:
        R = D9()
        orc.bao_exp_so3(D3(*w), R)
        expect = Rotation.from_rotvec(w).as_matrix()
        assert np.abs(np.array(R[:]).reshape(3, 3) - expect).max() < 1e-15 * 8


def test_s2_retraction_stays_on_the_sphere(orc):
    t = D3(0.3, -2.0, 0.5)
    basis = D6()
    orc.bao_s2_tangent_basis(t, basis)
    b = np.array(basis[:]).reshape(2, 3)
    tv = np.array(t[:]) / np.linalg.norm(t[:])
    assert np.abs(b @ tv).max() < 1e-15 and np.abs(b @ b.T - np.eye(2)).max() < 1e-15
    out = D3()
    for d in [(0.0, 0.0), (1e-12, 0.0), (0.3, -0.2)]:
        orc.bao_retract_s2(t, basis, D2(*d), out)
        o = np.array(out[:])
        assert abs(np.linalg.norm(o) - 1.0) < 1e-15
        th = math.hypot(*d)
        assert abs(float(o @ tv) - math.cos(th)) < 1e-15
