"""COLMAP text models (openmvg_ba.colmap) against the conventions of openmvg-io's exporter
(crates/openmvg-io/src/colmap.rs:53-124, 300-380): the camera-model table, the 0.5-pixel offset,
world-to-camera quaternions with t = -R C, ID numbering, and write/read round trips whose solves
equal the original's (no GPU needed)."""

from __future__ import annotations

import numpy as np
import pytest

from openmvg_ba import SolveConfig, _abi as A, colmap, synthetic
from openmvg_ba.problem import BundleProblem, Intrinsics, solve_with


def _sorted_obs(p):
    order = np.lexsort((p.obs_pose, p.obs_point))
    return p.obs_point[order], p.obs_pose[order], p.obs_pixel[order], p.obs_intrinsic[order]


@pytest.mark.parametrize("name", ["ring", "brown"])
def test_write_read_round_trip(tmp_path, name):
    sc = synthetic.ring_scene() if name == "ring" else synthetic.calibrated_brown_scene()
    p = sc.problem
    out = colmap.write_colmap(p, tmp_path)
    assert out == tmp_path / "sparse" / "0"
    q, names = colmap.read_colmap(tmp_path, free_intrinsics=(name == "ring"))
    assert names == [f"image{i}.jpg" for i in range(p.n_poses)]
    assert q.n_poses == p.n_poses and q.n_points == p.n_points and q.n_observations == p.n_observations
    assert np.max(np.abs(q.pose_rotation - p.pose_rotation)) < 1e-15
    assert np.max(np.abs(q.pose_center - p.pose_center)) < 1e-12 * max(1.0, np.abs(p.pose_center).max())
    assert np.array_equal(q.point_xyz, p.point_xyz)  # shortest round-trip text
    a, b = _sorted_obs(q), _sorted_obs(p)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])
    assert np.max(np.abs(a[2] - b[2])) < 1e-12 * np.abs(b[2]).max()  # +0.5 then -0.5
    for x, y in zip(q.intrinsics, p.intrinsics):
        assert (x.focal, x.cx, x.cy, x.width, x.height) == (y.focal, y.cx, y.cy, y.width, y.height)
        if y.model != A.CAMERA_PINHOLE:
            assert (x.k1, x.k2) == (y.k1, y.k2)


def test_files_follow_the_exporter_conventions(tmp_path):
    sc = synthetic.ring_scene()
    p = sc.problem
    out = colmap.write_colmap(p, tmp_path)
    cams = [ln for ln in (out / "cameras.txt").read_text().splitlines() if not ln.startswith("#")]
    assert cams[0].split()[:4] == ["1", "SIMPLE_PINHOLE", "1024", "768"]
    assert [float(v) for v in cams[0].split()[4:]] == [1000.0, 512.5, 384.5]
    imgs = [ln for ln in (out / "images.txt").read_text().splitlines() if not ln.startswith("#")]
    head = imgs[0].split()
    qw, qx, qy, qz = (float(v) for v in head[1:5])
    R = colmap.quaternion_to_rotation([qw, qx, qy, qz])
    assert np.max(np.abs(R - p.pose_rotation[0])) < 1e-15
    t = np.array([float(v) for v in head[5:8]])
    assert np.max(np.abs(t + p.pose_rotation[0] @ p.pose_center[0])) < 1e-12
    assert head[8] == "1" and head[9] == "image0.jpg"
    # the first listed 2-D point of image 1 is its first observation in track order, shifted by 0.5
    first = imgs[1].split()[:3]
    obs = [o for o in np.argsort(p.obs_point, kind="stable") if p.obs_pose[o] == 0][0]
    assert float(first[0]) == p.obs_pixel[obs][0] + 0.5 and float(first[1]) == p.obs_pixel[obs][1] + 0.5
    assert int(first[2]) == p.obs_point[obs] + 1
    pts = [ln for ln in (out / "points3D.txt").read_text().splitlines() if not ln.startswith("#")]
    f = pts[0].split()
    assert f[0] == "1" and f[4:7] == ["128", "128", "128"]
    # ERROR: mean reprojection error of the point's observations, in pixels
    k = np.flatnonzero(p.obs_point == 0)
    err = []
    for o in k:
        pc = p.pose_rotation[p.obs_pose[o]] @ (p.point_xyz[0] - p.pose_center[p.obs_pose[o]])
        err.append(np.hypot(*(1000.0 * pc[:2] / pc[2] + [512.0, 384.0] - p.obs_pixel[o])))
    assert abs(float(f[7]) - np.mean(err)) < 1e-9


def test_camera_model_table():
    base = dict(focal=800.0, cx=319.5, cy=239.5, width=640, height=480)
    cases = [
        (Intrinsics(**base), "SIMPLE_PINHOLE", [800.0, 320.0, 240.0]),
        (Intrinsics(**base, focal_y=810.0), "PINHOLE", [800.0, 810.0, 320.0, 240.0]),
        (Intrinsics(**base, model=A.CAMERA_BROWN_K1K2), "SIMPLE_PINHOLE", [800.0, 320.0, 240.0]),
        (Intrinsics(**base, model=A.CAMERA_BROWN_K1, k1=-0.1), "SIMPLE_RADIAL", [800.0, 320.0, 240.0, -0.1]),
        (Intrinsics(**base, model=A.CAMERA_BROWN_K1K2, k1=-0.1, k2=0.02), "RADIAL", [800.0, 320.0, 240.0, -0.1, 0.02]),
        (Intrinsics(**base, model=A.CAMERA_BROWN_K1K2, k1=-0.1, p1=1e-3), "OPENCV",
         [800.0, 800.0, 320.0, 240.0, -0.1, 0.0, 1e-3, 0.0]),
        (Intrinsics(**base, model=A.CAMERA_BROWN_K1K2K3, k1=-0.1, k3=1e-3), "FULL_OPENCV",
         [800.0, 800.0, 320.0, 240.0, -0.1, 0.0, 0.0, 0.0, 1e-3, 0.0, 0.0, 0.0]),
    ]
    for intr, model, params in cases:
        assert colmap.colmap_camera(intr) == (model, params)
        back = colmap.intrinsic_from_colmap(model, 640, 480, params, free=False)
        assert colmap.colmap_camera(back) == (model, params)
    with pytest.raises(ValueError):
        colmap.colmap_camera(Intrinsics(**base, model=A.CAMERA_SPHERICAL))
    with pytest.raises(ValueError):
        colmap.intrinsic_from_colmap("FULL_OPENCV", 640, 480, [1, 1, 1, 1, 0, 0, 0, 0, 0, 0.1, 0, 0], False)
    with pytest.raises(ValueError):
        colmap.intrinsic_from_colmap("FOV", 640, 480, [1, 1, 1, 1, 0], False)


def test_hand_written_model(tmp_path):
    """A model in COLMAP's documented text layout, with a 2-D point without a 3-D point (-1)."""
    d = tmp_path / "model"
    d.mkdir()
    (d / "cameras.txt").write_text("# Camera list\n1 SIMPLE_RADIAL 3072 2304 2559.81 1536.5 1152.5 -0.0204\n")
    (d / "images.txt").write_text(
        "# Image list\n"
        "1 1 0 0 0 0 0 0 1 a.jpg\n"
        "100.5 200.5 -1 1600.5 1200.5 7\n"
        "2 0.9999875 0 0.005 0 -1 0 0 1 b.jpg\n"
        "1612.5 1188.5 7\n")
    (d / "points3D.txt").write_text("# 3D points\n7 0.1 0.05 4.0 10 20 30 0.5 1 1 2 0\n")
    p, names = colmap.read_colmap(d)
    assert names == ["a.jpg", "b.jpg"]
    assert p.n_poses == 2 and p.n_points == 1 and p.n_observations == 2
    assert np.array_equal(p.obs_pose, [0, 1]) and np.array_equal(p.obs_point, [0, 0])
    assert np.array_equal(p.obs_pixel, [[1600.0, 1200.0], [1612.0, 1188.0]])
    intr = p.intrinsics[0]
    assert (intr.model, intr.focal, intr.cx, intr.cy, intr.k1) == (A.CAMERA_BROWN_K1, 2559.81, 1536.0, 1152.0, -0.0204)
    # camera centre of image 2: C = -R^T t with t = (-1, 0, 0)
    assert np.allclose(p.pose_center[1], p.pose_rotation[1].T @ np.array([1.0, 0.0, 0.0]))


def test_read_back_solve_equals_original(tmp_path, oracle):
    sc = synthetic.calibrated_ring_scene()
    p = sc.problem
    colmap.write_colmap(p, tmp_path)
    q, _ = colmap.read_colmap(tmp_path)
    cfg = SolveConfig(max_iterations=20, linear_solver=A.SOLVER_DIRECT)
    a = solve_with(oracle, "bao_solve", p, cfg)
    b = solve_with(oracle, "bao_solve", q, cfg)
    assert a.report.iterations == b.report.iterations
    assert abs(a.report.final_cost - b.report.final_cost) <= 1e-9 * a.report.final_cost
    assert np.max(np.abs(a.point_xyz - b.point_xyz)) < 1e-8


def test_writer_refuses_a_pose_with_two_cameras(tmp_path):
    p = synthetic.ring_scene().problem
    oi = p.obs_intrinsic.copy()
    oi[np.flatnonzero(p.obs_pose == 0)[0]] = 1
    bad = BundleProblem(p.pose_rotation, p.pose_center, p.intrinsics, p.point_xyz, p.obs_pose, oi, p.obs_point,
                        p.obs_pixel)
    with pytest.raises(ValueError):
        colmap.write_colmap(bad, tmp_path)


def test_quaternion_conversion_on_every_branch():
    rng = np.random.default_rng(7)
    for w in list(rng.normal(size=(50, 3)) * 2.0) + [np.array([np.pi - 1e-9, 0, 0]), np.array([0, np.pi, 0]),
                                                       np.array([0, 0, np.pi - 1e-3]), np.zeros(3)]:
        R = synthetic.exp_so3(w)
        q = colmap.rotation_to_quaternion(R)
        assert q[0] >= 0.0 and abs(np.linalg.norm(q) - 1.0) < 1e-15
        assert np.max(np.abs(colmap.quaternion_to_rotation(q) - R)) < 5e-15
