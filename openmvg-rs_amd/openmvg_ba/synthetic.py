"""Seeded synthetic bundle-adjustment scenes.

The five workloads restate the configurations this path is measured on:

* ``ring_scene`` — 10 pinhole cameras on a radius-10 ring around 1,000 points (~6k observations),
  focal free, squared loss.
* ``grid_scene`` — cameras on a noisy planar grid looking along +z at a point field at depth
  8–15, radial (f, k1, k2) free; 200 cameras / 50k points (~300k observations) for the medium
  problem, 2,000 cameras / 1M points (~6M observations, geometric track lengths, 2 % outliers,
  Huber) for the large one.
* ``window_batch`` — many small local-BA windows.
* ``bal_scene`` — a BAL-style scene (angle-axis, t, f, k1, k2 per camera) for the cost/gradient
  (differentiation) form.

Randomness is numpy's PCG64 (``numpy.random.default_rng(seed)``); the same seed gives the same
scene on every machine with the same numpy.  Camera convention: world-to-camera rotation R,
centre C, camera point R (X - C), +z forward, +y down; pixels are principal point plus
f (1 + k1 r^2 + k2 r^4) (x/z, y/z), the radial model of
crates/openmvg-ba/src/kernel/multi_view/self_calibration/model.rs:26-75.
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from . import _abi as A
from .problem import BundleProblem, Intrinsics

WIDTH, HEIGHT = 1024, 768
CX, CY = 512.0, 384.0


def exp_so3(w: np.ndarray) -> np.ndarray:
    """Rodrigues exponential of (..., 3) rotation vectors, (..., 3, 3)."""
    w = np.asarray(w, dtype=np.float64)
    th = np.linalg.norm(w, axis=-1, keepdims=True)
    safe = np.where(th < 1e-12, 1.0, th)
    k = w / safe
    K = np.zeros(w.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2] = -k[..., 2], k[..., 1]
    K[..., 1, 0], K[..., 1, 2] = k[..., 2], -k[..., 0]
    K[..., 2, 0], K[..., 2, 1] = -k[..., 1], k[..., 0]
    s = np.sin(th)[..., None]
    c = np.cos(th)[..., None]
    eye = np.broadcast_to(np.eye(3), K.shape)
    return eye + s * K + (1.0 - c) * (K @ K)


def log_so3(R: np.ndarray) -> np.ndarray:
    """Principal rotation vector of (..., 3, 3) rotations away from pi."""
    tr = np.trace(R, axis1=-2, axis2=-1)
    c = np.clip((tr - 1.0) * 0.5, -1.0, 1.0)
    th = np.arccos(c)
    v = np.stack([R[..., 2, 1] - R[..., 1, 2], R[..., 0, 2] - R[..., 2, 0], R[..., 1, 0] - R[..., 0, 1]], -1)
    s = np.sin(th)
    scale = np.where(th < 1e-8, 0.5, th / (2.0 * np.where(s == 0, 1.0, s)))
    return v * scale[..., None]


def look_at(center: np.ndarray, target: np.ndarray, up=(0.0, -1.0, 0.0)) -> np.ndarray:
    """World-to-camera rotation whose +z points from `center` to `target` and +y is `up`."""
    z = target - center
    z = z / np.linalg.norm(z)
    y = np.asarray(up, dtype=np.float64)
    y = y - np.dot(y, z) * z
    y = y / np.linalg.norm(y)
    x = np.cross(y, z)
    return np.stack([x, y, z])


def radial_project(R, C, f, k1, k2, X):
    """Pixels (principal point included) and depths of points X (..., 3) in cameras (R, C)."""
    pc = np.einsum("...ij,...j->...i", R, X - C)
    z = pc[..., 2]
    safe = np.where(np.abs(z) < 1e-12, 1e-12, z)
    n = pc[..., :2] / safe[..., None]
    r2 = np.sum(n * n, axis=-1)
    d = 1.0 + k1 * r2 + k2 * r2 * r2
    uv = (f * d)[..., None] * n
    return uv + np.array([CX, CY]), z


@dataclass
class Scene:
    problem: BundleProblem
    truth_rotation: np.ndarray
    truth_center: np.ndarray
    truth_points: np.ndarray
    truth_focal: np.ndarray
    truth_k1: np.ndarray
    truth_k2: np.ndarray
    name: str = ""


def _observe(R, C, f, k1, k2, X, rng, cap_of_point, noise, outlier_fraction=0.0,
             candidates=None, chunk=65536):
    """Observations of points X by the cameras that see them inside the image.

    `candidates(i0, i1)` returns, for points i0..i1, a (n, m) array of candidate camera
    indices (or all cameras).  Each point keeps a random subset of its visible cameras of size
    `cap_of_point[i]` (at least 2); points with fewer than 2 visible cameras are dropped.
    Returns (obs_pose, obs_point, obs_pixel, kept_point_mask).
    """
    n_pts = len(X)
    poses, points, pixels = [], [], []
    kept = np.zeros(n_pts, dtype=bool)
    for i0 in range(0, n_pts, chunk):
        i1 = min(n_pts, i0 + chunk)
        cand = candidates(i0, i1) if candidates is not None else np.broadcast_to(
            np.arange(len(R)), (i1 - i0, len(R)))
        Xc = X[i0:i1, None, :]
        uv, z = radial_project(R[cand], C[cand], f[cand], k1[cand], k2[cand], Xc)
        vis = (z > 0.1) & (uv[..., 0] >= 0) & (uv[..., 0] < WIDTH) & (uv[..., 1] >= 0) & (uv[..., 1] < HEIGHT)
        vis &= cand >= 0
        # random priority among visible cameras, keep the first `cap` of them
        pri = rng.random(vis.shape)
        pri[~vis] = 2.0
        order = np.argsort(pri, axis=1, kind="stable")
        nvis = vis.sum(axis=1)
        cap = np.minimum(cap_of_point[i0:i1], nvis)
        ok = cap >= 2
        kept[i0:i1] = ok
        rank = np.empty_like(order)
        rows = np.arange(order.shape[0])[:, None]
        rank[rows, order] = np.arange(order.shape[1])[None, :]
        take = vis & (rank < cap[:, None]) & ok[:, None]
        pi, ci = np.nonzero(take)  # row-major: grouped by point, cameras in candidate order
        cams = cand[pi, ci]
        pix = uv[pi, ci]
        pixels.append(pix)
        poses.append(cams)
        points.append(pi + i0)
    obs_pose = np.concatenate(poses).astype(np.uint32)
    obs_point = np.concatenate(points).astype(np.int64)
    pix = np.concatenate(pixels)
    pix = pix + rng.normal(0.0, noise, pix.shape)
    if outlier_fraction > 0:
        out = rng.random(len(pix)) < outlier_fraction
        pix[out] = np.stack([rng.uniform(0, WIDTH, out.sum()), rng.uniform(0, HEIGHT, out.sum())], -1)
    return obs_pose, obs_point, pix, kept


def _observe_grid(R, C, f, k1, k2, X, rng, cap, noise, outlier_fraction, nx, ny, chunk=131072):
    """Tracks for cameras on an nx-by-ny unit grid looking along +z: each point draws up to
    cap[i] distinct cameras uniformly from the grid window that sees it (half-angles 0.46 and
    0.34 of the depth, inside the 0.512/0.384 field of view), keeps the ones whose exact
    projection falls in the image, and is dropped below 2 views."""
    n_pts = len(X)
    kmax = int(cap.max())
    poses, points, pixels = [], [], []
    kept = np.zeros(n_pts, dtype=bool)
    for i0 in range(0, n_pts, chunk):
        i1 = min(n_pts, i0 + chunk)
        Xc = X[i0:i1]
        gx = Xc[:, 0] + (nx - 1) / 2.0
        gy = Xc[:, 1] + (ny - 1) / 2.0
        ix0 = np.clip(np.ceil(gx - 0.46 * Xc[:, 2]), 0, nx - 1).astype(np.int64)
        ix1 = np.clip(np.floor(gx + 0.46 * Xc[:, 2]), 0, nx - 1).astype(np.int64)
        iy0 = np.clip(np.ceil(gy - 0.34 * Xc[:, 2]), 0, ny - 1).astype(np.int64)
        iy1 = np.clip(np.floor(gy + 0.34 * Xc[:, 2]), 0, ny - 1).astype(np.int64)
        wx = np.maximum(ix1 - ix0 + 1, 0)
        wy = np.maximum(iy1 - iy0 + 1, 0)
        count = wx * wy
        wmax = int(count.max())
        keys = rng.random((i1 - i0, wmax), dtype=np.float32)
        keys[np.arange(wmax)[None, :] >= count[:, None]] = 2.0
        kk = min(kmax, wmax)
        part = np.argpartition(keys, kk - 1, axis=1)[:, :kk] if kk < wmax else np.broadcast_to(
            np.arange(wmax), (i1 - i0, wmax)).copy()
        order = np.argsort(np.take_along_axis(keys, part, axis=1), axis=1, kind="stable")
        sel = np.take_along_axis(part, order, axis=1)
        take = (np.arange(sel.shape[1])[None, :] < np.minimum(cap[i0:i1], count)[:, None])
        cam = (iy0[:, None] + sel // np.maximum(wx, 1)[:, None]) * nx + ix0[:, None] + sel % np.maximum(wx, 1)[:, None]
        cam = np.where(take, cam, 0)
        uv, z = radial_project(R[cam], C[cam], f[cam], k1[cam], k2[cam], Xc[:, None, :])
        vis = take & (z > 0.1) & (uv[..., 0] >= 0) & (uv[..., 0] < WIDTH) & (uv[..., 1] >= 0) & (uv[..., 1] < HEIGHT)
        ok = vis.sum(axis=1) >= 2
        kept[i0:i1] = ok
        vis &= ok[:, None]
        pi, ci = np.nonzero(vis)
        poses.append(cam[pi, ci])
        points.append(pi + i0)
        pixels.append(uv[pi, ci])
    obs_pose = np.concatenate(poses).astype(np.uint32)
    obs_point = np.concatenate(points).astype(np.int64)
    pix = np.concatenate(pixels)
    pix = pix + rng.normal(0.0, noise, pix.shape)
    if outlier_fraction > 0:
        out = rng.random(len(pix)) < outlier_fraction
        pix[out] = np.stack([rng.uniform(0, WIDTH, out.sum()), rng.uniform(0, HEIGHT, out.sum())], -1)
    return obs_pose, obs_point, pix, kept


def _finish(name, R, C, f, k1, k2, X, obs_pose, obs_point, pix, kept, init, intr_model, var_radial,
            loss=A.LOSS_SQUARED, loss_scale=0.0, waived=True):
    """Renumber kept points, perturb the state and build the BundleProblem."""
    new = -np.ones(len(X), dtype=np.int64)
    new[kept] = np.arange(kept.sum())
    Xk = X[kept]
    rot0, cen0, pts0, f0, k10, k20 = init(R, C, Xk, f, k1, k2)
    intr = [Intrinsics(focal=float(f0[i]), cx=CX, cy=CY, width=WIDTH, height=HEIGHT, model=intr_model,
                       k1=float(k10[i]), k2=float(k20[i]), var_focal=True, var_radial=var_radial,
                       certification_waived=waived) for i in range(len(R))]
    prob = BundleProblem(
        pose_rotation=rot0, pose_center=cen0, intrinsics=intr, point_xyz=pts0,
        obs_pose=obs_pose, obs_intrinsic=obs_pose.copy(), obs_point=new[obs_point].astype(np.uint32),
        obs_pixel=pix, robust_loss=loss, loss_scale=loss_scale)
    return Scene(prob, R, C, Xk, f, k1, k2, name)


def ring_scene(seed: int = 0xBA0001, n_cameras: int = 10, n_points: int = 1000, noise: float = 0.5) -> Scene:
    """10 pinhole cameras (f ~ 1000) on a radius-10 ring in the XZ plane looking at the origin;
    points U[-3, 3]^3; each point keeps 2–10 of the cameras that see it."""
    rng = np.random.default_rng(seed)
    th = 2.0 * np.pi * np.arange(n_cameras) / n_cameras
    C = np.stack([10.0 * np.cos(th), np.zeros_like(th), 10.0 * np.sin(th)], -1)
    R = np.stack([look_at(c, np.zeros(3)) for c in C])
    f = np.full(n_cameras, 1000.0)
    k = np.zeros(n_cameras)
    X = rng.uniform(-3.0, 3.0, (n_points, 3))
    cap = rng.integers(2, 11, n_points)
    op, opt, pix, kept = _observe(R, C, f, k, k, X, rng, cap, noise)

    def init(R, C, Xk, f, k1, k2):
        dR = exp_so3(rng.normal(0.0, 0.01, (len(R), 3)))
        return (dR @ R, C + rng.normal(0.0, 0.05, C.shape), Xk + rng.normal(0.0, 0.05, Xk.shape),
                f.copy(), k1.copy(), k2.copy())

    return _finish("ring", R, C, f, k, k, X, op, opt, pix, kept, init, A.CAMERA_PINHOLE, False)


def grid_scene(seed: int = 0xBA0002, nx: int = 20, ny: int = 10, n_points: int = 50_000, noise: float = 0.5,
               track: str = "binomial", outlier_fraction: float = 0.0, loss: int = A.LOSS_SQUARED,
               loss_scale: float = 0.0, perturb: float = 1.0) -> Scene:
    """Cameras on an nx-by-ny grid (spacing 1, z = 0) looking along +z with rotation jitter
    N(0, 0.02); radial intrinsics f = 1000 (1 +- 5 %), k1, k2 ~ U[-0.1, 0.1]; points at depth
    8–15 over the grid's footprint.  Track lengths are 2 + Binomial(10, 0.4) (2–12, mean 6) or,
    with track="geometric", geometric with mean 6 clipped to [2, 50]."""
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid(np.arange(nx) - (nx - 1) / 2.0, np.arange(ny) - (ny - 1) / 2.0, indexing="xy")
    C = np.stack([gx.ravel(), gy.ravel(), np.zeros(nx * ny)], -1) + rng.normal(0.0, 0.05, (nx * ny, 3))
    R = exp_so3(rng.normal(0.0, 0.02, (nx * ny, 3)))
    m = nx * ny
    f = 1000.0 * (1.0 + rng.uniform(-0.05, 0.05, m))
    k1 = rng.uniform(-0.1, 0.1, m)
    k2 = rng.uniform(-0.1, 0.1, m)
    X = np.stack([rng.uniform(-nx / 2.0 - 2.0, nx / 2.0 + 2.0, n_points),
                  rng.uniform(-ny / 2.0 - 2.0, ny / 2.0 + 2.0, n_points),
                  rng.uniform(8.0, 15.0, n_points)], -1)
    if track == "geometric":
        cap = np.clip(rng.geometric(1.0 / 6.0, n_points), 2, 50)
    elif track == "window":
        cap = rng.integers(2, 9, n_points)
    else:
        cap = 2 + rng.binomial(10, 0.4, n_points)
    op, opt, pix, kept = _observe_grid(R, C, f, k1, k2, X, rng, cap, noise, outlier_fraction, nx, ny)

    def init(R, C, Xk, f, k1, k2):
        dR = exp_so3(rng.normal(0.0, 0.002 * perturb, (len(R), 3)))
        return (dR @ R, C + rng.normal(0.0, 0.02 * perturb, C.shape),
                Xk + rng.normal(0.0, 0.02 * perturb, Xk.shape),
                f * (1.0 + rng.normal(0.0, 0.002 * perturb, f.shape)),
                k1 + rng.normal(0.0, 0.002 * perturb, k1.shape), k2 + rng.normal(0.0, 0.002 * perturb, k2.shape))

    return _finish(f"grid{m}", R, C, f, k1, k2, X, op, opt, pix, kept, init, A.CAMERA_BROWN_K1K2, True,
                   loss=loss, loss_scale=loss_scale)


def window_scenes(seed: int = 0xBA0004, n_windows: int = 512, cameras=(5, 40), points=(200, 5000)) -> list:
    """Local bundle-adjustment windows: window w (seed + w) has U{5..40} cameras on a line
    (spacing 1, looking along +z; a stretch of trajectory), U{200..5000} points at depth 8–15 and
    2–8 views per point; free f, k1, k2; squared loss."""
    rng = np.random.default_rng(seed)
    sizes = [(int(rng.integers(cameras[0], cameras[1] + 1)), int(rng.integers(points[0], points[1] + 1)))
             for _ in range(n_windows)]
    return [grid_scene(seed + 1 + w, m, 1, npts, track="window") for w, (m, npts) in enumerate(sizes)]


def candidate_rule_scene(seed: int = 0xBA0001):
    """The ring scene plus two points that drive the candidate rule of
    crates/openmvg-ba/src/kernel/point_geometry.rs:153-168 into its other branches.

    * A point 0.01 behind camera 3's true centre on its optical axis, seen by camera 3 at the
      principal point and by cameras 7, 8 and 9, which are fixed at their true poses.  Camera 3
      starts 0.05 behind its true centre (the point 0.04 in front of it); a step that carries the
      camera past the point leaves the point behind the candidate camera both at its stepped and
      at its previous position while the current cameras still project it, so the step is
      rejected; otherwise the stepped point is degenerate (nearly on a camera centre) and the
      point is held at its previous position.
    * A far point seen by cameras 0 and 1 at the direction of their bisecting axis at infinity,
      placed where the parallax sine is 1.75e-5 (just above TrackGeometry's 1.7e-5).

    Returns (problem, index of the near point, index of the far point)."""
    sc = ring_scene(seed)
    p = sc.problem
    Rt, Ct = sc.truth_rotation, sc.truth_center
    rot, cen = p.pose_rotation.copy(), p.pose_center.copy()
    pts, op, opt, pix = list(p.point_xyz), list(p.obs_pose), list(p.obs_point), list(p.obs_pixel)
    var = np.ones(len(rot), dtype=np.uint8)
    c, opposite = 3, (7, 8, 9)
    axis = Rt[c][2]
    near = Ct[c] - 0.01 * axis
    cen[c], rot[c] = Ct[c] - 0.05 * axis, Rt[c].copy()
    for k in opposite:
        var[k] = 0
        rot[k], cen[k] = Rt[k], Ct[k]
    q_near = len(pts)
    pts.append(near.copy())
    for cam in (c,) + opposite:
        if cam == c:
            uv = np.array([CX, CY])
        else:
            pc = Rt[cam] @ (near - Ct[cam])
            uv = np.array([1000.0 * pc[0] / pc[2] + CX, 1000.0 * pc[1] / pc[2] + CY])
        op.append(cam)
        opt.append(q_near)
        pix.append(uv)
    a, b = 0, 1
    u = Rt[a][2] + Rt[b][2]
    u /= np.linalg.norm(u)
    far = 0.5 * (Ct[a] + Ct[b]) + np.linalg.norm(Ct[a] - Ct[b]) / 1.75e-5 * u
    q_far = len(pts)
    pts.append(far)
    for cam in (a, b):
        pc = Rt[cam] @ u
        op.append(cam)
        opt.append(q_far)
        pix.append(np.array([1000.0 * pc[0] / pc[2] + CX, 1000.0 * pc[1] / pc[2] + CY]))
    prob = BundleProblem(pose_rotation=rot, pose_center=cen, intrinsics=p.intrinsics, point_xyz=np.array(pts),
                         obs_pose=np.array(op), obs_intrinsic=np.array(op), obs_point=np.array(opt),
                         obs_pixel=np.array(pix), pose_var_rotation=var, pose_var_translation=var.copy())
    return prob, q_near, q_far


def fixed_intrinsics(scene: Scene) -> Scene:
    """The same scene with every intrinsic fixed at its initial value: the calibrated pixel path
    (crates/openmvg-ba/src/native/calibrated.rs) instead of self-calibration."""
    intr = [replace(i, var_focal=False, var_radial=False, var_principal_point=False, var_tangential=False)
            for i in scene.problem.intrinsics]
    return replace(scene, problem=replace(scene.problem, intrinsics=intr), name=scene.name + "-calibrated")


def calibrated_ring_scene(seed: int = 0xBA0001, noise: float = 0.5, loss: int = A.LOSS_SQUARED,
                          loss_scale: float = 0.0) -> Scene:
    """The ring scene with its pinhole cameras (f = 1000) fixed: the calibrated pixel path with
    6-parameter pose blocks (the gauge camera 5)."""
    sc = fixed_intrinsics(ring_scene(seed, noise=noise))
    sc.problem.robust_loss, sc.problem.loss_scale = loss, loss_scale
    return sc


def calibrated_brown_scene(seed: int = 0xBA0006, nx: int = 6, ny: int = 5, n_points: int = 3000) -> Scene:
    """A small grid scene whose Brown-Conrady (k1, k2) cameras are fixed at their true values
    (k1, k2 ~ U[-0.1, 0.1]), so the calibrated residuals carry the radial distortion."""
    sc = grid_scene(seed, nx, ny, n_points, perturb=3.0)
    intr = [replace(i, focal=float(f), k1=float(a), k2=float(b)) for i, f, a, b in
            zip(sc.problem.intrinsics, sc.truth_focal, sc.truth_k1, sc.truth_k2)]
    sc = replace(sc, problem=replace(sc.problem, intrinsics=intr))
    return fixed_intrinsics(sc)


def step_limit_scene(seed: int = 0xBA0001, roll: float = 0.6, camera: int = 3) -> Scene:
    """The calibrated ring with one camera rolled about its optical axis by `roll` radians: the
    Gauss-Newton correction of that rotation exceeds the calibrated step limit of 0.5
    (crates/openmvg-ba/src/kernel/multi_view/optimizer.rs:677-702), so the first steps are scaled."""
    sc = calibrated_ring_scene(seed)
    rot = sc.problem.pose_rotation.copy()
    rot[camera] = exp_so3(np.array([0.0, 0.0, roll])) @ rot[camera]
    return replace(sc, problem=replace(sc.problem, pose_rotation=rot))


def controls_scene(seed: int = 0xBA0001, n_controls: int = 40) -> Scene:
    """The calibrated ring with `n_controls` points fixed at their true positions (survey
    controls), no fixed pose and a free gauge: the controls fix the frame, so the calibrated
    gauge plan leaves every camera free, the anchor included (native/gauge.rs:153-162)."""
    sc = calibrated_ring_scene(seed)
    pts = sc.problem.point_xyz.copy()
    fixed = np.zeros(len(pts), dtype=np.uint8)
    fixed[:n_controls] = 1
    pts[:n_controls] = sc.truth_points[:n_controls]
    return replace(sc, problem=replace(sc.problem, point_xyz=pts, point_fixed=fixed))


def narrow_field_scene(seed: int = 0xBA0031, n_points: int = 600, spread: float = 0.3) -> Scene:
    """The ring's 10 cameras with Brown (k1, k2) free and certification required, looking at
    points in a cube of half-width `spread` at the ring's centre: every point projects within
    ~0.04 (normalised) of the principal point, where the radial terms move a pixel by ~1e-4 per
    unit k1, so k1 and k2 stay unconstrained at the image corner and the certificate holds them
    at their input values (crates/openmvg-ba/src/native/self_calibration.rs:215-303)."""
    rng = np.random.default_rng(seed)
    n_cameras = 10
    th = 2.0 * np.pi * np.arange(n_cameras) / n_cameras
    C = np.stack([10.0 * np.cos(th), np.zeros_like(th), 10.0 * np.sin(th)], -1)
    R = np.stack([look_at(c, np.zeros(3)) for c in C])
    f = np.full(n_cameras, 1000.0)
    k1 = np.full(n_cameras, -0.05)
    k2 = np.full(n_cameras, 0.01)
    X = rng.uniform(-spread, spread, (n_points, 3))
    cap = rng.integers(2, 11, n_points)
    op, opt, pix, kept = _observe(R, C, f, k1, k2, X, rng, cap, 0.5)

    def init(R, C, Xk, f, k1, k2):
        dR = exp_so3(rng.normal(0.0, 0.002, (len(R), 3)))
        return (dR @ R, C + rng.normal(0.0, 0.01, C.shape), Xk + rng.normal(0.0, 0.005, Xk.shape),
                f * 1.01, k1 + 0.02, k2.copy())

    return _finish("narrow", R, C, f, k1, k2, X, op, opt, pix, kept, init, A.CAMERA_BROWN_K1K2, True, waived=False)


def rank_deficient_scene(seed: int = 0xBA0001) -> BundleProblem:
    """The ring scene plus an 11th camera (its own intrinsic group, focal free) that sees one
    point only: two residuals cannot fix its seven parameters, so the undamped reduced camera
    system is singular and the rank certificate fails (direct.rs:288-334)."""
    sc = ring_scene(seed)
    p = sc.problem
    C11 = np.array([0.0, 10.0, 0.0])
    R11 = look_at(C11, np.zeros(3), up=(0.0, 0.0, 1.0))
    q = int(p.obs_point[0])
    pc = R11 @ (sc.truth_points[q] - C11)
    uv = np.array([1000.0 * pc[0] / pc[2] + CX, 1000.0 * pc[1] / pc[2] + CY])
    intr = list(p.intrinsics) + [Intrinsics(focal=1000.0, cx=CX, cy=CY, width=WIDTH, height=HEIGHT,
                                            model=A.CAMERA_PINHOLE, var_focal=True, certification_waived=True)]
    return BundleProblem(
        pose_rotation=np.concatenate([p.pose_rotation, R11[None]]), pose_center=np.concatenate([p.pose_center, C11[None]]),
        intrinsics=intr, point_xyz=p.point_xyz.copy(),
        obs_pose=np.concatenate([p.obs_pose, [len(p.pose_rotation)]]).astype(np.uint32),
        obs_intrinsic=np.concatenate([p.obs_intrinsic, [len(p.pose_rotation)]]).astype(np.uint32),
        obs_point=np.concatenate([p.obs_point, [q]]).astype(np.uint32),
        obs_pixel=np.concatenate([p.obs_pixel, uv[None]]))


def bal_file_text(seed: int = 0xBA0041, n_cameras: int = 12, n_points: int = 800, noise: float = 0.5) -> str:
    """A BAL file written with BAL's own conventions (not through bal.py): cameras on an arc
    looking down their -z axis at points 8-14 units away, P = R(w) X + t with P.z < 0, measurement
    f (1 + k1 |p|^2 + k2 |p|^4) p, p = -P.xy / P.z, plus Gaussian noise; every point seen by 2-6
    cameras; the state perturbed.  noise = 0 gives exact measurements of the unperturbed state,
    whose cost is then zero."""
    rng = np.random.default_rng(seed)
    th = np.linspace(-0.4, 0.4, n_cameras)
    # camera i at (6 sin th, 0, -6 cos th) looking at the origin: its -z axis points there
    C = np.stack([6.0 * np.sin(th), rng.normal(0, 0.2, n_cameras), -6.0 * np.cos(th)], -1)
    R = np.stack([look_at(c, np.zeros(3)) for c in C])  # +z toward the origin
    Rb = np.diag([1.0, -1.0, -1.0])[None] @ R           # BAL: -z toward the origin
    t = -np.einsum("nij,nj->ni", Rb, C)
    f = rng.uniform(900.0, 1100.0, n_cameras)
    k1 = rng.uniform(-0.05, 0.05, n_cameras)
    k2 = rng.uniform(-0.01, 0.01, n_cameras)
    X = rng.uniform(-1.5, 1.5, (n_points, 3)) + np.array([0.0, 0.0, 4.0])
    obs = []
    for j in range(n_points):
        for c in rng.choice(n_cameras, int(rng.integers(2, min(7, n_cameras + 1))), replace=False):
            P = Rb[c] @ X[j] + t[c]
            p = -P[:2] / P[2]
            r2 = p @ p
            uv = f[c] * (1.0 + k1[c] * r2 + k2[c] * r2 * r2) * p + (rng.normal(0.0, noise, 2) if noise else 0.0)
            obs.append((int(c), j, uv[0], uv[1]))
    w = log_so3(Rb)
    if noise:
        w = w + rng.normal(0.0, 0.002, w.shape)
        t = t + rng.normal(0.0, 0.01, t.shape)
        X = X + rng.normal(0.0, 0.01, X.shape)
    lines = [f"{n_cameras} {n_points} {len(obs)}"]
    lines += [f"{c} {j} {u:.17g} {v:.17g}" for c, j, u, v in obs]
    for i in range(n_cameras):
        lines += [f"{v:.17g}" for v in (*w[i], *t[i], f[i], k1[i], k2[i])]
    lines += [f"{v:.17g}" for v in X.reshape(-1)]
    return "\n".join(lines) + "\n"


def medium_scene(seed: int = 0xBA0002) -> Scene:
    """200 cameras, 50,000 points, ~300k observations, (f, k1, k2) free, squared loss."""
    return grid_scene(seed, 20, 10, 50_000)


def large_scene(seed: int = 0xBA0003, n_points: int = 1_000_000) -> Scene:
    """2,000 cameras, 1M points, ~6M observations, geometric tracks, 2 % outliers, Huber 1 px."""
    return grid_scene(seed, 50, 40, n_points, track="geometric", outlier_fraction=0.02,
                      loss=A.LOSS_HUBER, loss_scale=1.0)


def reference_fixture_scene(camera_noise: bool = True):
    """The 4-camera, 36-point noise-free scene of the reference's sparse self-calibration tests
    (tests/sparse_self_calibration/main.rs:48-135, dense_parity.rs:104-122) as a BundleProblem
    with Gauge::FixedBaseline(0, 1), which lowers to the test's FixedBaseline{camera: 1}."""

    def rel(w, t):
        return exp_so3(np.asarray(w)), np.asarray(t, dtype=np.float64)

    truth = [(np.eye(3), np.zeros(3)), rel([0.018, 0.065, -0.012], [0.72, 0.025, 0.08]),
             rel([-0.042, -0.052, 0.024], [-0.61, 0.075, 0.13]), rel([0.052, -0.024, 0.032], [0.19, -0.51, 0.11])]
    pts = []
    for i in range(36):
        col, row = i % 6, i // 6
        pts.append([-1.3 + 0.5 * col, -0.95 + 0.38 * row, 3.25 + 0.28 * (i % 7) + 0.05 * row])
    pts = np.array(pts)
    cams_true = np.array([[680.0, 0.055, -0.022], [725.0, -0.041, 0.017], [695.0, 0.032, -0.013],
                          [740.0, -0.048, 0.020]])
    cams_init = np.array([[688.0, 0.058, -0.023], [718.0, -0.038, 0.016], [701.0, 0.030, -0.012],
                          [733.0, -0.045, 0.019]])
    obs_pose, obs_point, pix = [], [], []
    for pi, X in enumerate(pts):
        for ci, (R, t) in enumerate(truth):
            pc = R @ X + t
            n = pc[:2] / pc[2]
            r2 = n @ n
            d = 1.0 + cams_true[ci, 1] * r2 + cams_true[ci, 2] * r2 * r2
            obs_pose.append(ci)
            obs_point.append(pi)
            pix.append(cams_true[ci, 0] * d * n + np.array([CX, CY]))
    # perturbed_scene, main.rs:99-129
    init = []
    for ci, (R, t) in enumerate(truth):
        R = R.copy()
        t = t.copy()
        if ci > 0:
            R = exp_so3(np.array([0.002 * ci, -0.003, 0.0015 * ci])) @ R
            t[0] += 0.006 * ci
            t[1] -= 0.004
            t[2] += 0.003
        init.append([R, t])
    tb = np.linalg.norm(truth[1][1])
    ib = np.linalg.norm(init[1][1])
    init[1][1] = init[1][1] * tb / ib
    phase = np.arange(36, dtype=np.float64)
    pts0 = pts + np.stack([0.006 * np.sin(phase), -0.005 * np.cos(0.7 * phase), 0.008 * np.sin(0.4 * phase)], -1)
    rot = np.stack([R for R, _ in init])
    cen = np.stack([-(R.T @ t) for R, t in init])
    start = cams_init if camera_noise else cams_true
    intr = [Intrinsics(focal=float(start[i, 0]), cx=CX, cy=CY, width=WIDTH, height=HEIGHT,
                       model=A.CAMERA_BROWN_K1K2, k1=float(start[i, 1]), k2=float(start[i, 2]),
                       var_focal=True, var_radial=True, certification_waived=True) for i in range(4)]
    prob = BundleProblem(pose_rotation=rot, pose_center=cen, intrinsics=intr, point_xyz=pts0,
                         obs_pose=np.array(obs_pose), obs_intrinsic=np.array(obs_pose),
                         obs_point=np.array(obs_point), obs_pixel=np.array(pix),
                         gauge=A.GAUGE_FIXED_BASELINE, gauge_first=0, gauge_second=1,
                         gauge_distance=float(np.linalg.norm(cen[1] - cen[0])))
    tr = np.stack([R for R, _ in truth])
    tc = np.stack([-(R.T @ t) for R, t in truth])
    return Scene(prob, tr, tc, pts, cams_true[:, 0], cams_true[:, 1], cams_true[:, 2], "reference4")


def bal_scene(seed: int = 0xBA0005, n_cameras: int = 1000, n_points: int = 200_000, n_obs: int = 1_000_000,
              dtype=np.float64):
    """BAL-style parameters for the differentiation form: per camera (w(3), t(3), f, k1, k2),
    per point X(3), observations (camera, point, uv) with uv relative to the principal point.
    Every observation has positive depth and lies inside its camera's radial branch."""
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, 0.1, (n_cameras, 3))
    t = np.stack([rng.uniform(-1, 1, n_cameras), rng.uniform(-1, 1, n_cameras), rng.uniform(-0.5, 0.5, n_cameras)], -1)
    f = rng.uniform(800.0, 1200.0, n_cameras)
    k1 = rng.uniform(-0.1, 0.1, n_cameras)
    k2 = rng.uniform(-0.05, 0.05, n_cameras)
    cams = np.concatenate([w, t, f[:, None], k1[:, None], k2[:, None]], -1)
    X = np.stack([rng.uniform(-3, 3, n_points), rng.uniform(-2, 2, n_points), rng.uniform(8, 14, n_points)], -1)
    # every point at least twice, the rest uniform
    base = np.repeat(np.arange(n_points), 2)
    extra = rng.integers(0, n_points, n_obs - len(base))
    op = np.sort(np.concatenate([base, extra]))
    oc = rng.integers(0, n_cameras, n_obs)
    R = exp_so3(w)
    pc = np.einsum("nij,nj->ni", R[oc], X[op]) + t[oc]
    n = pc[:, :2] / pc[:, 2:3]
    r2 = np.sum(n * n, -1)
    d = 1.0 + k1[oc] * r2 + k2[oc] * r2 * r2
    uv = (f[oc] * d)[:, None] * n + rng.normal(0, 0.5, (n_obs, 2))
    out = rng.random(n_obs) < 0.02
    uv[out] += rng.normal(0, 20.0, (out.sum(), 2))
    # perturb the state so the gradient is not at a stationary point
    cams0 = cams.copy()
    cams0[:, :3] += rng.normal(0, 0.002, (n_cameras, 3))
    cams0[:, 3:6] += rng.normal(0, 0.01, (n_cameras, 3))
    X0 = X + rng.normal(0, 0.01, X.shape)
    return (cams0.astype(dtype), X0.astype(dtype), oc.astype(np.uint32), op.astype(np.uint32),
            uv.astype(dtype))
