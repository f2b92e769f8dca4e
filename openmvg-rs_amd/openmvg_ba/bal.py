"""Bundle Adjustment in the Large (BAL) problems as BundleProblem.

A BAL file lists, after its counts line "n_cameras n_points n_observations", one line per
observation "camera point x y", then 9 values per camera (Rodrigues rotation w, translation t,
focal f, radial k1, k2) and 3 per point.  Its camera looks down -z: with P = R(w) X + t,
the predicted measurement is f (1 + k1 |p|^2 + k2 |p|^4) p with p = -P.xy / P.z, in image
coordinates centred on the principal point with y up.  This is the radial model of
crates/openmvg-ba/src/kernel/multi_view/self_calibration/model.rs:26-129 after the reference's
conversion of BAL's negative-depth, negated-image convention (self_calibration/mod.rs:71-93):
the camera frame turned by F = diag(1, -1, -1) looks down +z, so

    R' = F R(w),   centre C = -R(w)^T t,   pixel = (x, -y) + (cx, cy),

with one Brown (k1, k2) intrinsic per camera, f = fx = fy, focal and radial terms free, and the
principal point (cx, cy) at the centre of an image just large enough for every measurement.
Points behind a camera (P.z >= 0 in BAL's frame) are left in the problem: the native solve
refuses them at its projection-domain check, as the reference's BundleAdjuster does (only its
kernel-level signed-depth variant, model.rs:131-161, accepts them).
"""

from __future__ import annotations

import bz2
import gzip
import io
import math
import os
import pathlib

import numpy as np

from . import _abi as A
from .problem import BundleProblem, Intrinsics
from .synthetic import exp_so3, log_so3

F = np.diag([1.0, -1.0, -1.0])


def _open_text(source) -> str:
    if isinstance(source, pathlib.Path) or (isinstance(source, str) and "\n" not in source and os.path.exists(source)):
        path = pathlib.Path(source)
        raw = path.read_bytes()
        if path.suffix == ".bz2":
            raw = bz2.decompress(raw)
        elif path.suffix == ".gz":
            raw = gzip.decompress(raw)
        return raw.decode("ascii")
    if isinstance(source, bytes):
        return source.decode("ascii")
    return str(source)


def parse_bal(source):
    """(cameras (n, 9), points (m, 3), obs_camera, obs_point, obs_xy (k, 2)) from BAL text, a
    path (.bz2 and .gz read directly) or bytes."""
    tokens = np.array(_open_text(source).split())
    if len(tokens) < 3:
        raise ValueError("BAL: missing counts line")
    nc, npt, no = (int(x) for x in tokens[:3])
    want = 3 + 4 * no + 9 * nc + 3 * npt
    if len(tokens) != want:
        raise ValueError(f"BAL: expected {want} values for {nc} cameras, {npt} points, {no} observations, "
                         f"found {len(tokens)}")
    obs = tokens[3:3 + 4 * no].reshape(no, 4)
    cam = obs[:, 0].astype(np.int64)
    pt = obs[:, 1].astype(np.int64)
    xy = obs[:, 2:].astype(np.float64)
    rest = tokens[3 + 4 * no:].astype(np.float64)
    cameras = rest[:9 * nc].reshape(nc, 9)
    points = rest[9 * nc:].reshape(npt, 3)
    if no and (cam.min() < 0 or cam.max() >= nc or pt.min() < 0 or pt.max() >= npt):
        raise ValueError("BAL: observation index out of range")
    return cameras, points, cam.astype(np.uint32), pt.astype(np.uint32), xy


def read_bal(source, certification_waived: bool = False, robust_loss: int = A.LOSS_SQUARED,
             loss_scale: float = 0.0) -> BundleProblem:
    """A BundleProblem from a BAL file (path, text or bytes)."""
    cameras, points, cam, pt, xy = parse_bal(source)
    R = exp_so3(cameras[:, :3])
    t = cameras[:, 3:6]
    rot = F[None] @ R
    centre = -np.einsum("nji,nj->ni", R, t)
    half_w = math.ceil(float(np.max(np.abs(xy[:, 0])))) + 1 if len(xy) else 1
    half_h = math.ceil(float(np.max(np.abs(xy[:, 1])))) + 1 if len(xy) else 1
    cx, cy = float(half_w), float(half_h)
    intr = [Intrinsics(focal=float(c[6]), cx=cx, cy=cy, width=2 * half_w, height=2 * half_h,
                       model=A.CAMERA_BROWN_K1K2, k1=float(c[7]), k2=float(c[8]), var_focal=True, var_radial=True,
                       certification_waived=certification_waived) for c in cameras]
    pixel = np.stack([xy[:, 0] + cx, -xy[:, 1] + cy], -1)
    return BundleProblem(pose_rotation=rot, pose_center=centre, intrinsics=intr, point_xyz=points,
                         obs_pose=cam, obs_intrinsic=cam.copy(), obs_point=pt, obs_pixel=pixel,
                         robust_loss=robust_loss, loss_scale=loss_scale)


def bal_text(problem: BundleProblem) -> str:
    """BAL text of a problem with one radial intrinsic per camera (the inverse of read_bal): every
    value written with 17 significant digits, so that read_bal(bal_text(p)) restores p's state
    to rounding (the rotations pass through their Rodrigues vectors)."""
    P = problem.n_poses
    if not np.array_equal(problem.obs_intrinsic, problem.obs_pose):
        raise ValueError("BAL stores one intrinsic per camera")
    Rb = F[None] @ problem.pose_rotation
    w = log_so3(Rb)
    t = -np.einsum("nij,nj->ni", Rb, problem.pose_center)
    out = io.StringIO()
    out.write(f"{P} {problem.n_points} {problem.n_observations}\n")
    cx = np.array([problem.intrinsics[i].cx for i in range(P)])
    cy = np.array([problem.intrinsics[i].cy for i in range(P)])
    cam = problem.obs_pose.astype(np.int64)
    x = problem.obs_pixel[:, 0] - cx[cam]
    y = -(problem.obs_pixel[:, 1] - cy[cam])
    for c, p, u, v in zip(cam, problem.obs_point, x, y):
        out.write(f"{c} {int(p)} {u:.17g} {v:.17g}\n")
    for i in range(P):
        intr = problem.intrinsics[i]
        vals = list(w[i]) + list(t[i]) + [intr.focal, intr.k1, intr.k2]
        out.write("".join(f"{v:.17g}\n" for v in vals))
    for X in problem.point_xyz:
        out.write("".join(f"{v:.17g}\n" for v in X))
    return out.getvalue()


def write_bal(problem: BundleProblem, path) -> None:
    path = pathlib.Path(path)
    text = bal_text(problem).encode("ascii")
    if path.suffix == ".bz2":
        text = bz2.compress(text)
    elif path.suffix == ".gz":
        text = gzip.compress(text)
    path.write_bytes(text)
