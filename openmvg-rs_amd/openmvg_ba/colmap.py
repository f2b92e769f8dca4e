"""COLMAP text models as BundleProblem: the writer follows the conventions of openmvg-io's
exporter (crates/openmvg-io/src/colmap.rs:53-124, write_colmap) and the reader is its inverse.

A model is the directory ``<path>/sparse/0`` with three files:

    cameras.txt    CAMERA_ID MODEL WIDTH HEIGHT PARAMS[]
    images.txt     IMAGE_ID QW QX QY QZ TX TY TZ CAMERA_ID NAME, then POINTS2D[] as (X, Y, POINT3D_ID)
    points3D.txt   POINT3D_ID X Y Z R G B ERROR TRACK[] as (IMAGE_ID, POINT2D_IDX)

Conventions (as colmap.rs:66-98 states them for a reconstruction):

- each intrinsic becomes one COLMAP camera numbered by its index plus one; each pose one image
  numbered by its index plus one, using the camera of its observations (a pose observed through
  several intrinsics cannot be written);
- poses are world to camera: ``QW QX QY QZ`` is R and ``TX TY TZ`` is t = -R C;
- each point is numbered by its index plus one; points are written gray (128 128 128) with
  ``ERROR`` the mean reprojection error of its observations in pixels; an image lists the points it
  observes in track order (observations grouped by point, in input order) and ``POINT2D_IDX``
  indexes that list;
- COLMAP puts the centre of the top-left pixel at (0.5, 0.5) and this library at (0, 0): 0.5 is
  added to every principal point and observation coordinate on writing and taken off on reading;
- values are written in the shortest round-trip form (Python's ``repr``), so they parse back to
  the same floats;
- each camera takes the smallest COLMAP model that holds it without loss (colmap.rs:300-380):
  pinhole (or Brown without distortion) as SIMPLE_PINHOLE when fx = fy else PINHOLE; Brown with
  k1 only and fx = fy as SIMPLE_RADIAL, with k1, k2 and fx = fy as RADIAL, without k3 as OPENCV,
  else FULL_OPENCV (k4..k6 = 0).  Fisheye and spherical cameras are not written.

The reader accepts those six models (FULL_OPENCV with k4..k6 = 0), skips 2-D points whose
POINT3D_ID is -1 and takes every intrinsic fixed unless asked to free the focal length and the
radial terms (the self-calibrating path).
"""

from __future__ import annotations

import math
import pathlib
from dataclasses import replace

import numpy as np

from . import _abi as A
from .problem import BundleProblem, Intrinsics

PIXEL_OFFSET = 0.5


def rotation_to_quaternion(R: np.ndarray) -> np.ndarray:
    """(w, x, y, z) of the rotation R, w >= 0: the largest-diagonal branch of the standard
    conversion (Shepperd's method), which stays accurate for every rotation."""
    R = np.asarray(R, dtype=np.float64)
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0.0:
        s = 2.0 * math.sqrt(tr + 1.0)
        q = np.array([0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s])
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = 2.0 * math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2])
        q = np.array([(R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s])
    elif R[1, 1] > R[2, 2]:
        s = 2.0 * math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2])
        q = np.array([(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s])
    else:
        s = 2.0 * math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1])
        q = np.array([(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s])
    q /= np.linalg.norm(q)
    return -q if q[0] < 0.0 else q


def quaternion_to_rotation(q) -> np.ndarray:
    """The rotation of the quaternion (w, x, y, z), normalised first."""
    w, x, y, z = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def colmap_camera(intr: Intrinsics) -> tuple[str, list]:
    """COLMAP model name and parameters of an intrinsic (colmap.rs:300-380)."""
    fx = float(intr.focal)
    fy = fx if intr.focal_y is None else float(intr.focal_y)
    cx, cy = float(intr.cx) + PIXEL_OFFSET, float(intr.cy) + PIXEL_OFFSET
    m = intr.model
    if m in (A.CAMERA_FISHEYE, A.CAMERA_SPHERICAL):
        raise ValueError("COLMAP export covers pinhole and Brown-Conrady cameras")
    k1 = float(intr.k1) if m >= A.CAMERA_BROWN_K1 else 0.0
    k2 = float(intr.k2) if m >= A.CAMERA_BROWN_K1K2 else 0.0
    k3 = float(intr.k3) if m >= A.CAMERA_BROWN_K1K2K3 else 0.0
    p1, p2 = (float(intr.p1), float(intr.p2)) if m != A.CAMERA_PINHOLE else (0.0, 0.0)
    if all(v == 0.0 for v in (k1, k2, k3, p1, p2)):
        return ("SIMPLE_PINHOLE", [fx, cx, cy]) if fx == fy else ("PINHOLE", [fx, fy, cx, cy])
    if fx == fy and all(v == 0.0 for v in (k2, k3, p1, p2)):
        return "SIMPLE_RADIAL", [fx, cx, cy, k1]
    if fx == fy and all(v == 0.0 for v in (k3, p1, p2)):
        return "RADIAL", [fx, cx, cy, k1, k2]
    if k3 == 0.0:
        return "OPENCV", [fx, fy, cx, cy, k1, k2, p1, p2]
    return "FULL_OPENCV", [fx, fy, cx, cy, k1, k2, p1, p2, k3, 0.0, 0.0, 0.0]


def intrinsic_from_colmap(model: str, width: int, height: int, params: list, free: bool) -> Intrinsics:
    """The Intrinsics of a COLMAP camera line; `free` frees the focal length and radial terms."""
    p = [float(v) for v in params]
    want = {"SIMPLE_PINHOLE": 3, "PINHOLE": 4, "SIMPLE_RADIAL": 4, "RADIAL": 5, "OPENCV": 8, "FULL_OPENCV": 12}
    if model not in want:
        raise ValueError(f"COLMAP camera model {model} is not read (pinhole and radial/OpenCV models only)")
    if len(p) != want[model]:
        raise ValueError(f"COLMAP {model} camera has {len(p)} parameters, expected {want[model]}")
    fx = fy = k1 = k2 = k3 = q1 = q2 = 0.0
    if model == "SIMPLE_PINHOLE":
        fx, cx, cy = p
        fy = fx
    elif model == "PINHOLE":
        fx, fy, cx, cy = p
    elif model == "SIMPLE_RADIAL":
        fx, cx, cy, k1 = p
        fy = fx
    elif model == "RADIAL":
        fx, cx, cy, k1, k2 = p
        fy = fx
    elif model == "OPENCV":
        fx, fy, cx, cy, k1, k2, q1, q2 = p
    else:
        fx, fy, cx, cy, k1, k2, q1, q2, k3, k4, k5, k6 = p
        if (k4, k5, k6) != (0.0, 0.0, 0.0):
            raise ValueError("FULL_OPENCV cameras with k4..k6 are not read (no rational distortion here)")
    if model in ("SIMPLE_PINHOLE", "PINHOLE"):
        camera_model = A.CAMERA_PINHOLE
    elif k3 != 0.0:
        camera_model = A.CAMERA_BROWN_K1K2K3
    elif model == "SIMPLE_RADIAL":
        camera_model = A.CAMERA_BROWN_K1
    else:
        camera_model = A.CAMERA_BROWN_K1K2
    return Intrinsics(focal=fx, focal_y=None if fy == fx else fy, cx=cx - PIXEL_OFFSET, cy=cy - PIXEL_OFFSET,
                      width=int(width), height=int(height), model=camera_model, k1=k1, k2=k2, k3=k3, p1=q1, p2=q2,
                      var_focal=free, var_radial=free and camera_model != A.CAMERA_PINHOLE)


def _pose_intrinsics(problem: BundleProblem) -> np.ndarray:
    cam = np.full(problem.n_poses, -1, dtype=np.int64)
    for pose, intr in zip(problem.obs_pose, problem.obs_intrinsic):
        if cam[pose] >= 0 and cam[pose] != intr:
            raise ValueError(f"pose {pose} is observed through several intrinsics; a COLMAP image has one camera")
        cam[pose] = intr
    cam[cam < 0] = 0
    return cam


def _project(R, C, intr: Intrinsics, X) -> np.ndarray:
    """Pixel of X through camera (R, C) and the Brown-Conrady model of `intr`."""
    pc = R @ (X - C)
    x, y = pc[0] / pc[2], pc[1] / pc[2]
    m = intr.model
    k1 = intr.k1 if m >= A.CAMERA_BROWN_K1 else 0.0
    k2 = intr.k2 if m >= A.CAMERA_BROWN_K1K2 else 0.0
    k3 = intr.k3 if m >= A.CAMERA_BROWN_K1K2K3 else 0.0
    p1, p2 = (intr.p1, intr.p2) if m != A.CAMERA_PINHOLE else (0.0, 0.0)
    r2 = x * x + y * y
    d = 1.0 + r2 * (k1 + r2 * (k2 + r2 * k3))
    xd = x * d + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x)
    yd = y * d + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y
    fy = intr.focal if intr.focal_y is None else intr.focal_y
    return np.array([intr.focal * xd + intr.cx, fy * yd + intr.cy])


def write_colmap(problem: BundleProblem, path, solution=None, names=None) -> pathlib.Path:
    """Write `problem` (or the state of `solution`, a Solution of it) as a COLMAP text model in
    ``path/sparse/0``; returns that directory.  `names` are the image names (default
    ``image<index>.jpg``)."""
    rot = problem.pose_rotation if solution is None else solution.pose_rotation
    cen = problem.pose_center if solution is None else solution.pose_center
    xyz = problem.point_xyz if solution is None else solution.point_xyz
    intrinsics = list(problem.intrinsics)
    if solution is not None:
        for g, intr in enumerate(intrinsics):
            intrinsics[g] = replace(intr, focal=float(solution.intrinsic_focal[g]),
                                    focal_y=None if intr.focal_y is None or intr.focal_y == intr.focal
                                    else intr.focal_y,
                                    k1=float(solution.intrinsic_k1[g]), k2=float(solution.intrinsic_k2[g]))
    names = names or [f"image{i}.jpg" for i in range(problem.n_poses)]
    if any(("\n" in n or "\r" in n or "\0" in n) for n in names):
        raise ValueError("image names with a line break or NUL cannot be written")
    cam_of = _pose_intrinsics(problem)
    order = np.argsort(problem.obs_point, kind="stable")  # track order: grouped by point, input order
    image_obs = [[] for _ in range(problem.n_poses)]
    tracks = [[] for _ in range(problem.n_points)]
    errors = np.zeros(problem.n_points)
    counts = np.zeros(problem.n_points, dtype=np.int64)
    for o in order:
        pose, q = int(problem.obs_pose[o]), int(problem.obs_point[o])
        uv = problem.obs_pixel[o]
        tracks[q].append((pose + 1, len(image_obs[pose])))
        image_obs[pose].append((float(uv[0]) + PIXEL_OFFSET, float(uv[1]) + PIXEL_OFFSET, q + 1))
        pred = _project(rot[pose], cen[pose], intrinsics[problem.obs_intrinsic[o]], xyz[q])
        errors[q] += float(np.hypot(*(pred - uv)))
        counts[q] += 1
    out = pathlib.Path(path) / "sparse" / "0"
    out.mkdir(parents=True, exist_ok=True)
    used = sorted(set(int(c) for c in cam_of))
    lines = ["# Camera list with one line of data per camera:", "#   CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]",
             f"# Number of cameras: {len(used)}"]
    for g in used:
        model, params = colmap_camera(intrinsics[g])
        lines.append(" ".join([str(g + 1), model, str(intrinsics[g].width), str(intrinsics[g].height)] +
                              [repr(float(v)) for v in params]))
    (out / "cameras.txt").write_text("\n".join(lines) + "\n")
    lines = ["# Image list with two lines of data per image:", "#   IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME",
             "#   POINTS2D[] as (X, Y, POINT3D_ID)", f"# Number of images: {problem.n_poses}"]
    for i in range(problem.n_poses):
        q = rotation_to_quaternion(rot[i])
        t = -rot[i] @ cen[i]
        lines.append(" ".join([str(i + 1)] + [repr(float(v)) for v in q] + [repr(float(v)) for v in t] +
                              [str(int(cam_of[i]) + 1), names[i]]))
        lines.append(" ".join(f"{repr(x)} {repr(y)} {pid}" for x, y, pid in image_obs[i]))
    (out / "images.txt").write_text("\n".join(lines) + "\n")
    lines = ["# 3D point list with one line of data per point:",
             "#   POINT3D_ID, X, Y, Z, R, G, B, ERROR, TRACK[] as (IMAGE_ID, POINT2D_IDX)",
             f"# Number of points: {problem.n_points}"]
    for q in range(problem.n_points):
        err = errors[q] / counts[q] if counts[q] else 0.0
        lines.append(" ".join([str(q + 1)] + [repr(float(v)) for v in xyz[q]] + ["128", "128", "128", repr(float(err))] +
                              [f"{img} {idx}" for img, idx in tracks[q]]))
    (out / "points3D.txt").write_text("\n".join(lines) + "\n")
    return out


def _data_lines(path: pathlib.Path) -> list:
    return [ln.rstrip("\r") for ln in path.read_text().split("\n") if ln.strip() and not ln.lstrip().startswith("#")]


def read_colmap(path, free_intrinsics: bool = False, robust_loss: int = A.LOSS_SQUARED,
                loss_scale: float = 0.0) -> tuple[BundleProblem, list]:
    """The BundleProblem of a COLMAP text model (``path`` is the model directory, or a directory
    holding ``sparse/0``) and its image names.  Poses are numbered by ascending IMAGE_ID,
    intrinsics by ascending CAMERA_ID, points by ascending POINT3D_ID; observations are listed
    point by point in track order."""
    root = pathlib.Path(path)
    if not (root / "cameras.txt").exists() and (root / "sparse" / "0" / "cameras.txt").exists():
        root = root / "sparse" / "0"
    cams = {}
    for ln in _data_lines(root / "cameras.txt"):
        f = ln.split()
        cams[int(f[0])] = intrinsic_from_colmap(f[1], int(f[2]), int(f[3]), f[4:], free_intrinsics)
    cam_ids = sorted(cams)
    cam_index = {c: i for i, c in enumerate(cam_ids)}
    lines = [ln for ln in (root / "images.txt").read_text().split("\n") if not ln.lstrip().startswith("#")]
    while lines and not lines[-1].strip():
        lines.pop()
    if len(lines) % 2:
        raise ValueError("images.txt: every image takes two lines")
    images = {}
    for k in range(0, len(lines), 2):
        f = lines[k].split()
        if len(f) < 10:
            raise ValueError(f"images.txt: malformed image line {lines[k]!r}")
        pts = lines[k + 1].split()
        if len(pts) % 3:
            raise ValueError("images.txt: POINTS2D[] is not a list of (X, Y, POINT3D_ID)")
        obs = [(float(pts[j]), float(pts[j + 1]), int(pts[j + 2])) for j in range(0, len(pts), 3)]
        images[int(f[0])] = ([float(v) for v in f[1:5]], [float(v) for v in f[5:8]], int(f[8]),
                             " ".join(f[9:]), obs)
    image_ids = sorted(images)
    image_index = {i: k for k, i in enumerate(image_ids)}
    points = {}
    for ln in _data_lines(root / "points3D.txt"):
        f = ln.split()
        track = [(int(f[j]), int(f[j + 1])) for j in range(8, len(f), 2)]
        points[int(f[0])] = ([float(v) for v in f[1:4]], track)
    point_ids = sorted(points)
    point_index = {p: k for k, p in enumerate(point_ids)}
    n = len(image_ids)
    rot = np.zeros((n, 3, 3))
    cen = np.zeros((n, 3))
    names = []
    for k, i in enumerate(image_ids):
        q, t, cam, name, _ = images[i]
        if cam not in cam_index:
            raise ValueError(f"image {i} uses camera {cam}, which cameras.txt does not list")
        rot[k] = quaternion_to_rotation(q)
        cen[k] = -rot[k].T @ np.array(t)
        names.append(name)
    op, oi, opt, pix = [], [], [], []
    for p in point_ids:
        _, track = points[p]
        for img, idx in track:
            if img not in images:
                raise ValueError(f"point {p} is observed in image {img}, which images.txt does not list")
            q, t, cam, _, obs = images[img]
            if idx >= len(obs) or obs[idx][2] != p:
                raise ValueError(f"point {p}: POINT2D_IDX {idx} of image {img} does not observe it")
            op.append(image_index[img])
            oi.append(cam_index[cam])
            opt.append(point_index[p])
            pix.append((obs[idx][0] - PIXEL_OFFSET, obs[idx][1] - PIXEL_OFFSET))
    problem = BundleProblem(
        pose_rotation=rot, pose_center=cen, intrinsics=[cams[c] for c in cam_ids],
        point_xyz=np.array([points[p][0] for p in point_ids]).reshape(-1, 3),
        obs_pose=np.array(op, dtype=np.uint32), obs_intrinsic=np.array(oi, dtype=np.uint32),
        obs_point=np.array(opt, dtype=np.uint32), obs_pixel=np.array(pix).reshape(-1, 2),
        robust_loss=robust_loss, loss_scale=loss_scale)
    return problem, names
