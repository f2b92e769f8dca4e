"""Loading of the golden fixtures in tests/golden/ (see tests/golden/make_golden.py)."""

from __future__ import annotations

import hashlib
import json
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def load(name: str) -> dict:
    return json.loads((GOLDEN / f"{name}.json").read_text())


def scene_digest(problem) -> str:
    """SHA-256 of a problem's input arrays, as make_golden.py records it."""
    h = hashlib.sha256()
    for name in ("pose_rotation", "pose_center", "point_xyz", "obs_pose", "obs_intrinsic", "obs_point",
                 "obs_pixel"):
        h.update(np.ascontiguousarray(getattr(problem, name)).tobytes())
    for it in problem.intrinsics:
        h.update(repr(sorted(vars(it).items())).encode())
    return h.hexdigest()


def arr(values) -> np.ndarray:
    return np.asarray(values, dtype=np.float64)
