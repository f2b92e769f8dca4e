"""Shared test helpers: native (lowered) problems as host or device views, oracle calls."""

from __future__ import annotations

import ctypes as C

import numpy as np

from openmvg_ba import _abi as A
from openmvg_ba.problem import device_problem, obs_point_index, prepare_with

TOPO_KEYS = ("camera_rotation", "camera_translation", "group_params", "camera_group", "point_xyz", "obs_camera",
             "obs_uv", "point_offsets", "point_fixed", "camera_offset", "camera_dim", "group_offset", "group_dim")


def native_host(oracle, problem):
    """The oracle-lowered problem as host arrays plus its obs->point index."""
    nat = prepare_with(oracle, "bao_prepare", problem)
    arrays = {k: np.ascontiguousarray(v) for k, v in nat.arrays.items()}
    arrays["obs_point"] = obs_point_index(arrays["point_offsets"])
    return nat, arrays


def host_view(nat, arrays, loss, loss_scale, state=None):
    return device_problem(arrays, nat.sizes, loss, loss_scale, nat.gauge_camera, nat.scene_scale, state=state,
                          obs_point=arrays["obs_point"])


def to_device(arrays):
    import torch

    out = {}
    for k in TOPO_KEYS + ("obs_point",):
        out[k] = torch.from_numpy(np.ascontiguousarray(arrays[k])).to("cuda")
    return out


def device_view(nat, dev_arrays, host_arrays, loss, loss_scale):
    """Device view; the gauge baseline/basis come from the host copy of the translations."""
    d = device_problem(dev_arrays, nat.sizes, loss, loss_scale, nat.gauge_camera, nat.scene_scale,
                       state={"camera_translation": host_arrays["camera_translation"]},
                       obs_point=dev_arrays["obs_point"])
    d.camera_translation = A.vptr(dev_arrays["camera_translation"])
    return d


def oracle_assemble(oracle, view, D, np_, n_params, lam):
    S = np.zeros((D, D))
    g = np.zeros(D)
    fg = np.zeros(D)
    dg = np.zeros(D)
    sc = A.SchurScalars()
    oracle.bao_assemble(C.byref(view), lam, S.ctypes.data, g.ctypes.data, fg.ctypes.data, dg.ctypes.data,
                        None, None, None, C.byref(sc))
    return S, g, fg, dg, sc


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)
