"""Differentiable robust reprojection cost on the GPU.

``ReprojectionCost`` is a ``torch.autograd.Function`` over the HIP entry point ``ba_hip_cost_grad``:
the forward evaluates C = sum_o rho(|r_o|) with the radial model of
crates/openmvg-ba/src/kernel/multi_view/self_calibration/model.rs:26-129 and the robust losses of
numerics/robust.rs:8-44; the backward is a hand-written kernel that recomputes the Jacobians and
returns dC/dcameras (9 per camera: angle-axis w, t, f, k1, k2) and dC/dpoints (3 per point).
Pixels are data: no gradient flows to them.  There is no CPU path.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A
from ._lib import Handle
from .problem import raise_for


class ReprojectionTopology:
    """Observation index (camera and point per observation) resident on one GPU, with the
    workspace of every cost/gradient evaluation over it."""

    def __init__(self, obs_camera, obs_point, n_cameras: int, n_points: int, device: int = 0,
                 handle: Handle | None = None):
        cam = np.ascontiguousarray(np.asarray(obs_camera), dtype=np.uint32)
        pt = np.ascontiguousarray(np.asarray(obs_point), dtype=np.uint32)
        if cam.shape != pt.shape or cam.ndim != 1:
            raise ValueError("obs_camera and obs_point must be 1-D arrays of the same length")
        self.handle = handle if handle is not None else Handle(device)
        self.device = device
        self.n_cameras = int(n_cameras)
        self.n_points = int(n_points)
        self.n_observations = int(cam.shape[0])
        self.ptr = C.c_void_p()
        msg = A.message_buffer()
        code = self.handle.lib.ba_hip_reprojection_topology_create(
            self.handle.ptr, self.n_cameras, self.n_points, self.n_observations, cam.ctypes.data, pt.ctypes.data,
            C.byref(self.ptr), msg, len(msg))
        raise_for(code, msg.value)

    def close(self):
        if getattr(self, "ptr", None):
            self.handle.lib.ba_hip_reprojection_topology_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check(topology: ReprojectionTopology, cameras, points, obs_uv):
    import torch

    dtype = cameras.dtype
    if dtype not in (torch.float32, torch.float64):
        raise TypeError("cameras must be float32 or float64")
    for name, t, shape in (("cameras", cameras, (topology.n_cameras, 9)), ("points", points, (topology.n_points, 3)),
                           ("obs_uv", obs_uv, (topology.n_observations, 2))):
        if t.dtype != dtype:
            raise TypeError(f"{name} must be {dtype}")
        if not t.is_cuda or t.device.index != topology.device:
            raise ValueError(f"{name} must live on cuda:{topology.device}")
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} must have shape {shape}, got {tuple(t.shape)}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    return dtype


def cost_grad(topology: ReprojectionTopology, cameras, points, obs_uv, loss: int = A.LOSS_SQUARED,
              loss_scale: float = 1.0, grad_scale=None, with_grad: bool = True):
    """One call of ba_hip_cost_grad on torch's current stream.

    Returns (cost, invalid, grad_cameras, grad_points); the gradients are None when with_grad is
    False.  `invalid` counts observations outside the projection domain (they contribute nothing).
    """
    import torch

    dtype = _check(topology, cameras, points, obs_uv)
    dev = cameras.device
    cost = torch.empty((), dtype=dtype, device=dev)
    invalid = torch.empty((), dtype=torch.int32, device=dev)
    gc = torch.empty_like(cameras) if with_grad else None
    gp = torch.empty_like(points) if with_grad else None
    gs = None
    if grad_scale is not None:
        gs = grad_scale.to(device=dev, dtype=dtype).reshape(1).contiguous()
    stream = torch.cuda.current_stream(dev).cuda_stream
    msg = A.message_buffer()
    code = topology.handle.lib.ba_hip_cost_grad(
        topology.handle.ptr, topology.ptr, int(dtype == torch.float32), cameras.data_ptr(), points.data_ptr(),
        obs_uv.data_ptr(), int(loss), float(loss_scale), gs.data_ptr() if gs is not None else None, cost.data_ptr(),
        invalid.data_ptr(), gc.data_ptr() if gc is not None else None, gp.data_ptr() if gp is not None else None,
        C.c_void_p(stream), msg, len(msg))
    raise_for(code, msg.value)
    return cost, invalid, gc, gp


def _function():
    import torch

    class ReprojectionCost(torch.autograd.Function):
        """C = sum_o rho(|r_o|); backward by the hand-written HIP kernel (Jacobians recomputed)."""

        @staticmethod
        def forward(ctx, cameras, points, obs_uv, topology, loss, loss_scale):
            cameras = cameras.contiguous()
            points = points.contiguous()
            obs_uv = obs_uv.contiguous()
            cost, _, _, _ = cost_grad(topology, cameras, points, obs_uv, loss, loss_scale, with_grad=False)
            ctx.save_for_backward(cameras, points, obs_uv)
            ctx.topology = topology
            ctx.loss = loss
            ctx.loss_scale = loss_scale
            return cost

        @staticmethod
        def backward(ctx, grad_cost):
            cameras, points, obs_uv = ctx.saved_tensors
            _, _, gc, gp = cost_grad(ctx.topology, cameras, points, obs_uv, ctx.loss, ctx.loss_scale,
                                     grad_scale=grad_cost, with_grad=True)
            return gc, gp, None, None, None, None

    return ReprojectionCost


_FN = None


def reprojection_cost(cameras, points, obs_uv, topology: ReprojectionTopology, loss: int = A.LOSS_SQUARED,
                      loss_scale: float = 1.0):
    """Differentiable robust reprojection cost (a 0-d tensor of the cameras' dtype)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(cameras, points, obs_uv, topology, loss, loss_scale)
