"""Python mirror of openmvg-ba's public problem and solver contract.

The classes follow crates/openmvg-ba/src/problem.rs (BundleProblem, BundlePose, BundleIntrinsic,
BundlePoint, BundleObservation, RobustLoss, Gauge) and crates/openmvg-ba/src/solver.rs
(SolveConfig :74-155, SolveStatus :172-185, SolveReport :461-473, Solution :803-876,
SolveError :1474-1493).  Storage is structure-of-arrays numpy, which is what the C ABI
(include/ba_hip.h) reads.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


# ---- errors: SolveError (solver.rs:1474-1493) ---------------------------------------------
class SolveError(Exception):
    code = A.BA_ERR_INTERNAL


class InvalidConfig(SolveError):
    code = A.BA_ERR_INVALID_CONFIG


class Unsupported(SolveError):
    """The sanctioned decline: a caller may hand the problem to another adjuster."""

    code = A.BA_ERR_UNSUPPORTED


class ResourceExhausted(SolveError):
    code = A.BA_ERR_RESOURCE_EXHAUSTED


class Numerical(SolveError):
    code = A.BA_ERR_NUMERICAL


class Internal(SolveError):
    code = A.BA_ERR_INTERNAL


class ProblemError(SolveError):
    """A problem rejected by the checks of BundleProblemBuilder::build (problem.rs:492-576)."""

    code = A.BA_ERR_INVALID_PROBLEM


_ERRORS = {
    A.BA_ERR_INVALID_CONFIG: InvalidConfig,
    A.BA_ERR_UNSUPPORTED: Unsupported,
    A.BA_ERR_RESOURCE_EXHAUSTED: ResourceExhausted,
    A.BA_ERR_NUMERICAL: Numerical,
    A.BA_ERR_INTERNAL: Internal,
    A.BA_ERR_INVALID_PROBLEM: ProblemError,
}


def raise_for(code: int, message: bytes | str) -> None:
    if code == A.BA_OK:
        return
    if isinstance(message, bytes):
        message = message.decode(errors="replace")
    raise _ERRORS.get(code, Internal)(message)


# ---- problem ------------------------------------------------------------------------------
@dataclass
class Intrinsics:
    """BundleIntrinsic: a perspective camera, its variable mask and certification policy."""

    focal: float
    cx: float
    cy: float
    width: int
    height: int
    model: int = A.CAMERA_PINHOLE
    k1: float = 0.0
    k2: float = 0.0
    k3: float = 0.0
    p1: float = 0.0
    p2: float = 0.0
    focal_y: float | None = None
    var_focal: bool = False
    var_principal_point: bool = False
    var_radial: bool = False
    var_tangential: bool = False
    fixed_calibration: bool = False
    certification_waived: bool = False

    def to_c(self) -> A.Intrinsic:
        out = A.Intrinsic()
        out.model = self.model
        out.fx = self.focal
        out.fy = self.focal if self.focal_y is None else self.focal_y
        out.cx, out.cy = self.cx, self.cy
        out.k1, out.k2, out.k3, out.p1, out.p2 = self.k1, self.k2, self.k3, self.p1, self.p2
        out.width, out.height = self.width, self.height
        out.fixed_calibration = int(self.fixed_calibration)
        out.var_focal = int(self.var_focal)
        out.var_principal_point = int(self.var_principal_point)
        out.var_radial = int(self.var_radial)
        out.var_tangential = int(self.var_tangential)
        out.certification_waived = int(self.certification_waived)
        return out


@dataclass
class BundleProblem:
    """BundleProblem (problem.rs:447-455) as structure-of-arrays.

    pose_rotation: (P,3,3) world-to-camera; pose_center: (P,3) world camera centres.
    A pose is fixed when neither its rotation nor its translation is variable.
    """

    pose_rotation: np.ndarray
    pose_center: np.ndarray
    intrinsics: list
    point_xyz: np.ndarray
    obs_pose: np.ndarray
    obs_intrinsic: np.ndarray
    obs_point: np.ndarray
    obs_pixel: np.ndarray
    pose_var_rotation: np.ndarray | None = None
    pose_var_translation: np.ndarray | None = None
    point_fixed: np.ndarray | None = None
    residual_space: int = A.RESIDUAL_PIXELS
    robust_loss: int = A.LOSS_SQUARED
    loss_scale: float = 0.0
    gauge: int = A.GAUGE_FREE
    gauge_first: int = 0
    gauge_second: int = 0
    gauge_distance: float = 0.0

    def __post_init__(self):
        P = len(self.pose_rotation)
        self.pose_rotation = np.ascontiguousarray(self.pose_rotation, dtype=np.float64).reshape(P, 3, 3)
        self.pose_center = np.ascontiguousarray(self.pose_center, dtype=np.float64).reshape(P, 3)
        self.point_xyz = np.ascontiguousarray(self.point_xyz, dtype=np.float64).reshape(-1, 3)
        self.obs_pose = np.ascontiguousarray(self.obs_pose, dtype=np.uint32)
        self.obs_intrinsic = np.ascontiguousarray(self.obs_intrinsic, dtype=np.uint32)
        self.obs_point = np.ascontiguousarray(self.obs_point, dtype=np.uint32)
        self.obs_pixel = np.ascontiguousarray(self.obs_pixel, dtype=np.float64).reshape(-1, 2)
        if self.pose_var_rotation is None:
            self.pose_var_rotation = np.ones(P, dtype=np.uint8)
        if self.pose_var_translation is None:
            self.pose_var_translation = np.ones(P, dtype=np.uint8)
        self.pose_var_rotation = np.ascontiguousarray(self.pose_var_rotation, dtype=np.uint8)
        self.pose_var_translation = np.ascontiguousarray(self.pose_var_translation, dtype=np.uint8)
        if self.point_fixed is not None:
            self.point_fixed = np.ascontiguousarray(self.point_fixed, dtype=np.uint8)

    @property
    def n_poses(self) -> int:
        return len(self.pose_rotation)

    @property
    def n_points(self) -> int:
        return len(self.point_xyz)

    @property
    def n_observations(self) -> int:
        return len(self.obs_point)

    def to_c(self):
        """The C view and the objects that keep its memory alive."""
        intr = (A.Intrinsic * len(self.intrinsics))(*[i.to_c() for i in self.intrinsics])
        p = A.Problem()
        p.n_poses = self.n_poses
        p.pose_rotation = A.ptr(self.pose_rotation, C.c_double)
        p.pose_center = A.ptr(self.pose_center, C.c_double)
        p.pose_var_rotation = A.ptr(self.pose_var_rotation, C.c_uint8)
        p.pose_var_translation = A.ptr(self.pose_var_translation, C.c_uint8)
        p.n_intrinsics = len(self.intrinsics)
        p.intrinsics = C.cast(intr, C.POINTER(A.Intrinsic))
        p.n_points = self.n_points
        p.point_xyz = A.ptr(self.point_xyz, C.c_double)
        p.point_fixed = A.ptr(self.point_fixed, C.c_uint8)
        p.n_observations = self.n_observations
        p.obs_pose = A.ptr(self.obs_pose, C.c_uint32)
        p.obs_intrinsic = A.ptr(self.obs_intrinsic, C.c_uint32)
        p.obs_point = A.ptr(self.obs_point, C.c_uint32)
        p.obs_pixel = A.ptr(self.obs_pixel, C.c_double)
        p.residual_space = self.residual_space
        p.robust_loss = self.robust_loss
        p.loss_scale = self.loss_scale
        p.gauge = self.gauge
        p.gauge_first = self.gauge_first
        p.gauge_second = self.gauge_second
        p.gauge_distance = self.gauge_distance
        return p, (intr, self)


@dataclass(frozen=True)
class SolveConfig:
    """SolveConfig (solver.rs:74-155) plus this implementation's precision and PCG cap."""

    max_iterations: int = 50
    function_tolerance: float = 1.0e-6
    gradient_tolerance: float = 1.0e-10
    parameter_tolerance: float = 1.0e-8
    linear_solver: int = A.SOLVER_AUTO
    workspace_budget_bytes: int = 0
    precision: int = A.PRECISION_FP64
    pcg_max_iterations: int = 0
    pcg_relative_tolerance: float = 0.0
    profile_kernel: int = A.PROFILE_NONE
    certification: int = A.CERTIFY_FULL

    def with_(self, **kw) -> "SolveConfig":
        return dataclasses.replace(self, **kw)

    def to_c(self) -> A.SolveConfigC:
        c = A.SolveConfigC()
        c.max_iterations = self.max_iterations
        c.function_tolerance = self.function_tolerance
        c.gradient_tolerance = self.gradient_tolerance
        c.parameter_tolerance = self.parameter_tolerance
        c.linear_solver = self.linear_solver
        c.workspace_budget_bytes = self.workspace_budget_bytes
        c.precision = self.precision
        c.pcg_max_iterations = self.pcg_max_iterations
        c.pcg_relative_tolerance = self.pcg_relative_tolerance
        c.profile_kernel = self.profile_kernel
        c.certification = self.certification
        return c


@dataclass
class SolveReport:
    status: str
    fields: dict

    def __getattr__(self, name):
        try:
            return self.__dict__["fields"][name]
        except KeyError as exc:
            raise AttributeError(name) from exc


@dataclass
class Solution:
    pose_rotation: np.ndarray
    pose_center: np.ndarray
    intrinsic_focal: np.ndarray
    intrinsic_k1: np.ndarray
    intrinsic_k2: np.ndarray
    point_xyz: np.ndarray
    point_quarantined: np.ndarray
    report: SolveReport
    trace: dict = field(default_factory=dict)
    intrinsic_unconstrained: np.ndarray | None = None
    intrinsic_covariance: np.ndarray | None = None  # (G, 3, 3), NaN where none was formed


def _solution_buffers(problem: BundleProblem, sol: A.SolutionC) -> dict:
    """Caller-allocated solution arrays for `problem`, wired into `sol`."""
    P, N, G = problem.n_poses, problem.n_points, len(problem.intrinsics)
    out = dict(
        pose_rotation=np.zeros((P, 3, 3)), pose_center=np.zeros((P, 3)),
        intrinsic_focal=np.zeros(G), intrinsic_k1=np.zeros(G), intrinsic_k2=np.zeros(G),
        point_xyz=np.zeros((N, 3)), point_quarantined=np.zeros(N, dtype=np.uint8),
        intrinsic_unconstrained=np.zeros(G, dtype=np.uint8), intrinsic_covariance=np.full((G, 3, 3), np.nan),
    )
    for name in ("pose_rotation", "pose_center", "intrinsic_focal", "intrinsic_k1", "intrinsic_k2",
                 "point_xyz"):
        setattr(sol, name, A.ptr(out[name], C.c_double))
    sol.point_quarantined = A.ptr(out["point_quarantined"], C.c_uint8)
    sol.intrinsic_unconstrained = A.ptr(out["intrinsic_unconstrained"], C.c_uint8)
    sol.intrinsic_covariance = A.ptr(out["intrinsic_covariance"], C.c_double)
    return out


def _report(rep: A.SolveReportC) -> SolveReport:
    fields = {name: getattr(rep, name) for name, _ in A.SolveReportC._fields_}
    return SolveReport(status=A.STATUS_NAMES.get(rep.status, str(rep.status)), fields=fields)


def solve_with(lib, function: str, problem: BundleProblem, config: SolveConfig, *handle,
               trace_capacity: int | None = None) -> Solution:
    """Call a `ba_solve`-shaped C entry point of `lib` and wrap its outputs."""
    cp, keep = problem.to_c()
    cfg = config.to_c()
    sol = A.SolutionC()
    out = _solution_buffers(problem, sol)
    rep = A.SolveReportC()
    cap = trace_capacity if trace_capacity is not None else max(config.max_iterations, 1)
    tr_arrays = dict(cost=np.full(cap, np.nan), damping=np.full(cap, np.nan),
                     candidate_cost=np.full(cap, np.nan), predicted_reduction=np.full(cap, np.nan),
                     gain_ratio=np.full(cap, np.nan), accepted=np.zeros(cap, dtype=np.uint8),
                     pcg_iterations=np.zeros(cap, dtype=np.uint32), seconds=np.full(cap, np.nan))
    tr = A.IterationTrace()
    tr.capacity = cap
    tr.count = 0
    for name in ("cost", "damping", "candidate_cost", "predicted_reduction", "gain_ratio", "seconds"):
        setattr(tr, name, A.ptr(tr_arrays[name], C.c_double))
    tr.accepted = A.ptr(tr_arrays["accepted"], C.c_uint8)
    tr.pcg_iterations = A.ptr(tr_arrays["pcg_iterations"], C.c_uint32)
    msg = A.message_buffer()
    fn = getattr(lib, function)
    code = fn(*handle, C.byref(cp), C.byref(cfg), C.byref(sol), C.byref(rep), C.byref(tr), msg, len(msg))
    del keep
    raise_for(code, msg.value)
    report = _report(rep)
    n = tr.count
    trace = {k: v[:n].copy() for k, v in tr_arrays.items()}
    return Solution(out["pose_rotation"], out["pose_center"], out["intrinsic_focal"],
                    out["intrinsic_k1"], out["intrinsic_k2"], out["point_xyz"],
                    out["point_quarantined"], report, trace, out["intrinsic_unconstrained"],
                    out["intrinsic_covariance"])


def plan_with(lib, function: str, problem: BundleProblem, config: SolveConfig) -> dict:
    """The camera-step choice of a `ba_hip_plan`-shaped entry point (LinearSolver::Auto's
    execution_policy.rs:190-227) as a dict of the ba_execution_plan fields."""
    cp, keep = problem.to_c()
    cfg = config.to_c()
    plan = A.ExecutionPlanC()
    msg = A.message_buffer()
    code = getattr(lib, function)(C.byref(cp), C.byref(cfg), C.byref(plan), msg, len(msg))
    del keep
    raise_for(code, msg.value)
    return {name: getattr(plan, name) for name, _ in A.ExecutionPlanC._fields_ if name != "_pad"}


@dataclass
class NativeProblem:
    """The lowered problem and index maps that ba_hip_prepare / bao_prepare return."""

    sizes: dict
    arrays: dict
    gauge_camera: int
    scene_scale: float


def prepare_with(lib, function: str, problem: BundleProblem) -> NativeProblem:
    cp, keep = problem.to_c()
    sizes = A.NativeSizes()
    msg = A.message_buffer()
    fn = getattr(lib, function)
    raise_for(fn(C.byref(cp), C.byref(sizes), None, msg, len(msg)), msg.value)
    s = {name: int(getattr(sizes, name)) for name, _ in A.NativeSizes._fields_}
    nc, ng, np_, no, D = s["n_cameras"], s["n_groups"], s["n_points"], s["n_observations"], s["dimension"]
    shapes = dict(
        camera_rotation=((nc, 9), np.float64), camera_translation=((nc, 3), np.float64),
        group_params=((ng, 3), np.float64), camera_group=((nc,), np.uint32),
        camera_source=((nc,), np.uint32), point_xyz=((np_, 3), np.float64),
        point_source=((np_,), np.uint32), obs_camera=((no,), np.uint32), obs_source=((no,), np.uint32),
        obs_uv=((no, 2), np.float64), point_offsets=((np_ + 1,), np.uint32),
        camera_fixed=((nc,), np.uint8), point_fixed=((np_,), np.uint8),
        camera_offset=((nc,), np.uint32), camera_dim=((nc,), np.uint32),
        group_offset=((ng,), np.uint32), group_dim=((ng,), np.uint32),
        cross_offsets=((np_ + 1,), np.uint32), point_parameters=((s["n_point_parameters"],), np.uint32),
        pattern_col_ptr=((D + 1,), np.uint32), pattern_row_idx=((s["n_pattern_entries"],), np.uint32),
    )
    arrays = {k: np.zeros(shape, dtype=dt) for k, (shape, dt) in shapes.items()}
    native = A.NativeProblem()
    ctype = {np.float64: C.c_double, np.uint32: C.c_uint32, np.uint8: C.c_uint8}
    for k, (shape, dt) in shapes.items():
        setattr(native, k, A.ptr(arrays[k], ctype[dt]))
    raise_for(fn(C.byref(cp), C.byref(sizes), C.byref(native), msg, len(msg)), msg.value)
    del keep
    return NativeProblem(s, arrays, int(native.gauge_camera), float(native.scene_scale))


def s2_tangent_basis(t: np.ndarray) -> np.ndarray:
    """Rows b0, b1 of the S^2 tangent basis at t (numerics/manifold.rs:129-141)."""
    tv = t / np.sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2])
    a = np.abs(tv)
    if a[0] <= a[1] and a[0] <= a[2]:
        ref = np.array([1.0, 0.0, 0.0])
    elif a[1] <= a[2]:
        ref = np.array([0.0, 1.0, 0.0])
    else:
        ref = np.array([0.0, 0.0, 1.0])

    def cross(x, y):
        return np.array([x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]])

    def normalize(v):
        return v / np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])

    b0 = normalize(cross(tv, ref))
    b1 = normalize(cross(tv, b0))
    return np.concatenate([b0, b1])


def device_problem(arrays: dict, sizes: dict, loss: int, loss_scale: float, gauge_camera: int,
                   scene_scale: float, state: dict | None = None, obs_point=None) -> A.DeviceProblem:
    """A ba_device_problem view over `arrays` (numpy arrays for host calls, torch tensors for
    device calls).  `state` overrides camera_rotation/camera_translation/group_params/point_xyz.
    The gauge baseline and basis are computed from the (host copy of the) translations."""
    src = dict(arrays)
    if state:
        src.update(state)
    if obs_point is None:
        obs_point = src.get("obs_point")
    d = A.DeviceProblem()
    d.n_cameras = sizes["n_cameras"]
    d.n_groups = sizes["n_groups"]
    d.n_points = sizes["n_points"]
    d.n_observations = sizes["n_observations"]
    d.dimension = sizes["dimension"]
    for name in ("camera_rotation", "camera_translation", "group_params", "camera_group", "point_xyz",
                 "obs_camera", "obs_uv", "point_offsets", "point_fixed", "camera_offset", "camera_dim",
                 "group_offset", "group_dim"):
        setattr(d, name, A.vptr(src[name]))
    d.obs_point = A.vptr(obs_point)
    kernel_scale = loss_scale / np.sqrt(2.0) if loss == A.LOSS_CAUCHY else loss_scale
    d.loss = loss
    d.loss_scale = kernel_scale
    d.gauge_camera = gauge_camera
    d.scene_scale = scene_scale
    if gauge_camera >= 0:
        t = src["camera_translation"]
        t = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
        tg = t.reshape(-1, 3)[gauge_camera]
        d.baseline = float(np.sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]))
        basis = s2_tangent_basis(tg)
        for k in range(6):
            d.basis[k] = basis[k]
    return d


def obs_point_index(point_offsets: np.ndarray) -> np.ndarray:
    """Owning point of each observation of a point-major CSR."""
    counts = np.diff(point_offsets.astype(np.int64))
    return np.repeat(np.arange(len(counts), dtype=np.uint32), counts)


def solve_batched_with(lib, handle, problems: list, config: SolveConfig, mode: int = A.BATCH_GRAPH,
                       trace_capacity: int | None = None):
    """ba_hip_solve_batched over `problems`: a list of Solutions (each with its report) and a dict
    with the batch trace ([iteration, window] arrays, the chosen tile and loop time)."""
    W = len(problems)
    cps = [p.to_c() for p in problems]
    arr = (A.Problem * W)(*[c for c, _ in cps])
    sols = (A.SolutionC * W)()
    outs = [_solution_buffers(p, sols[w]) for w, p in enumerate(problems)]
    reps = (A.SolveReportC * W)()
    cap = trace_capacity if trace_capacity is not None else config.max_iterations
    tr_arrays = dict(cost=np.full((cap, W), np.nan), damping=np.full((cap, W), np.nan),
                     candidate_cost=np.full((cap, W), np.nan), accepted=np.zeros((cap, W), dtype=np.uint8))
    tr = A.BatchTrace()
    tr.capacity = cap
    for name in ("cost", "damping", "candidate_cost"):
        setattr(tr, name, A.ptr(tr_arrays[name], C.c_double))
    tr.accepted = A.ptr(tr_arrays["accepted"], C.c_uint8)
    cfg = config.to_c()
    msg = A.message_buffer()
    code = lib.ba_hip_solve_batched(handle, W, arr, C.byref(cfg), sols, reps, C.byref(tr), mode, msg, len(msg))
    del cps
    raise_for(code, msg.value)
    solutions = []
    for w in range(W):
        o = outs[w]
        iters = reps[w].iterations
        trace = {k: v[:iters, w].copy() for k, v in tr_arrays.items()}
        solutions.append(Solution(o["pose_rotation"], o["pose_center"], o["intrinsic_focal"], o["intrinsic_k1"],
                                  o["intrinsic_k2"], o["point_xyz"], o["point_quarantined"], _report(reps[w]), trace))
    info = dict(tile=tr.tile, seconds_loop=tr.seconds_loop, autotune_seconds=tuple(tr.autotune_seconds))
    return solutions, info


class HipBundleAdjuster:
    """BundleAdjuster (solver.rs:1562-1573) backed by the MI355X library.

    ``solve`` returns a Solution or raises the SolveError subclass the library reports;
    Unsupported is the sanctioned decline a caller may route to another adjuster.
    """

    def __init__(self, device: int = 0):
        from ._lib import Handle

        self._handle = Handle(device)

    def name(self) -> str:
        return self._handle.lib.ba_hip_name().decode()

    def solve(self, problem: BundleProblem, config: SolveConfig | None = None, trace_capacity=None) -> Solution:
        config = config or SolveConfig()
        return solve_with(self._handle.lib, "ba_hip_solve", problem, config, self._handle.ptr,
                          trace_capacity=trace_capacity)

    def prepare(self, problem: BundleProblem) -> NativeProblem:
        return prepare_with(self._handle.lib, "ba_hip_prepare", problem)

    def plan(self, problem: BundleProblem, config: SolveConfig | None = None) -> dict:
        """The camera step `solve` would take (host only)."""
        return plan_with(self._handle.lib, "ba_hip_plan", problem, config or SolveConfig())

    def solve_batched(self, problems: list, config: SolveConfig | None = None, mode: int = A.BATCH_GRAPH,
                      trace_capacity=None):
        """Many independent windows in one batched solve (direct step): (solutions, info)."""
        config = config or SolveConfig(linear_solver=A.SOLVER_DIRECT)
        return solve_batched_with(self._handle.lib, self._handle.ptr, problems, config, mode, trace_capacity)

    @property
    def handle(self):
        return self._handle
