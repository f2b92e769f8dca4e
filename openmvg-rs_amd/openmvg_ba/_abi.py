"""ctypes mirror of include/ba_hip.h.

Every structure here has the field order of its C declaration; ``tests/test_abi.py`` checks the
sizes against the C compiler's.  Nothing in this module touches a GPU.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

BA_OK = 0
BA_ERR_INVALID_CONFIG = 1
BA_ERR_UNSUPPORTED = 2
BA_ERR_RESOURCE_EXHAUSTED = 3
BA_ERR_NUMERICAL = 4
BA_ERR_INTERNAL = 5
BA_ERR_INVALID_PROBLEM = 6

CAMERA_PINHOLE = 0
CAMERA_BROWN_NONE = 1
CAMERA_BROWN_K1 = 2
CAMERA_BROWN_K1K2 = 3
CAMERA_BROWN_K1K2K3 = 4
CAMERA_FISHEYE = 5
CAMERA_SPHERICAL = 6

RESIDUAL_PIXELS = 0
RESIDUAL_ANGULAR = 1
LOSS_SQUARED = 0
LOSS_HUBER = 1
LOSS_CAUCHY = 2
GAUGE_FREE = 0
GAUGE_FIXED_BASELINE = 1

SOLVER_AUTO = 0
SOLVER_DIRECT = 1
SOLVER_ITERATIVE = 2
PRECISION_FP64 = 0
PRECISION_MIXED = 1

PROFILE_NONE = 0
PROFILE_LINEARIZE = 1
PROFILE_POINT_ELIM = 2
PROFILE_PCG_POINTS = 3
PROFILE_PCG_CAMERAS = 4
PROFILE_PAIRS = 5
PROFILE_CHOLESKY = 6
PROFILE_CANDIDATE = 7
PROFILE_SCHUR_PAIRS = 8
PROFILE_PCG_PRODUCT = 9

# ba_schur_product
SCHUR_AUTO = 0
SCHUR_IMPLICIT = 1
SCHUR_EXPLICIT = 2

STATUS_CONVERGED_GRADIENT = 0
STATUS_CONVERGED_STEP = 1
STATUS_CONVERGED_COST = 2
STATUS_MAX_ITERATIONS = 3
STATUS_STALLED = 4
STATUS_DEGENERATE = 5
STATUS_SKIPPED = 6
CERTIFY_FULL = 0
CERTIFY_LOOP_ONLY = 1
OBSERVABILITY_FULL_RANK = 0
OBSERVABILITY_ONLY_EXPECTED_GAUGE = 1
OBSERVABILITY_RANK_DEFICIENT = 2
OBSERVABILITY_NOT_ASSESSED = 3
STATUS_NAMES = {
    0: "converged(gradient)",
    1: "converged(step)",
    2: "converged(cost)",
    3: "max_iterations",
    4: "stalled",
    5: "degenerate",
    6: "skipped",
}

_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)


class Intrinsic(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("k1", C.c_double), ("k2", C.c_double), ("k3", C.c_double),
        ("p1", C.c_double), ("p2", C.c_double),
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("fixed_calibration", C.c_uint8),
        ("var_focal", C.c_uint8), ("var_principal_point", C.c_uint8),
        ("var_radial", C.c_uint8), ("var_tangential", C.c_uint8),
        ("certification_waived", C.c_uint8),
        ("_pad", C.c_uint8 * 2),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("n_poses", C.c_size_t),
        ("pose_rotation", _d), ("pose_center", _d),
        ("pose_var_rotation", _u8), ("pose_var_translation", _u8),
        ("n_intrinsics", C.c_size_t), ("intrinsics", C.POINTER(Intrinsic)),
        ("n_points", C.c_size_t), ("point_xyz", _d), ("point_fixed", _u8),
        ("n_observations", C.c_size_t),
        ("obs_pose", _u32), ("obs_intrinsic", _u32), ("obs_point", _u32), ("obs_pixel", _d),
        ("residual_space", C.c_int32), ("robust_loss", C.c_int32), ("loss_scale", C.c_double),
        ("gauge", C.c_int32), ("gauge_first", C.c_uint32), ("gauge_second", C.c_uint32),
        ("gauge_distance", C.c_double),
    ]


class SolveConfigC(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint32),
        ("function_tolerance", C.c_double),
        ("gradient_tolerance", C.c_double),
        ("parameter_tolerance", C.c_double),
        ("linear_solver", C.c_int32),
        ("workspace_budget_bytes", C.c_uint64),
        ("precision", C.c_int32),
        ("pcg_max_iterations", C.c_uint32),
        ("pcg_relative_tolerance", C.c_double),
        ("profile_kernel", C.c_int32),
        ("certification", C.c_int32),
    ]


class SolveReportC(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("linear_solver_used", C.c_int32),
        ("iterations", C.c_uint32), ("accepted_steps", C.c_uint32), ("rejected_steps", C.c_uint32),
        ("rejected_linearization", C.c_uint32), ("rejected_step_solve", C.c_uint32),
        ("rejected_candidate", C.c_uint32), ("rejected_insufficient_reduction", C.c_uint32),
        ("initial_cost", C.c_double), ("final_cost", C.c_double),
        ("initial_rmse", C.c_double), ("final_rmse", C.c_double),
        ("gradient_inf_norm", C.c_double), ("step_norm", C.c_double),
        ("initial_damping", C.c_double), ("final_damping", C.c_double),
        ("last_gain_ratio", C.c_double),
        ("reduced_dimension", C.c_uint32), ("weak_point_count", C.c_uint32),
        ("minimum_point_rcond", C.c_double),
        ("pcg_iterations", C.c_uint32), ("quarantined_points", C.c_uint32),
        ("maximum_pcg_relative_residual", C.c_double),
        ("normalization_scale", C.c_double),
        ("observation_count", C.c_uint64), ("point_count", C.c_uint64), ("pose_count", C.c_uint64),
        ("seconds_setup", C.c_double), ("seconds_iterations", C.c_double), ("seconds_final", C.c_double),
        ("profile_launches", C.c_uint32), ("profile_seconds", C.c_double),
        ("precision_fallbacks", C.c_uint32), ("refinement_residual", C.c_double),
        ("plan_reason", C.c_int32), ("workspace_bytes", C.c_uint64),
        ("inlier_count", C.c_uint64), ("final_inlier_rmse", C.c_double),
        ("observability", C.c_int32), ("pose_rank", C.c_uint32), ("pose_dimension", C.c_uint32),
        ("unconstrained_intrinsics", C.c_uint32), ("weak_point_resolves", C.c_uint32),
        ("_pad_certify", C.c_uint32),
    ]


# ba_plan_reason (SparseExecutionReason)
PLAN_EXPLICIT_SOLVER = 0
PLAN_DIRECT_FITS_BUDGET = 1
PLAN_DIRECT_EXCEEDS_BUDGET = 2
PLAN_LOWEST_MEMORY = 3
PLAN_WORKLOAD_PREFERS_RETAINED = 4


class ExecutionPlanC(C.Structure):
    _fields_ = [
        ("linear_solver", C.c_int32), ("reason", C.c_int32),
        ("budget_bytes", C.c_uint64), ("estimated_direct_bytes", C.c_uint64),
        ("estimated_retained_bytes", C.c_uint64), ("reduced_nonzeros", C.c_uint64),
        ("reduced_dimension", C.c_uint32), ("_pad", C.c_uint32),
    ]


class IterationTrace(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint32), ("count", C.c_uint32),
        ("cost", _d), ("damping", _d), ("candidate_cost", _d), ("predicted_reduction", _d),
        ("gain_ratio", _d), ("accepted", _u8), ("pcg_iterations", _u32), ("seconds", _d),
    ]


class SolutionC(C.Structure):
    _fields_ = [
        ("pose_rotation", _d), ("pose_center", _d),
        ("intrinsic_focal", _d), ("intrinsic_k1", _d), ("intrinsic_k2", _d),
        ("point_xyz", _d), ("point_quarantined", _u8), ("intrinsic_unconstrained", _u8),
        ("intrinsic_covariance", _d),
    ]


class NativeSizes(C.Structure):
    _fields_ = [
        ("n_cameras", C.c_size_t), ("n_groups", C.c_size_t), ("n_points", C.c_size_t),
        ("n_observations", C.c_size_t), ("dimension", C.c_size_t),
        ("n_point_parameters", C.c_size_t), ("n_pattern_entries", C.c_size_t),
        ("n_quarantined", C.c_size_t),
    ]


class NativeProblem(C.Structure):
    _fields_ = [
        ("camera_rotation", _d), ("camera_translation", _d), ("group_params", _d),
        ("camera_group", _u32), ("camera_source", _u32), ("point_xyz", _d), ("point_source", _u32),
        ("obs_camera", _u32), ("obs_source", _u32), ("obs_uv", _d), ("point_offsets", _u32),
        ("camera_fixed", _u8), ("point_fixed", _u8),
        ("camera_offset", _u32), ("camera_dim", _u32), ("group_offset", _u32), ("group_dim", _u32),
        ("cross_offsets", _u32), ("point_parameters", _u32),
        ("pattern_col_ptr", _u32), ("pattern_row_idx", _u32),
        ("gauge_camera", C.c_int32), ("_pad", C.c_int32),
        ("scene_scale", C.c_double),
    ]


class DeviceProblem(C.Structure):
    _fields_ = [
        ("n_cameras", C.c_size_t), ("n_groups", C.c_size_t), ("n_points", C.c_size_t),
        ("n_observations", C.c_size_t), ("dimension", C.c_size_t),
        ("camera_rotation", C.c_void_p), ("camera_translation", C.c_void_p),
        ("group_params", C.c_void_p), ("camera_group", C.c_void_p),
        ("point_xyz", C.c_void_p), ("obs_camera", C.c_void_p), ("obs_point", C.c_void_p),
        ("obs_uv", C.c_void_p), ("point_offsets", C.c_void_p), ("point_fixed", C.c_void_p),
        ("camera_offset", C.c_void_p), ("camera_dim", C.c_void_p),
        ("group_offset", C.c_void_p), ("group_dim", C.c_void_p),
        ("loss", C.c_int32), ("loss_scale", C.c_double),
        ("gauge_camera", C.c_int32), ("scene_scale", C.c_double), ("baseline", C.c_double),
        ("basis", C.c_double * 6),
    ]


class SchurScalars(C.Structure):
    _fields_ = [
        ("cost", C.c_double), ("squared_residual", C.c_double),
        ("point_model_reduction", C.c_double), ("scaled_gradient_inf_norm", C.c_double),
        ("minimum_point_rcond", C.c_double),
        ("weak_point_count", C.c_uint32), ("valid", C.c_uint32),
    ]


def ptr(array: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array's data (NULL for None)."""
    if array is None:
        return C.cast(None, C.POINTER(ctype))
    assert array.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return array.ctypes.data_as(C.POINTER(ctype))


def vptr(array) -> int | None:
    """Raw address of a numpy array or a torch tensor (host or device), None for None."""
    if array is None:
        return None
    if isinstance(array, np.ndarray):
        assert array.flags["C_CONTIGUOUS"]
        return array.ctypes.data
    assert array.is_contiguous(), "tensor must be contiguous"
    return array.data_ptr()


def message_buffer(size: int = 512):
    return C.create_string_buffer(size)


BATCH_GRAPH = 0
BATCH_STREAM = 1


class BatchTrace(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint32), ("cost", C.POINTER(C.c_double)), ("damping", C.POINTER(C.c_double)),
        ("candidate_cost", C.POINTER(C.c_double)), ("accepted", C.POINTER(C.c_uint8)),
        ("tile", C.c_uint32), ("seconds_loop", C.c_double), ("autotune_seconds", C.c_double * 2),
    ]

