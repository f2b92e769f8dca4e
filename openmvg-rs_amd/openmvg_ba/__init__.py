"""openmvg_ba — MI355X (gfx950) Levenberg–Marquardt bundle adjustment behind openmvg-ba's
BundleAdjuster::solve contract (crates/openmvg-ba/src/solver.rs:1562-1573).

The compute path is the HIP library built from ``csrc/`` (``libba_hip.so``); this package is its
host-side mirror of the reference's problem/solver types and a ctypes binding.
"""

from . import _abi
from .problem import (
    HipBundleAdjuster,
    BundleProblem,
    Internal,
    Intrinsics,
    InvalidConfig,
    Numerical,
    ProblemError,
    ResourceExhausted,
    Solution,
    SolveConfig,
    SolveError,
    SolveReport,
    Unsupported,
)

__all__ = [
    "HipBundleAdjuster", "BundleProblem", "Intrinsics", "SolveConfig", "SolveReport", "Solution", "SolveError",
    "InvalidConfig", "Unsupported", "ResourceExhausted", "Numerical", "Internal", "ProblemError",
    "_abi",
]
