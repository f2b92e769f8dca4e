"""Loader of the CPU oracle (oracle/_build/libba_oracle.so) for tests and the CPU baseline.

The oracle is the checker: tests compare the HIP path against it; nothing in the product
imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "_build" / "libba_oracle.so"


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


def load_oracle() -> C.CDLL:
    if not ORACLE_SO.exists():
        build_oracle()
    lib = C.CDLL(str(ORACLE_SO))
    d = C.c_double
    pd = C.POINTER(C.c_double)
    lib.bao_loss_cost.restype = d
    lib.bao_loss_cost.argtypes = [C.c_int32, d, d]
    lib.bao_loss_sqrt_weight.restype = d
    lib.bao_loss_sqrt_weight.argtypes = [C.c_int32, d, d]
    lib.bao_first_positive_radial_turn.restype = d
    lib.bao_point_block_rcond.restype = d
    lib.bao_native_cost.restype = d
    lib.bao_nielsen_accept.argtypes = [pd, pd, d, d, pd]
    lib.bao_nielsen_reject.argtypes = [pd, pd]
    lib.bao_cholesky_solve.argtypes = [C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.bao_assemble.argtypes = [C.c_void_p, d] + [C.c_void_p] * 7 + [C.c_void_p]
    lib.bao_pcg.argtypes = [C.c_void_p, d, d, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.bao_track_geometry.argtypes = [C.c_size_t, C.c_void_p, d]
    lib.bao_exp_so3.argtypes = [C.c_void_p, C.c_void_p]
    lib.bao_radial_projection.argtypes = [C.c_void_p] * 5
    lib.bao_radial_linearization.argtypes = [C.c_void_p] * 6
    lib.bao_cost_grad.argtypes = [C.c_size_t] * 3 + [C.c_void_p] * 5 + [C.c_int32, d] + [C.c_void_p] * 4
    lib.bao_point_block_factor.argtypes = [C.c_void_p, d, C.c_void_p]
    lib.bao_point_block_rcond.argtypes = [C.c_void_p]
    return lib
