"""Loader of libba_hip.so, the HIP library this package drives.

There is no CPU fallback: if the library is missing or cannot load, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import pathlib

from . import _abi as A

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libba_hip.so"

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load() -> C.CDLL:
    """The loaded HIP library (raises LibraryMissing when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LibraryMissing(f"{LIB_PATH} is not built; run `make -C openmvg-rs_amd` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    lib.ba_hip_name.restype = C.c_char_p
    lib.ba_hip_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.ba_hip_destroy.argtypes = [vp]
    lib.ba_hip_destroy.restype = None
    lib.ba_hip_solve.argtypes = [vp] + [vp] * 5 + [C.c_char_p, C.c_size_t]
    lib.ba_hip_prepare.argtypes = [vp, vp, vp, C.c_char_p, C.c_size_t]
    lib.ba_hip_plan.argtypes = [vp, vp, vp, C.c_char_p, C.c_size_t]
    lib.ba_hip_linearize.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.ba_hip_schur.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp, vp]
    lib.ba_hip_pcg.argtypes = [vp, vp, C.c_double, C.c_double, C.c_uint32, vp, vp, vp, vp]
    lib.ba_hip_cholesky_solve.argtypes = [vp, C.c_size_t, vp, vp, vp, C.c_int32]
    lib.ba_hip_cholesky_solve_noalloc.argtypes = [vp, C.c_size_t, vp, vp, vp, C.c_int32]
    lib.ba_hip_set_schur_product.argtypes = [vp, C.c_int32]
    lib.ba_hip_last_schur_product.argtypes = [vp]
    lib.ba_hip_last_schur_product.restype = C.c_int32
    sz = C.c_size_t
    lib.ba_hip_reprojection_topology_create.argtypes = [vp, sz, sz, sz, vp, vp, C.POINTER(vp), C.c_char_p, sz]
    lib.ba_hip_reprojection_topology_destroy.argtypes = [vp]
    lib.ba_hip_reprojection_topology_destroy.restype = None
    lib.ba_hip_solve_batched.argtypes = [vp, sz, vp, vp, vp, vp, vp, C.c_int32, C.c_char_p, sz]
    lib.ba_hip_cost_grad.argtypes = [vp, vp, C.c_int32, vp, vp, vp, C.c_int32, C.c_double, vp, vp, vp, vp, vp, vp,
                                     C.c_char_p, sz]
    _lib = lib
    return lib


class Handle:
    """A ba_hip_handle bound to one HIP device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        code = self.lib.ba_hip_create(device, C.byref(h))
        if code != A.BA_OK:
            raise RuntimeError(f"ba_hip_create({device}) failed with code {code}: no usable HIP device")
        self.ptr = h

    def close(self):
        if self.ptr:
            self.lib.ba_hip_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
