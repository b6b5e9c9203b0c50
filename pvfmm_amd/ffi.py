"""ctypes bindings of libpvfmm (the C ABI declared in include/pvfmm.h).

Mirrors python/src/pvfmm/ffi.py of the reference: the library is found through the ``PVFMM``
environment variable (a directory holding libpvfmm.so, python/src/pvfmm/ffi.py:51-63) and
otherwise taken from this package directory, where ``__graft_entry__.build()`` /
``make -C pvfmm_amd/csrc`` put it.  mpi4py is not required: the communicator argument is a
pointer-sized integer that the library ignores (one GPU, the MPI_COMM_SELF share).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    custom = os.getenv("PVFMM")
    if custom:
        return os.path.join(custom, "libpvfmm.so")
    return os.path.join(_HERE, "libpvfmm.so")


class LibraryNotBuilt(ImportError):
    pass


_LIB = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
PVFMMKernel = ctypes.c_uint  # enum
Comm = ctypes.c_ssize_t


def _declare(L):
    vp, i, i64, l, d, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_long, ctypes.c_double, ctypes.c_float
    sig = {
        "PVFMMCreateContextD": (vp, [d, i, i, PVFMMKernel, Comm]),
        "PVFMMCreateContextF": (vp, [f, i, i, PVFMMKernel, Comm]),
        "PVFMMCreateContextDWorld": (vp, [d, i, i, PVFMMKernel]),
        "PVFMMCreateContextFWorld": (vp, [f, i, i, PVFMMKernel]),
        "PVFMMEvalD": (None, [c_double_p, c_double_p, c_double_p, l, c_double_p, c_double_p, l, vp, i]),
        "PVFMMEvalF": (None, [c_float_p, c_float_p, c_float_p, l, c_float_p, c_float_p, l, vp, i]),
        "PVFMMDestroyContextD": (None, [ctypes.POINTER(vp)]),
        "PVFMMDestroyContextF": (None, [ctypes.POINTER(vp)]),
        "pvfmmcreatecontextd_": (None, [ctypes.POINTER(vp), c_double_p, c_int32_p, c_int32_p, c_int32_p, c_int32_p]),
        "pvfmmcreatecontextf_": (None, [ctypes.POINTER(vp), c_float_p, c_int32_p, c_int32_p, c_int32_p, c_int32_p]),
        "pvfmmevald_": (None, [c_double_p, c_double_p, c_int64_p, c_double_p, c_double_p, c_int64_p, ctypes.POINTER(vp),
                               c_int32_p]),
        "pvfmmevalf_": (None, [c_float_p, c_float_p, c_int64_p, c_float_p, c_float_p, c_int64_p, ctypes.POINTER(vp),
                               c_int32_p]),
        "pvfmmdestroycontextd_": (None, [ctypes.POINTER(vp)]),
        "pvfmmdestroycontextf_": (None, [ctypes.POINTER(vp)]),
        "PVFMMGetLastError": (i, []),
        "PVFMMErrorString": (ctypes.c_char_p, [i]),
        "PVFMMGetLastErrorMessage": (ctypes.c_char_p, []),
        "PVFMMSetMixedPrecision": (i, [vp, i]),
        "pvfmm_amd_fmm_create": (vp, [i, i, i, d]),
        "pvfmm_amd_fmm_destroy": (None, [vp]),
        "pvfmm_amd_tree_create": (vp, [c_double_p, i64, c_double_p, i64, i]),
        "pvfmm_amd_tree_create_pts": (vp, [c_double_p, i64, c_double_p, i64, c_double_p, i64, i]),
        "pvfmm_amd_tree_create_bndry": (vp, [c_double_p, i64, c_double_p, i64, c_double_p, i64, i, i]),
        "pvfmm_amd_tree_destroy": (None, [vp]),
        "pvfmm_amd_tree_setup": (i, [vp, vp, i]),
        "pvfmm_amd_tree_evaluate": (i, [vp, c_double_p, c_double_p, i]),
        "pvfmm_amd_tree_evaluate_device": (i, [vp, vp, vp, i, i]),
        "pvfmm_amd_tree_evaluate_adjoint": (i, [vp, c_double_p, c_double_p, i]),
        "pvfmm_amd_tree_evaluate_adjoint_device": (i, [vp, vp, vp, i, i]),
        "pvfmm_amd_tree_stream": (vp, [vp]),
        "pvfmm_amd_tree_autotune": (i, [vp, vp, vp, i, c_int32_p, i]),
        "pvfmm_amd_tree_get_tuning": (i, [vp, c_int32_p, i]),
        "pvfmm_amd_tree_set_tuning": (i, [vp, c_int32_p, i]),
        "pvfmm_amd_batch_create": (vp, [ctypes.POINTER(vp), i]),
        "pvfmm_amd_batch_evaluate_device": (i, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), i]),
        "pvfmm_amd_batch_stream": (vp, [vp]),
        "pvfmm_amd_batch_records": (i, [vp]),
        "pvfmm_amd_batch_destroy": (None, [vp]),
        "pvfmm_amd_tree_num_nodes": (i64, [vp]),
        "pvfmm_amd_tree_nodes": (i, [vp, c_int64_p]),
        "pvfmm_amd_tree_point_order": (i, [vp, i, c_int64_p]),
        "pvfmm_amd_list_count": (i, [i]),
        "pvfmm_amd_tree_lists": (i, [vp, i, c_int32_p]),
        "pvfmm_amd_tree_set_timing": (i, [vp, i]),
        "pvfmm_amd_tree_phase_times": (i, [vp, c_double_p, i]),
        "pvfmm_amd_phase_name": (ctypes.c_char_p, [i]),
        "pvfmm_amd_tree_stats": (i, [vp, c_int64_p, i]),
        "pvfmm_amd_fmm_operator": (i64, [vp, i, c_double_p]),
        "pvfmm_amd_device_count": (i, []),
        "pvfmm_amd_direct": (i, [i, c_double_p, i64, c_double_p, c_double_p, i64, c_double_p]),
        "pvfmm_amd_direct_device_noalloc": (i, [i, vp, i64, vp, vp, i64, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return sorted(sig)


EXPORTED = None


def lib():
    """Load libpvfmm.so; raises LibraryNotBuilt when it is missing (no fallback path exists)."""
    global _LIB, EXPORTED
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise LibraryNotBuilt(
                f"libpvfmm.so not found at {path}; build it with `make -C pvfmm_amd/csrc` "
                "or __graft_entry__.build()")
        L = ctypes.CDLL(path)
        EXPORTED = _declare(L)
        _LIB = L
    return _LIB


class PVFMMError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"pvfmm error {code}: {message}")
        self.code = code


def check(code: int | None = None):
    L = lib()
    c = L.PVFMMGetLastError() if code is None else code
    if c:
        msg = L.PVFMMGetLastErrorMessage().decode(errors="replace")
        raise PVFMMError(c, msg)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(c_double_p) if a is not None else None


def fptr(a: np.ndarray):
    return a.ctypes.data_as(c_float_p) if a is not None else None
