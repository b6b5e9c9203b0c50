"""C ABI and Python boundary checks that need no GPU: the library loads, exports every entry
point declared in include/pvfmm.h, reports errors through PVFMMGetLastError instead of
aborting, and the Python wrapper validates like python/src/pvfmm/wrapper.py:121-165."""
import ctypes
import os
import re

import numpy as np
import pytest

import pvfmm_amd as P
from pvfmm_amd import ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pvfmm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:PVFMM|pvfmm_amd_|pvfmm)\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ffi.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(ffi.EXPORTED) <= set(names)


def test_error_codes_without_abort():
    L = ffi.lib()
    # unsupported kernel (StokesPressure) -> NULL and PVFMM_ERR_UNSUPPORTED
    assert not L.PVFMMCreateContextD(-1.0, 100, 6, 2, 0)
    assert L.PVFMMGetLastError() == 2
    assert L.PVFMMErrorString(2) == b"unsupported feature"
    # a periodic box (box_size > 0) is a valid context
    c = L.PVFMMCreateContextD(1.0, 100, 6, 0, 0)
    assert c and L.PVFMMGetLastError() == 0
    L.PVFMMDestroyContextD(ctypes.byref(ctypes.c_void_p(c)))
    # odd multipole order
    assert not L.pvfmm_amd_fmm_create(0, 5, 0, 0.0)
    assert L.PVFMMGetLastError() == 1
    # evaluation before setup is a state error, not a crash
    x = np.random.default_rng(0).random((10, 3))
    t = P.PtFMMTree(x, x, 4)
    out = np.zeros(10)
    assert L.pvfmm_amd_tree_evaluate(t._h, ffi.dptr(np.ones(10)), ffi.dptr(out), 1) == 5
    # tuning and batching need a set-up tree: state errors, no HIP work
    p2 = (ctypes.c_int32 * 2)(16, 512)
    assert L.pvfmm_amd_tree_set_tuning(t._h, p2, 2) == 5
    assert L.pvfmm_amd_tree_get_tuning(t._h, p2, 2) == 5
    arr = (ctypes.c_void_p * 1)(t._h)
    assert not L.pvfmm_amd_batch_create(arr, 1)
    assert L.PVFMMGetLastError() == 5
    assert not L.pvfmm_amd_batch_create(arr, 0)
    assert L.PVFMMGetLastError() == 1
    # destroy on NULL is a no-op and sets the handle to NULL
    p = ctypes.c_void_p(None)
    L.PVFMMDestroyContextD(ctypes.byref(p))
    assert p.value is None


def test_fortran_context_create_and_errors():
    """Fortran-callable create/destroy (host work only: operators, no evaluation)."""
    L = ffi.lib()
    i32 = ctypes.c_int32
    ctx = ctypes.c_void_p()
    box = ctypes.c_double(-1.0)
    L.pvfmmcreatecontextd_(ctypes.byref(ctx), ctypes.byref(box), ctypes.byref(i32(100)), ctypes.byref(i32(4)),
                           ctypes.byref(i32(0)), ctypes.byref(i32(0)))
    assert L.PVFMMGetLastError() == 0 and ctx.value
    L.pvfmmdestroycontextd_(ctypes.byref(ctx))
    assert ctx.value is None
    # StokesPressure is unsupported: NULL context and the error code, no abort
    L.pvfmmcreatecontextd_(ctypes.byref(ctx), ctypes.byref(box), ctypes.byref(i32(100)), ctypes.byref(i32(4)),
                           ctypes.byref(i32(2)), ctypes.byref(i32(0)))
    assert ctx.value is None and L.PVFMMGetLastError() == 2
    # evaluation on a NULL context is an argument error
    z = ctypes.c_int64(1)
    L.pvfmmevald_(None, None, ctypes.byref(z), None, None, ctypes.byref(z), ctypes.byref(ctx), ctypes.byref(i32(1)))
    assert L.PVFMMGetLastError() == 1


def test_list_counts_match_reference_tables():
    L = ffi.lib()
    # include/interac_list.txx:39-56 with the Mat_Type numbering of include/precomp_mat.hpp:21-39
    # (U0 8, U1 9, U2 10, V 11, W 12, X 13, V1 14): U0 56, U1 27, U2 56, V 316, W 152, X 152, V1 26
    assert [L.pvfmm_amd_list_count(t) for t in (8, 9, 10, 11, 12, 13, 14)] == [56, 27, 56, 316, 152, 152, 26]


def test_python_context_validation():
    ctx = None
    with pytest.raises(ValueError):
        P.FMMParticleContext(-1.0, 100, 5, P.FMMKernel.LaplacePotential)
    with pytest.raises(ffi.PVFMMError):
        P.FMMParticleContext(-1.0, 100, 6, P.FMMKernel.StokesPressure)
    ctx = P.FMMParticleContext(-1.0, 64, 4, P.FMMKernel.LaplacePotential)
    src = np.random.default_rng(0).random(30)
    with pytest.raises(ValueError):
        ctx.evaluate(src.astype(np.float32), src, None, src)
    with pytest.raises(ValueError):
        ctx.evaluate(src[:29], src[:29], None, src)
    with pytest.raises(ValueError):
        ctx.evaluate(src, src[:27], None, src)
    with pytest.raises(ValueError):
        ctx.evaluate(src, src, src, src)  # dl_den must have 2x the length
    with pytest.raises(ValueError):
        ctx.evaluate(src, src, None, src[:28])


def test_kernel_dims_mirror_reference():
    assert P.KERNEL_DIMS[P.FMMKernel.LaplaceGradient] == (1, 3)
    assert P.KERNEL_DIMS[P.FMMKernel.StokesVelocity] == (3, 3)
    assert P.FMMKernel(3) is P.FMMKernel.StokesVelocity
