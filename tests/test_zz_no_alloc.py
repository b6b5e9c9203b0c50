"""pvfmm_amd_direct_device_noalloc: the FP64 direct sum on device arrays with its partial sums
in the library's static device scratch (no device allocation).  Its results must equal those
of pvfmm_amd_direct (the same kernels with an allocated partial-sum buffer) bit for bit, on two
calls in a row on one device, the second with a smaller input, so that the second call's
result cannot lean on what the first left in the shared scratch.  The first result is also
checked against the oracle's FP64 direct sum."""
import ctypes

import numpy as np
import pytest

import oracle as O
import pvfmm_amd as P
from pvfmm_amd import ffi

pytestmark = pytest.mark.gpu


def _noalloc(gpu, kernel, src, den, trg):
    import torch

    k1 = 1 if kernel == 0 else 3  # C ids: 0 laplace_poten, 1 laplace_grad, 3 stokes_vel
    ds = torch.from_numpy(np.ascontiguousarray(src)).to(gpu)
    dd = torch.from_numpy(np.ascontiguousarray(den)).to(gpu)
    dt = torch.from_numpy(np.ascontiguousarray(trg)).to(gpu)
    out = torch.full((k1 * len(trg),), float("nan"), dtype=torch.float64, device=gpu)
    stream = torch.cuda.current_stream(gpu)
    ffi.check(ffi.lib().pvfmm_amd_direct_device_noalloc(kernel, ctypes.c_void_p(ds.data_ptr()), len(src),
                                                        ctypes.c_void_p(dd.data_ptr()), ctypes.c_void_p(dt.data_ptr()),
                                                        len(trg), ctypes.c_void_p(out.data_ptr()),
                                                        ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return out.cpu().numpy()


def test_noalloc_direct_matches_allocating_entry_point(gpu):
    rng = np.random.default_rng(77)
    # laplace_grad: 200,000 sources onto 4,096 targets splits the sources into 112 chunks
    # (11 MB of partials); then a smaller Stokes call (3 density components) of 30,000 sources
    # onto 1,000 targets (118 chunks, 2.8 MB)
    cases = [(1, 200_000, 4096), (3, 30_000, 1000)]
    for i, (kernel, n_src, n_trg) in enumerate(cases):
        k0 = 3 if kernel == 3 else 1
        src = rng.random((n_src, 3))
        den = rng.standard_normal((n_src, k0))
        trg = rng.random((n_trg, 3))
        print(f"noalloc direct call {i}: kernel {kernel}, {n_src} sources, {n_trg} targets", flush=True)
        got = _noalloc(gpu, kernel, src, den, trg)
        ref = P.direct(kernel, src, den, trg).ravel()
        assert np.array_equal(got, ref), (i, float(np.abs(got - ref).max()))
        if i == 0:
            exact = O.direct("laplace_grad", src, den.ravel(), trg)
            e = float(np.linalg.norm(got - exact) / np.linalg.norm(exact))
            print(f"noalloc direct vs oracle direct sum: {e:.2e}", flush=True)
            assert e < 1e-12


def test_noalloc_direct_no_sources_gives_zeros(gpu):
    rng = np.random.default_rng(78)
    trg = rng.random((300, 3))
    got = _noalloc(gpu, 0, np.zeros((0, 3)), np.zeros(0), trg)
    assert np.array_equal(got, np.zeros(300))
