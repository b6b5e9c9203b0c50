"""GPU parity: the HIP evaluator (through the C ABI) against the oracle on the same tree and
m, and against an FP64 direct sum.  Tolerances are stated per test; the FP64 evaluator must
match the oracle to rounding amplified by the check-to-equivalent conditioning (~1e-7 at
m=6), the mixed evaluator to the FP32 bar."""
import ctypes

import numpy as np
import pytest

import oracle as O
import pvfmm_amd as P
from pvfmm_amd import ffi
from pvfmm_amd import workloads as W

pytestmark = pytest.mark.gpu

KID = {"laplace_poten": 0, "laplace_grad": 1, "stokes_vel": 2}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def run_product(src, trg, q, kernel, m, max_pts, prec="fp64"):
    t = P.PtFMMTree(src, trg, max_pts)
    t.setup(P.PtFMM(kernel, m, prec))
    return t.evaluate(q), t


def run_oracle(src, trg, q, kernel, m, max_pts, real="fp64"):
    """The oracle on the same tree; real="fp32" takes the pseudo-inverse cutoff of the
    reference's single-precision build (4 eps_32 max_S), which the mixed evaluator uses."""
    ot = O.OracleTree(src, trg, max_pts)
    ot.set_operators(O.precompute(kernel, m, real))
    return ot.evaluate(q)


CASES = [
    ("laplace_poten", "uniform", 6, 64),
    ("laplace_grad", "sphere", 6, 48),
    ("stokes_vel", "uniform", 6, 64),
    ("laplace_poten", "blob", 8, 32),
]


def points(kind, n, rng):
    if kind == "uniform":
        return rng.random((n, 3))
    if kind == "sphere":
        return W.sphere_surface(rng, n)
    return W.gaussian_blob(rng, n, sigma=0.05)


@pytest.mark.parametrize("kernel,kind,m,max_pts", CASES)
def test_fp64_matches_oracle(gpu, kernel, kind, m, max_pts):
    rng = np.random.default_rng(11)
    n = 6000
    x = points(kind, n, rng)
    k0, k1 = O.KDIM[KID[kernel]]
    q = rng.standard_normal(k0 * n)
    u, _ = run_product(x, x, q, kernel, m, max_pts)
    uo = run_oracle(x, x, q, kernel, m, max_pts)
    idx = rng.choice(n, 400, replace=False)
    ud = O.direct(kernel, x, q, x[idx])
    e_or = rel(u, uo)
    e_dir = rel(u.reshape(n, k1)[idx].ravel(), ud)
    e_or_dir = rel(uo.reshape(n, k1)[idx].ravel(), ud)
    print(f"{kernel} {kind} m={m}: vs oracle {e_or:.2e}, vs direct {e_dir:.2e} (oracle vs direct {e_or_dir:.2e})")
    assert e_or < 1e-10  # FP64, same tree/lists/m; ill-conditioned C2E amplifies rounding
    assert abs(e_dir - e_or_dir) < 1e-3 * e_or_dir


@pytest.mark.parametrize("kernel,kind,m,max_pts", CASES[:3])
def test_mixed_matches_oracle(gpu, kernel, kind, m, max_pts):
    rng = np.random.default_rng(12)
    n = 6000
    x = points(kind, n, rng)
    k0, k1 = O.KDIM[KID[kernel]]
    q = rng.standard_normal(k0 * n)
    u, _ = run_product(x, x, q, kernel, m, max_pts, "mixed")
    uo = run_oracle(x, x, q, kernel, m, max_pts, real="fp32")
    e = rel(u, uo)
    print(f"mixed {kernel}: vs oracle (FP32 cutoff) {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("m", [2, 4, 12, 14])
def test_other_orders_match_oracle(gpu, m):
    """Multipole orders outside the headline set: m = 4 and 12 use the compile-time DFTs,
    m = 2 and 14 the generic ones.  FP64 against the oracle at the same m (m = 14 against the
    direct sum only: its oracle precompute is slow), mixed against the FP32 bar."""
    rng = np.random.default_rng(40 + m)
    n = 3000
    x = rng.random((n, 3))
    q = rng.standard_normal(n)
    idx = rng.choice(n, 300, replace=False)
    ud = O.direct("laplace_poten", x, q, x[idx])
    u, _ = run_product(x, x, q, "laplace_poten", m, 64)
    e_dir = rel(u[idx], ud)
    if m <= 12:
        uo = run_oracle(x, x, q, "laplace_poten", m, 64)
        e_or = rel(u, uo)
        print(f"m={m} fp64: vs oracle {e_or:.2e}, vs direct {e_dir:.2e}")
        assert e_or < 1e-10
    else:
        print(f"m={m} fp64: vs direct {e_dir:.2e}")
    # expected accuracy of the expansion order (kernel-independent FMM, max_pts = 64)
    assert e_dir < {2: 2e-1, 4: 1e-2, 12: 1e-6, 14: 1e-7}[m]
    um, _ = run_product(x, x, q, "laplace_poten", m, 64, "mixed")
    e_mix = rel(um[idx], ud)
    print(f"m={m} mixed: vs direct {e_mix:.2e}")
    assert e_mix < max(2 * e_dir, 2e-6)


def test_m10_fp64_and_mixed_vs_direct(gpu):
    rng = np.random.default_rng(13)
    n = 30000
    x = np.concatenate([W.sphere_surface(rng, n // 2), rng.random((n - n // 2, 3))])
    q = rng.standard_normal(n)
    idx = rng.choice(n, 1000, replace=False)
    ud = O.direct("laplace_poten", x, q, x[idx])
    t = P.PtFMMTree(x, x, 128)
    for prec, tol in (("fp64", 1e-8), ("mixed", 1e-6)):
        t.setup(P.PtFMM("laplace_poten", 10, prec))
        e = rel(t.evaluate(q)[idx], ud)
        print(f"m=10 {prec}: rel-L2 vs direct {e:.2e}")
        assert e < tol


def test_deterministic(gpu):
    rng = np.random.default_rng(14)
    x = rng.random((5000, 3))
    q = rng.standard_normal(5000)
    for prec in ("fp64", "mixed"):
        t = P.PtFMMTree(x, x, 64).setup(P.PtFMM("laplace_grad", 6, prec))
        assert np.array_equal(t.evaluate(q), t.evaluate(q))


def test_linearity_and_multiple_densities(gpu):
    rng = np.random.default_rng(15)
    n = 8000
    x = rng.random((n, 3))
    q = rng.standard_normal((3, n))
    t = P.PtFMMTree(x, x, 64).setup(P.PtFMM("laplace_poten", 6), n_vec_max=3)
    u = t.evaluate(q)
    for i in range(3):
        assert np.allclose(u[i], t.evaluate(q[i]), rtol=0, atol=1e-12 * np.abs(u[i]).max())
    lin = t.evaluate(2.0 * q[0] - 0.5 * q[1])
    assert rel(lin, 2.0 * u[0] - 0.5 * u[1]) < 1e-12


@pytest.mark.parametrize("kernel,prec,tol", [("laplace_grad", "fp64", 1e-12), ("laplace_grad", "mixed", 1e-6),
                                              ("stokes_vel", "mixed", 1e-6), ("laplace_poten", "fp64", 1e-12)])
def test_batched_density_passes(gpu, kernel, prec, tol):
    """n_vec_max = 8 batches the densities: 11 vectors run as passes of 8 and 3 (P2P chunks of
    8, 2 and 1 sharing the pair geometry; one GEMM of 8 or 3 row blocks per up/down step).
    Each vector must match its own single-vector evaluation on a tree set up for one vector
    (the per-pair arithmetic is the same; a GEMM K split may differ, so to rounding)."""
    rng = np.random.default_rng(31)
    n = 6000
    x = rng.random((n, 3))
    k0, k1 = P.kernel_dims(kernel)
    q = rng.standard_normal((11, k0 * n))
    fmm = P.PtFMM(kernel, 6, prec)
    tb = P.PtFMMTree(x, x, 64).setup(fmm, n_vec_max=8)
    t1 = P.PtFMMTree(x, x, 64).setup(fmm, n_vec_max=1)
    ub = tb.evaluate(q)
    for i in range(11):
        r = rel(ub[i], t1.evaluate(q[i]))
        assert r < tol, (i, r)
    # the batched values are the FMM's values: one vector against the direct sum
    assert rel(ub[9], O.direct(KID[kernel], x, q[9], x)) < (1e-4 if prec == "mixed" else 1e-5)


def test_separate_targets_and_edge_cases(gpu):
    rng = np.random.default_rng(16)
    src = rng.random((3000, 3))
    trg = rng.random((1234, 3))
    q = rng.standard_normal(3000)
    u, _ = run_product(src, trg, q, "laplace_poten", 6, 32)
    assert rel(u, O.direct(0, src, q, trg)) < 1e-5
    # fewer points than max_pts: root-only tree, everything is near field
    s = rng.random((7, 3))
    qs = rng.standard_normal(7)
    u, _ = run_product(s, s, qs, "laplace_grad", 6, 100)
    assert rel(u, O.direct(1, s, qs, s)) < 1e-13
    # single source, single target
    u, _ = run_product(np.array([[0.2, 0.3, 0.4]]), np.array([[0.9, 0.1, 0.5]]), np.array([1.0]), "laplace_poten", 6, 4)
    assert rel(u, O.direct(0, np.array([[0.2, 0.3, 0.4]]), np.array([1.0]), np.array([[0.9, 0.1, 0.5]]))) < 1e-14
    # duplicated points: r = 0 pairs contribute nothing
    d = rng.random((500, 3))
    dd = np.concatenate([d, d])
    qd = rng.standard_normal(1000)
    u, _ = run_product(dd, dd, qd, "laplace_poten", 6, 32)
    assert rel(u, O.direct(0, dd, qd, dd)) < 1e-5


def test_c_abi_context_roundtrip(gpu):
    """PVFMMCreateContextD / PVFMMEvalD / PVFMMDestroyContextD with the reference's bounding
    box rescaling and kernel scale exponents, setup=1 then setup=0 with new densities."""
    L = ffi.lib()
    rng = np.random.default_rng(17)
    n = 5000
    src = rng.uniform(-3, 5, (n, 3))
    trg = rng.uniform(-2, 4, (2000, 3))
    # m=8 accuracy bars of each kernel (the gradient and Stokes kernels converge slower)
    for kernel, kid, k0, k1, tol in ((0, 0, 1, 1, 1e-5), (1, 1, 1, 3, 1e-4), (3, 2, 3, 3, 1e-4)):
        ctx = L.PVFMMCreateContextD(-1.0, 64, 8, kernel, 0)
        ffi.check()
        q = rng.standard_normal(k0 * n)
        out = np.zeros(k1 * len(trg))
        L.PVFMMEvalD(ffi.dptr(src.ravel().copy()), ffi.dptr(q), None, n, ffi.dptr(trg.ravel().copy()), ffi.dptr(out),
                     len(trg), ctx, 1)
        ffi.check()
        ud = O.direct(kid, src, q, trg)
        assert rel(out, ud) < tol
        q2 = rng.standard_normal(k0 * n)
        L.PVFMMEvalD(ffi.dptr(src.ravel().copy()), ffi.dptr(q2), None, n, ffi.dptr(trg.ravel().copy()), ffi.dptr(out),
                     len(trg), ctx, 0)
        ffi.check()
        assert rel(out, O.direct(kid, src, q2, trg)) < tol
        p = ctypes.c_void_p(ctx)
        L.PVFMMDestroyContextD(ctypes.byref(p))
        assert p.value is None


def test_fortran_symbols_match_c_context(gpu):
    """pvfmmcreatecontextd_ / pvfmmevald_ / pvfmmdestroycontextd_ (arguments by reference,
    src/pvfmm-wrapper.cpp:831-876) give the C context's values bit for bit; the F variants too."""
    L = ffi.lib()
    rng = np.random.default_rng(23)
    n = 4000
    src = rng.uniform(-1, 2, (n, 3))
    q = rng.standard_normal(n)
    for real, ev, evc, cr, crc, dsf in ((np.float64, L.pvfmmevald_, L.PVFMMEvalD, L.pvfmmcreatecontextd_,
                                        L.PVFMMCreateContextD, L.pvfmmdestroycontextd_),
                                       (np.float32, L.pvfmmevalf_, L.PVFMMEvalF, L.pvfmmcreatecontextf_,
                                        L.PVFMMCreateContextF, L.pvfmmdestroycontextf_)):
        cptr = ctypes.POINTER(ctypes.c_double if real == np.float64 else ctypes.c_float)
        xs = np.ascontiguousarray(src.ravel(), dtype=real)
        qs = np.ascontiguousarray(q, dtype=real)
        out_f = np.zeros(3 * n, dtype=real)
        out_c = np.zeros(3 * n, dtype=real)
        fctx = ctypes.c_void_p()
        box = (ctypes.c_double if real == np.float64 else ctypes.c_float)(-1.0)
        i32 = ctypes.c_int32
        cr(ctypes.byref(fctx), ctypes.byref(box), ctypes.byref(i32(64)), ctypes.byref(i32(6)), ctypes.byref(i32(1)),
           ctypes.byref(i32(0)))
        ffi.check()
        assert fctx.value
        ns, nt, setup = ctypes.c_int64(n), ctypes.c_int64(n), i32(1)
        ev(xs.ctypes.data_as(cptr), qs.ctypes.data_as(cptr), ctypes.byref(ns), xs.ctypes.data_as(cptr),
           out_f.ctypes.data_as(cptr), ctypes.byref(nt), ctypes.byref(fctx), ctypes.byref(setup))
        ffi.check()
        cctx = crc(-1.0, 64, 6, 1, 0)
        ffi.check()
        evc(xs.ctypes.data_as(cptr), qs.ctypes.data_as(cptr), None, n, xs.ctypes.data_as(cptr), out_c.ctypes.data_as(cptr),
            n, cctx, 1)
        ffi.check()
        assert np.array_equal(out_f, out_c)
        assert rel(out_f, O.direct(1, src, q, src)) < (1e-3 if real == np.float32 else 1e-4)
        dsf(ctypes.byref(fctx))
        assert fctx.value is None
        p = ctypes.c_void_p(cctx)
        L.PVFMMDestroyContextD(ctypes.byref(p))


def test_python_context_float_and_mixed(gpu):
    rng = np.random.default_rng(18)
    n = 4000
    x = rng.random(3 * n)
    q = rng.standard_normal(3 * n)
    ud = O.direct(0, x.reshape(-1, 3), q[:n], x.reshape(-1, 3))
    ctx = P.FMMParticleContext(-1.0, 64, 8, P.FMMKernel.LaplacePotential)
    assert rel(ctx.evaluate(x, q, None, x)[:n], ud) < 1e-6
    ctx_m = P.FMMParticleContext(-1.0, 64, 8, P.FMMKernel.LaplacePotential, mixed_precision=True)
    assert rel(ctx_m.evaluate(x, q, None, x)[:n], ud) < 1e-5
    ctx_f = P.FMMParticleContext(-1.0, 64, 8, P.FMMKernel.LaplacePotential, dtype=np.float32)
    u32 = ctx_f.evaluate(x.astype(np.float32), q.astype(np.float32), None, x.astype(np.float32))
    assert rel(u32[:n].astype(np.float64), ud) < 1e-4


def test_device_tensor_evaluation(gpu):
    import torch

    rng = np.random.default_rng(19)
    n = 6000
    x = rng.random((n, 3))
    q = rng.standard_normal(n)
    t = P.PtFMMTree(x, x, 64).setup(P.PtFMM("laplace_poten", 6))
    u_dev = t.evaluate_device(torch.from_numpy(q).to(gpu)).cpu().numpy()
    assert np.array_equal(u_dev, t.evaluate(q))


def test_gpu_direct_sum_matches_oracle(gpu):
    """pvfmm_amd_direct (Kernel<T>::ker_poten) against the oracle's FP64 direct sum."""
    rng = np.random.default_rng(20)
    src = rng.random((3000, 3))
    trg = np.concatenate([rng.random((700, 3)), src[:5]])  # includes r = 0 pairs
    for kernel in ("laplace_poten", "laplace_grad", "stokes_vel"):
        k0, _ = O.KDIM[KID[kernel]]
        q = rng.standard_normal(k0 * len(src))
        u = P.direct(kernel, src, q, trg)
        ud = O.direct(kernel, src, q, trg)
        assert rel(u, ud) < 1e-13, kernel
