"""GPU: the adjoint FMM (pvfmm_amd_tree_evaluate_adjoint[_device]) is the exact transpose of
the evaluation's linear map.

On small adaptive trees every list is populated (U, V, W and X, including X sources handled
through the downward check surfaces of non-leaf targets), the dense matrix A of the
evaluation is assembled column by column (unit densities) and that of the adjoint row by row,
and A_adj must equal A^T to rounding.  A planted mistake in any transposed phase (a dropped
term, a transposed operator used untransposed, a wrong V-list offset) moves whole blocks of
the matrix and fails this check.  At the workload sizes the identity <g, A q> = <A^T g, q>
is checked, and A^T g against the direct sum (the kernels are symmetric, so A^T approximates
the same kernel matrix as A)."""
import numpy as np
import pytest

import oracle as O
import pvfmm_amd as P
from pvfmm_amd import ffi
from pvfmm_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def adaptive_points(rng, n):
    return np.concatenate([W.gaussian_blob(rng, n // 2, sigma=0.03), rng.random((n - n // 2, 3))])


@pytest.mark.parametrize("kernel,prec,sep,tol", [("laplace_poten", "fp64", False, 1e-12),
                                                 ("laplace_poten", "fp64", True, 1e-12),
                                                 ("stokes_vel", "fp64", False, 1e-12),
                                                 ("laplace_poten", "mixed", False, 1e-5),
                                                 ("stokes_vel", "mixed", True, 1e-5)])
def test_adjoint_is_dense_transpose(gpu, kernel, prec, sep, tol):
    rng = np.random.default_rng(77)
    k0, k1 = P.kernel_dims(kernel)
    n = 420 if k0 == 1 else 180
    src = adaptive_points(rng, n)
    trg = adaptive_points(rng, n - 37) if sep else src
    fmm = P.PtFMM(kernel, 4, prec)
    t = P.PtFMMTree(src, trg, 6).setup(fmm, n_vec_max=32)
    st = t.stats()
    print(f"{kernel} {prec} sep={sep}: v_pairs {st['v_pairs']}, far items {st['far_items']} over "
          f"{st['u_rows']} target leaves, x rows {st['x_rows']}, u pairs {st['u_pairs']}")
    assert st["v_pairs"] > 0 and st["far_items"] > st["u_rows"] and st["x_rows"] > 0
    A = t.evaluate(np.eye(k0 * len(src))).T            # (k1 n_trg, k0 n_src), column j = A e_j
    B = t.evaluate_adjoint(np.eye(k1 * len(trg)))      # row i = A^T e_i  -> B = A^T as (k1 n_trg, k0 n_src)
    e = rel(B, A)
    print(f"   ||A_adj - A^T|| / ||A|| = {e:.2e}")
    assert e < tol
    # the map is close to the kernel matrix itself (symmetric kernel): A^T g ~ K g
    g = rng.standard_normal(k1 * len(trg))
    assert rel(t.evaluate_adjoint(g), O.direct({"laplace_poten": 0, "stokes_vel": 2}[kernel], trg, g, src)) < 1e-2


@pytest.mark.parametrize("kernel,prec,bndry,tol", [("laplace_poten", "fp64", "pxyz", 1e-12),
                                                   ("laplace_poten", "fp64", "px", 1e-12),
                                                   ("stokes_vel", "fp64", "pxyz", 1e-12),
                                                   ("laplace_poten", "mixed", "pxy", 1e-5)])
def test_periodic_adjoint_is_dense_transpose(gpu, kernel, prec, bndry, tol):
    """Periodic trees: the transposed lists run through the images with their shifts negated,
    the lattice operator transposed and (PXYZ) the volume-potential correction's transpose; the
    assembled adjoint equals the transpose of the assembled evaluation to rounding."""
    rng = np.random.default_rng(78)
    k0, k1 = P.kernel_dims(kernel)
    n = 400 if k0 == 1 else 160
    src = adaptive_points(rng, n)
    trg = src if kernel == "stokes_vel" else adaptive_points(rng, n - 29)
    t = P.PtFMMTree(src, trg, 6, bndry=bndry).setup(P.PtFMM(kernel, 4, prec), n_vec_max=32)
    A = t.evaluate(np.eye(k0 * len(src))).T
    B = t.evaluate_adjoint(np.eye(k1 * len(trg)))
    e = rel(B, A)
    print(f"periodic {bndry} {kernel} {prec}: ||A_adj - A^T|| / ||A|| = {e:.2e}")
    assert e < tol


def test_periodic_energy_gradient_in_charges(gpu):
    """LaplaceFMM on a periodic tree: dE/dq from the adjoint FMM equals (A q + A^T q) / 2 and
    matches the central difference of the FMM energy to rounding (E is quadratic in q)."""
    import torch

    from pvfmm_amd.autograd import LaplaceFMM

    rng = np.random.default_rng(79)
    n = 3000
    x = rng.random((n, 3))
    q = rng.standard_normal(n)
    q -= q.mean()
    fmm = LaplaceFMM(x, m=8, max_pts=64, precision="fp64", bndry="pxyz")
    xt = torch.from_numpy(x).to("cuda:0").requires_grad_(True)
    qt = torch.from_numpy(q).to("cuda:0").requires_grad_(True)
    E = fmm.energy(xt, qt)
    E.backward()
    dq = qt.grad.cpu().numpy()
    d = rng.standard_normal(n)
    h = 1e-3
    Ep = fmm.energy(xt.detach(), torch.from_numpy(q + h * d).to("cuda:0")).item()
    Em = fmm.energy(xt.detach(), torch.from_numpy(q - h * d).to("cuda:0")).item()
    fd = (Ep - Em) / (2 * h)
    assert abs(fd - dq @ d) < 1e-9 * abs(fd)
    assert np.isfinite(xt.grad.cpu().numpy()).all()


def test_adjoint_rejects_laplace_grad(gpu):
    x = np.random.default_rng(1).random((500, 3))
    t = P.PtFMMTree(x, x, 32).setup(P.PtFMM("laplace_grad", 6))
    with pytest.raises(ffi.PVFMMError, match="symmetric kernel"):
        t.evaluate_adjoint(np.zeros(3 * 500))


@pytest.mark.parametrize("prec,tol,tol_dir", [("fp64", 1e-12, 1e-8), ("mixed", 1e-6, 1e-6)])
def test_adjoint_identity_at_energy_workload(gpu, prec, tol, tol_dir):
    """<g, A q> = <A^T g, q> on the energy workload's point set (2^20 uniform, m = 10,
    max_pts = 256), with several density pairs in one pass, and A^T g against the FP64 direct
    sum on 2048 sampled sources."""
    import torch

    c = W.uniform_energy_1m()
    x = c["src"]
    n = len(x)
    t = P.PtFMMTree(x, x, c["max_pts"]).setup(P.PtFMM(c["kernel"], c["m"], prec), n_vec_max=2)
    rng = np.random.default_rng(3)
    q = torch.tensor(rng.standard_normal((2, n)), device=gpu)
    g = torch.tensor(rng.standard_normal((2, n)), device=gpu)
    Aq = t.evaluate_device(q)
    Atg = t.evaluate_adjoint_device(g)
    for v in range(2):
        lhs = float(torch.dot(g[v], Aq[v]))
        rhs = float(torch.dot(Atg[v], q[v]))
        scale = float(torch.linalg.norm(g[v]) * torch.linalg.norm(Aq[v]))
        e = abs(lhs - rhs) / scale
        print(f"{prec} vector {v}: <g, Aq> = {lhs:.12e}, <A^T g, q> = {rhs:.12e}, |diff| / (|g||Aq|) = {e:.2e}")
        assert e < tol
    idx = np.sort(rng.choice(n, 2048, replace=False))
    gv = g[0].cpu().numpy()
    e_dir = rel(Atg[0].cpu().numpy()[idx], O.direct(0, x, gv, x[idx]))
    print(f"{prec}: A^T g vs direct sum {e_dir:.2e}")
    assert e_dir < tol_dir
