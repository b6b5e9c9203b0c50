"""The mixed evaluator on charges of very different size in one tree, against the FP64 evaluator
on the same tree (itself checked against the oracle in tests/test_workloads.py).

A Gaussian cluster (sigma 0.02) of 2^17 charges U[-1e3, 1e3] sits beside 2^17 charges
U[-1e-3, 1e-3] spread uniformly over the box: within one level the sources' spectra differ by
many powers of two, which is what the FP32 scalar Hadamard's run scales (the fp16 operand
split) must absorb.  The errors are stated per region; the bounds are about three times the
measured values, which equal those of the earlier three-part bf16 split of the same products
(for the gradient at the sparse targets 1.9e-5, the FP32 rounding of the cluster's far field,
which dominates their field).
"""
import numpy as np
import pytest

import oracle as O
import pvfmm_amd as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel,k1,tol_all,tol_far", [("laplace_grad", 3, 2e-6, 6e-5), ("laplace_poten", 1, 2e-6, 2e-6)])
def test_cluster_beside_sparse_small_charges(gpu, kernel, k1, tol_all, tol_far):
    rng = np.random.default_rng(7)
    na = nb = 1 << 17
    a = 0.3 + 0.02 * rng.standard_normal((na, 3))
    b = rng.uniform(0.05, 0.95, (nb, 3))
    src = np.clip(np.concatenate([a, b]), 0.001, 0.999)
    den = np.concatenate([1e3 * rng.uniform(-1, 1, na), 1e-3 * rng.uniform(-1, 1, nb)])
    far = na + np.nonzero(np.linalg.norm(src[na:] - 0.3, axis=1) > 0.3)[0]
    tree = P.PtFMMTree(src, src, 64)
    out = {}
    for prec in ("fp64", "mixed"):
        tree.setup(P.PtFMM(kernel, 10, prec))
        out[prec] = tree.evaluate(den).reshape(-1, k1)
    d = out["mixed"] - out["fp64"]

    def rel(sel):
        return np.linalg.norm(d[sel]) / np.linalg.norm(out["fp64"][sel])

    e_all, e_cluster, e_far = rel(slice(None)), rel(slice(0, na)), rel(far)
    per = np.linalg.norm(d[far], axis=1) / np.linalg.norm(out["fp64"][far], axis=1)
    print(f"{kernel}: mixed vs FP64 rel-L2 all {e_all:.2e}, cluster {e_cluster:.2e}, sparse targets away from it "
          f"{e_far:.2e} (per target median {np.median(per):.2e}, max {per.max():.2e})")
    assert e_all < tol_all and e_cluster < tol_all
    assert e_far < tol_far


def test_stokes_cluster_beside_sparse_small_forces(gpu):
    """The same geometry for stokes_vel at m = 8 (forces of size 1e3 in the cluster, 1e-3
    outside): the FP32 Stokes Hadamard scales each tile's partner spectra by one power of two
    from their per-component bounds, so the small forces' spectra sit far below the tile's
    scale wherever a tile lists partners of both kinds.  Against the oracle on the same tree
    at the pseudo-inverse cutoff the mixed evaluator uses (4 eps_32 max_S): the FP64 Stokes
    evaluator's own cutoff error (about 2e-5 at m = 8) would hide the FP32 arithmetic.
    Measured: 4.3e-7 overall and in the cluster, 9.8e-7 at the sparse targets."""
    rng = np.random.default_rng(8)
    na = nb = 1 << 15
    a = 0.3 + 0.02 * rng.standard_normal((na, 3))
    b = rng.uniform(0.05, 0.95, (nb, 3))
    src = np.clip(np.concatenate([a, b]), 0.001, 0.999)
    f = np.concatenate([1e3 * rng.uniform(-1, 1, (na, 3)), 1e-3 * rng.uniform(-1, 1, (nb, 3))])
    far = na + np.nonzero(np.linalg.norm(src[na:] - 0.3, axis=1) > 0.3)[0]
    tree = P.PtFMMTree(src, src, 64).setup(P.PtFMM("stokes_vel", 8, "mixed"))
    u = tree.evaluate(f.ravel()).reshape(-1, 3)
    ot = O.OracleTree(src, src, 64)
    ot.set_operators(O.precompute("stokes_vel", 8, "fp32"))
    uo = np.asarray(ot.evaluate(f.ravel())).reshape(-1, 3)
    d = u - uo

    def rel(sel):
        return np.linalg.norm(d[sel]) / np.linalg.norm(uo[sel])

    e_all, e_cluster, e_far = rel(slice(None)), rel(slice(0, na)), rel(far)
    print(f"stokes_vel: mixed vs oracle (FP32 cutoff) rel-L2 all {e_all:.2e}, cluster {e_cluster:.2e}, "
          f"sparse targets away from it {e_far:.2e}")
    assert e_all < 1.5e-6 and e_cluster < 1.5e-6
    assert e_far < 3e-6
