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
