"""Adjoint FMM (pvfmm_amd.autograd): gradients of the Laplace potential and energy against
central finite differences (tests/golden/fd_energy.npz) and against the direct-sum adjoint."""
import os

import numpy as np
import pytest

import oracle as O
from pvfmm_amd.autograd import LaplaceFMM

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def test_energy_gradients_match_finite_differences(gpu):
    import torch

    f = np.load(os.path.join(G, "fd_energy.npz"))
    fmm = LaplaceFMM(f["x"], m=8, max_pts=16, precision="fp64")
    x = torch.tensor(f["x"], device=gpu, requires_grad=True)
    q = torch.tensor(f["q"], device=gpu, requires_grad=True)
    E = fmm.energy(x, q)
    E.backward()
    eE = abs(E.item() - float(f["energy"])) / abs(float(f["energy"]))
    edq = rel(q.grad.cpu().numpy(), f["dE_dq"])
    edx = rel(x.grad.cpu().numpy(), f["dE_dx"])
    print(f"energy {eE:.2e}, dE/dq {edq:.2e}, dE/dx {edx:.2e}")
    assert eE < 1e-6 and edq < 1e-6 and edx < 1e-5


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-6), ("mixed", 1e-5)])
def test_potential_vjp_matches_direct_adjoint(gpu, prec, tol):
    import torch

    rng = np.random.default_rng(31)
    n = 4000
    xs = rng.random((n, 3))
    qs = rng.standard_normal(n)
    gs = rng.standard_normal(n)
    fmm = LaplaceFMM(xs, m=8, max_pts=32, precision=prec)
    x = torch.tensor(xs, device=gpu, requires_grad=True)
    q = torch.tensor(qs, device=gpu, requires_grad=True)
    phi = fmm.potential(x, q)
    phi.backward(torch.tensor(gs, device=gpu))
    dq_ref = O.direct(0, xs, gs, xs)
    uq = O.direct(1, xs, qs, xs).reshape(-1, 3)
    ug = O.direct(1, xs, gs, xs).reshape(-1, 3)
    dx_ref = -(gs[:, None] * uq + qs[:, None] * ug)
    e_phi = rel(phi.detach().cpu().numpy(), O.direct(0, xs, qs, xs))
    e_dq = rel(q.grad.cpu().numpy(), dq_ref)
    e_dx = rel(x.grad.cpu().numpy(), dx_ref)
    print(f"{prec}: phi {e_phi:.2e}, dL/dq {e_dq:.2e}, dL/dx {e_dx:.2e}")
    assert e_phi < tol and e_dq < tol and e_dx < 10 * tol


def test_energy_fast_path_equals_composition(gpu):
    import torch

    rng = np.random.default_rng(32)
    xs = rng.random((3000, 3))
    qs = rng.standard_normal(3000)
    fmm = LaplaceFMM(xs, m=6, max_pts=32, precision="fp64")
    x1 = torch.tensor(xs, device=gpu, requires_grad=True)
    q1 = torch.tensor(qs, device=gpu, requires_grad=True)
    fmm.energy(x1, q1).backward()
    x2 = torch.tensor(xs, device=gpu, requires_grad=True)
    q2 = torch.tensor(qs, device=gpu, requires_grad=True)
    (0.5 * torch.dot(q2, fmm.potential(x2, q2))).backward()
    # both give dE/dq = (A q + A^T q) / 2 (the composition through the potential's adjoint-FMM
    # VJP) and dE/dx = -q u[q]; they differ only in summation order
    assert rel(q1.grad.cpu().numpy(), q2.grad.cpu().numpy()) < 1e-13
    assert rel(x1.grad.cpu().numpy(), x2.grad.cpu().numpy()) < 1e-13
