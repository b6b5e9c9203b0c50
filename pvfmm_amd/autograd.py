"""Differentiable Laplace FMM (adjoint FMM behind torch.autograd.Function).

For sources = targets = x (n points) and charges q the FMM gives the potential
phi_i = sum_{j != i} q_j / (4 pi |x_i - x_j|) (laplace_poten, include/kernel.txx:1276-1287) and
the field u_i = sum_{j != i} q_j (x_i - x_j) / (4 pi |x_i - x_j|^3) = -grad phi(x_i)
(laplace_grad, include/kernel.txx:1316-1330).  The kernel is symmetric, so the adjoint of
the potential map is the map itself:

* potential phi(x, q) with upstream gradient g:
    dL/dq = K g                        (a potential FMM with density g)
    dL/dx_i = -g_i u[q]_i - q_i u[g]_i (one gradient FMM with the two densities q and g)
* energy E(x, q) = 1/2 q . phi:
    dE/dq = phi                        (kept from the forward pass)
    dE/dx_i = -q_i u[q]_i              (one gradient FMM)

The tree is built once for the point set given to LaplaceFMM (positions enter through the
tree); x passed to potential()/energy() must be those positions and only receives the
gradient.  All FMM work runs in libpvfmm.so on the GPU.
"""
from __future__ import annotations

import numpy as np

from .wrapper import PtFMM, PtFMMTree


class LaplaceFMM:
    """Potential and energy of a fixed point set with FMM forward and adjoint-FMM backward.

    precision: "fp64" or "mixed" (FP32 bulk, FP64 operators and merge)."""

    def __init__(self, points, m: int = 10, max_pts: int = 256, precision: str = "mixed"):
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        self.n = len(pts)
        self.points = pts
        self._tp = PtFMMTree(pts, pts, max_pts).setup(PtFMM("laplace_poten", m, precision), n_vec_max=2)
        self._tg = PtFMMTree(pts, pts, max_pts).setup(PtFMM("laplace_grad", m, precision), n_vec_max=2)

    # -- raw evaluations on float64 GPU tensors
    def _phi(self, q):
        return self._tp.evaluate_device(q.contiguous())

    def _field(self, dens):
        """dens (k, n) -> (k, n, 3) fields -grad phi of each density."""
        out = self._tg.evaluate_device(dens.contiguous())
        return out.reshape(dens.shape[0], self.n, 3)

    def potential(self, x, q):
        return _Potential.apply(x, q, self)

    def energy(self, x, q):
        return _Energy.apply(x, q, self)


def _check(fmm: LaplaceFMM, x, q):
    import torch

    if x.shape != (fmm.n, 3) or q.shape != (fmm.n,):
        raise ValueError("x must be (n, 3) and q (n,) for the point set of this LaplaceFMM")
    if not (q.is_cuda and q.dtype == torch.float64):
        raise ValueError("q must be a float64 GPU tensor")


def _make_functions():
    import torch

    class Potential(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, q, fmm):
            _check(fmm, x, q)
            ctx.fmm = fmm
            ctx.save_for_backward(q)
            return fmm._phi(q.detach())

        @staticmethod
        def backward(ctx, g):
            (q,) = ctx.saved_tensors
            fmm = ctx.fmm
            g = g.contiguous()
            dq = fmm._phi(g) if ctx.needs_input_grad[1] else None
            dx = None
            if ctx.needs_input_grad[0]:
                u = fmm._field(torch.stack([q.detach(), g]))
                dx = -(g[:, None] * u[0] + q.detach()[:, None] * u[1])
            return dx, dq, None

    class Energy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, q, fmm):
            _check(fmm, x, q)
            phi = fmm._phi(q.detach())
            ctx.fmm = fmm
            ctx.save_for_backward(q, phi)
            return 0.5 * torch.dot(q.detach(), phi)

        @staticmethod
        def backward(ctx, gE):
            q, phi = ctx.saved_tensors
            dq = gE * phi if ctx.needs_input_grad[1] else None
            dx = None
            if ctx.needs_input_grad[0]:
                u = ctx.fmm._field(q.detach()[None, :])[0]
                dx = -gE * q.detach()[:, None] * u
            return dx, dq, None

    return Potential, Energy


class _Lazy:
    """Defers the torch import (and Function class creation) to first use."""

    def __init__(self, idx):
        self.idx = idx
        self.fn = None

    def apply(self, *args):
        if self.fn is None:
            self.fn = _make_functions()[self.idx]
        return self.fn.apply(*args)


_Potential = _Lazy(0)
_Energy = _Lazy(1)
