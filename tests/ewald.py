"""Ewald sums for the unit cube lattice (test infrastructure): the independent checks of the
periodic (PXYZ) Laplace and Stokes FMM (Stokes: `stokes`, Hasimoto's splitting).

Potential of point charges q_j at x_j in [0,1)^3 with all their lattice images, plus (for a net
charge Q) a uniform neutralising background, with the tin-foil (no surface dipole) convention:
    phi(x) = 1/(4 pi) [ sum_j q_j sum_n erfc(a|r + n|) / |r + n|
                        + 4 pi sum_{k != 0} exp(-k^2 / 4a^2) / k^2 sum_j q_j cos(k . r)
                        - pi Q / a^2  -  2 a / sqrt(pi) q_t (target on a source, r = 0 left out) ]
with r = x - x_j and k = 2 pi (integer vector).  field(x) = -grad phi(x).  The FMM potential is
defined up to an additive constant, so comparisons remove the mean difference."""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc

ALPHA = 6.0   # erfc(6) ~ 2e-17: the 27 nearest images cover the real-space sum
KMAX = 12     # exp(-(2 pi 13)^2 / (4 * 36)) ~ 1e-20


def _kvecs():
    r = np.arange(-KMAX, KMAX + 1)
    m = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    m = m[np.any(m != 0, axis=1)]
    k = 2 * np.pi * m.astype(np.float64)
    k2 = (k * k).sum(1)
    return k, np.exp(-k2 / (4 * ALPHA * ALPHA)) / k2


def laplace(src, q, trg, grad=False):
    """Ewald potential (grad=False, (n_trg,)) or field -grad phi (grad=True, (n_trg, 3))."""
    src = np.asarray(src, dtype=np.float64)
    trg = np.asarray(trg, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    a = ALPHA
    out = np.zeros((len(trg), 3)) if grad else np.zeros(len(trg))
    imgs = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=np.float64)
    for t0 in range(0, len(trg), 256):
        T = trg[t0:t0 + 256]
        for n in imgs:
            r = T[:, None, :] - src[None, :, :] + n
            d = np.sqrt((r * r).sum(-1))
            ok = d > 0
            ds = np.where(ok, d, 1.0)
            if grad:
                f = np.where(ok, (erfc(a * ds) + 2 * a * ds / math.sqrt(math.pi) * np.exp(-(a * ds) ** 2)) / ds ** 3, 0.0)
                out[t0:t0 + 256] += np.einsum("ts,tsk,s->tk", f, r, q)
            else:
                out[t0:t0 + 256] += (np.where(ok, erfc(a * ds) / ds, 0.0) * q[None, :]).sum(1)
    k, w = _kvecs()
    for k0 in range(0, len(k), 2048):
        kk, ww = k[k0:k0 + 2048], w[k0:k0 + 2048]
        S = (q[None, :] * np.exp(-1j * (kk @ src.T))).sum(1)  # structure factor
        E = np.exp(1j * (trg @ kk.T))  # (n_trg, nk)
        if grad:
            out += 4 * np.pi * ((E * S[None, :]).imag * ww[None, :]) @ kk
        else:
            out += 4 * np.pi * ((E * S[None, :]).real @ ww)
    if not grad:
        out -= math.pi * q.sum() / (a * a)
        # targets sitting on sources: the r = 0 term's smooth part is in the k sum
        on = {tuple(x): i for i, x in enumerate(src)}
        for t, x in enumerate(trg):
            j = on.get(tuple(x))
            if j is not None:
                out[t] -= 2 * a / math.sqrt(math.pi) * q[j]
    return out / (4 * np.pi)


def stokes(src, f, trg, xi=6.0, kmax=16):
    """Hasimoto's Ewald sum of the periodic Stokeslet S = (I + r r / r^2) / (8 pi r) for forces
    of zero sum: the screening (1 + k^2 / 4 xi^2) exp(-k^2 / 4 xi^2) in Fourier space, whose
    real-space complement is, with r the image displacement,
        (erfc(xi r) / r - 2 xi / sqrt(pi) exp(-xi^2 r^2)) I
      + (erfc(xi r) / r + 2 xi / sqrt(pi) exp(-xi^2 r^2)) r r / r^2      (over 8 pi),
    and the Fourier part sum_{k != 0} screening(k) (I - k k / k^2) / k^2 f^(k) exp(i k.x); a
    target on a source loses the smooth part of its own r = 0 term, xi / (2 pi^(3/2)) f.  (Derived
    from S = (I lap - grad grad) r / (8 pi) and r convolved with the screening distribution; the
    sum does not depend on xi.)  Returns (n_trg, 3) velocities, defined up to a uniform flow."""
    src = np.asarray(src, dtype=np.float64)
    trg = np.asarray(trg, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64).reshape(-1, 3)
    c = 2 * xi / math.sqrt(math.pi)
    u = np.zeros((len(trg), 3))
    imgs = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=np.float64)
    for t0 in range(0, len(trg), 128):
        T = trg[t0:t0 + 128]
        for n in imgs:
            r = T[:, None, :] - src[None, :, :] + n
            d = np.sqrt((r * r).sum(-1))
            ok = d > 0
            ds = np.where(ok, d, 1.0)
            rh = r / ds[..., None]
            e, ec = np.exp(-(xi * ds) ** 2), erfc(xi * ds) / ds
            a, b = np.where(ok, ec - c * e, 0.0), np.where(ok, ec + c * e, 0.0)
            fr = (rh * f[None, :, :]).sum(-1)
            u[t0:t0 + 128] += (a[..., None] * f[None, :, :] + (b * fr)[..., None] * rh).sum(1) / (8 * np.pi)
    r_ = np.arange(-kmax, kmax + 1)
    m = np.stack(np.meshgrid(r_, r_, r_, indexing="ij"), -1).reshape(-1, 3)
    m = m[np.any(m != 0, axis=1)]
    for k0 in range(0, len(m), 4096):
        k = 2 * np.pi * m[k0:k0 + 4096].astype(np.float64)
        k2 = (k * k).sum(1)
        a2 = k2 / (4 * xi * xi)
        w = (1 + a2) * np.exp(-a2) / k2
        fh = np.exp(-1j * (k @ src.T)) @ f  # (nk, 3)
        proj = fh - k * ((k * fh).sum(1) / k2)[:, None]
        u += ((np.exp(1j * (trg @ k.T)) * w[None, :]) @ proj).real
    on = {tuple(x): i for i, x in enumerate(src)}
    for t, x in enumerate(trg):
        j = on.get(tuple(x))
        if j is not None:
            u[t] -= xi / (2 * math.pi ** 1.5) * f[j]
    return u
