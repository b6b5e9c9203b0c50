"""Ewald sums for the unit cube lattice (test infrastructure): the independent checks of the
periodic (PXYZ) Laplace and Stokes FMM (Stokes: `stokes`, Hasimoto's splitting), and the
lattice sums of neutral Laplace charges periodic along x only (`laplace_wire`) and along x and y
(`laplace_slab`).

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
from scipy.special import erf, erfc, erfcx

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


def laplace_wire(src, q, trg, n_near=100):
    """Potential of neutral charges repeated along x with period 1 (free along y and z): the
    images |n| <= n_near summed exactly, the rest by the midpoint Euler-Maclaurin formula
        sum_{n > N} h(n) = int_{N+1/2}^inf h + h'(N + 1/2) / 24 + O(third derivative)
    applied to h(u) = sum_j q_j / sqrt((u -+ d_j)^2 + rho_j^2) (d_j = x - x_j along the axis,
    rho_j the distance across it), whose integral is - sum_j q_j ln(X_j + sqrt(X_j^2 + rho_j^2)),
    X_j = N + 1/2 -+ d_j, for sum_j q_j = 0.  The images of the +n and -n sides pair up, which
    fixes the (otherwise conditionally convergent) dipole part the way a symmetric sum does."""
    src = np.asarray(src, dtype=np.float64)
    trg = np.asarray(trg, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    assert abs(q.sum()) < 1e-12 * np.abs(q).sum(), "laplace_wire needs zero net charge"
    out = np.zeros(len(trg))
    n = np.arange(-n_near, n_near + 1, dtype=np.float64)
    a = n_near + 0.5
    for t0 in range(0, len(trg), 64):
        T = trg[t0:t0 + 64]
        d = T[:, None, 0] - src[None, :, 0]  # (t, s)
        rho2 = ((T[:, None, 1:] - src[None, :, 1:]) ** 2).sum(-1)
        r = np.sqrt((d[..., None] - n) ** 2 + rho2[..., None])  # (t, s, images)
        with np.errstate(divide="ignore"):
            inv = np.where(r > 0, 1.0 / r, 0.0)
        acc = (inv.sum(-1) * q[None, :]).sum(1)
        for X in (a - d, a + d):
            R = np.sqrt(X * X + rho2)
            acc += (-np.log(X + R) * q[None, :]).sum(1)  # tail integral
            acc += (-(X / R ** 3) * q[None, :]).sum(1) / 24  # h'(N + 1/2) / 24
        out[t0:t0 + 64] = acc
    return out / (4 * np.pi)


def laplace_slab(src, q, trg, alpha=6.0, kmax=12):
    """Ewald sum of neutral charges repeated along x and y with period 1 (free along z; Parry's
    two-dimensional splitting): with r = x - x_j, z = z - z_j and k = 2 pi (integer pair),
        4 pi phi = sum_j q_j [ sum_n erfc(a|r + n|) / |r + n|
                               + pi sum_{k != 0} cos(k . r) / |k| exp(-a^2 z^2 - k^2 / 4a^2)
                                     (erfcx(a z + k / 2a) + erfcx(-a z + k / 2a))
                               - 2 sqrt(pi) (exp(-a^2 z^2) / a + sqrt(pi) z erf(a z)) ]
                   - 2 a / sqrt(pi) q_t (target on a source)
    (the k sum's terms written with erfcx to keep exp(k z) erfc(...) finite).  The square image
    shells pair up, the symmetric-sum convention for the in-plane dipole part; the sum does not
    depend on a."""
    src = np.asarray(src, dtype=np.float64)
    trg = np.asarray(trg, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    assert abs(q.sum()) < 1e-12 * np.abs(q).sum(), "laplace_slab needs zero net charge"
    out = np.zeros(len(trg))
    imgs = np.array([[i, j, 0] for i in (-1, 0, 1) for j in (-1, 0, 1)], dtype=np.float64)
    # k and -k give the same term: one half plane of the disc |m| <= kmax, counted twice; the
    # z factor depends on |k| only, so it is evaluated once per distinct |k|
    r_ = np.arange(-kmax, kmax + 1)
    m = np.stack(np.meshgrid(r_, r_, indexing="ij"), -1).reshape(-1, 2)
    m = m[((m[:, 0] > 0) | ((m[:, 0] == 0) & (m[:, 1] > 0))) & ((m * m).sum(1) <= kmax * kmax)]
    kv = 2 * np.pi * m.astype(np.float64)
    m2, inv = np.unique((m * m).sum(1), return_inverse=True)
    kn = 2 * np.pi * np.sqrt(m2.astype(np.float64))
    kk = kn / (2 * alpha)
    for t0 in range(0, len(trg), 32):
        T = trg[t0:t0 + 32]
        for sh in imgs:
            r = T[:, None, :] - src[None, :, :] + sh
            d = np.sqrt((r * r).sum(-1))
            ok = d > 0
            ds = np.where(ok, d, 1.0)
            out[t0:t0 + 32] += (np.where(ok, erfc(alpha * ds) / ds, 0.0) * q).sum(1)
        rxy = T[:, None, :2] - src[None, :, :2]  # (t, s, 2)
        z = T[:, None, 2] - src[None, :, 2]  # (t, s)
        az = alpha * z[..., None]
        g = np.exp(-az * az - kk * kk) * (erfcx(az + kk) + erfcx(-az + kk)) / kn  # (t, s, |k|)
        out[t0:t0 + 32] += 2 * np.pi * ((np.cos(rxy @ kv.T) * g[..., inv]).sum(-1) * q).sum(1)
        out[t0:t0 + 32] -= 2 * math.sqrt(math.pi) * ((np.exp(-(alpha * z) ** 2) / alpha
                                                      + math.sqrt(math.pi) * z * erf(alpha * z)) * q).sum(1)
    on = {tuple(x): i for i, x in enumerate(src)}
    for t, x in enumerate(trg):
        j = on.get(tuple(x))
        if j is not None:
            out[t] -= 2 * alpha / math.sqrt(math.pi) * q[j]
    return out / (4 * np.pi)
