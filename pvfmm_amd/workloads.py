"""Seeded point distributions of the benchmark workloads.

The shapes follow the reference's example generators (RandUnif, RandSphr, RandElps, RandGaus in
examples/include/utils.txx:286-372) but use NumPy PCG64 streams so that every run, test and
benchmark sees the same points.  Each workload has its own stream, seed 20261201 + k.
"""
from __future__ import annotations

import numpy as np

SEED0 = 20261201


def rng_for(k: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(SEED0 + k))


def uniform(rng, n):
    return rng.random((n, 3))


def sphere_surface(rng, n, center=(0.5, 0.5, 0.5), radius=0.4):
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.asarray(center) + radius * d


def gaussian_blob(rng, n, center=(0.3, 0.3, 0.7), sigma=0.02):
    """Gaussian points, resampled until they fall inside [0,1)^3."""
    out = np.empty((0, 3))
    while len(out) < n:
        p = np.asarray(center) + sigma * rng.standard_normal((2 * (n - len(out)) + 16, 3))
        p = p[np.all((p >= 0) & (p < 1), axis=1)]
        out = np.concatenate([out, p])
    return out[:n]


def ellipsoid_surface(rng, n, axes=(0.45, 0.3, 0.15), center=(0.5, 0.5, 0.5)):
    """Area-uniform points on an ellipsoid by rejection: a direction u on the unit sphere maps
    to a*u; it is kept with probability proportional to the local area factor
    |(u0/a0, u1/a1, u2/a2)| * a0*a1*a2, normalised by its maximum (a0*a1*a2/min(a))."""
    a = np.asarray(axes, dtype=np.float64)
    gmax = np.prod(a) / a.min()
    out = np.empty((0, 3))
    while len(out) < n:
        m = 2 * (n - len(out)) + 64
        u = rng.standard_normal((m, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        g = np.prod(a) * np.linalg.norm(u / a, axis=1)
        keep = rng.random(m) < g / gmax
        out = np.concatenate([out, np.asarray(center) + u[keep] * a])
    return out[:n]


def uniform_poten_64k(n=65536):
    """laplace_poten FP64, N=65,536 uniform in [0,1)^3, densities U[-1,1], m=6, max_pts=128."""
    rng = rng_for(1)
    x = uniform(rng, n)
    q = rng.uniform(-1.0, 1.0, n)
    return dict(kernel="laplace_poten", m=6, max_pts=128, src=x, trg=x, den=q)


def adaptive_grad_2m(n=2097152):
    """laplace_grad, N=2^21 adaptive: 50% sphere |x-0.5|=0.4, 30% Gaussian (sigma 0.02 at
    (0.3,0.3,0.7)), 20% uniform; densities N(0,1); m=10, max_pts=256."""
    rng = rng_for(2)
    n_s = n // 2
    n_g = int(round(0.3 * n))
    n_u = n - n_s - n_g
    x = np.concatenate([sphere_surface(rng, n_s), gaussian_blob(rng, n_g), uniform(rng, n_u)])
    q = rng.standard_normal(n)
    return dict(kernel="laplace_grad", m=10, max_pts=256, src=x, trg=x, den=q)


def ellipsoid_stokes_1m(n=1048576):
    """stokes_vel, N=2^20 on the ellipsoid (0.45,0.3,0.15) centred at 0.5; forces N(0,1); m=8, max_pts=200."""
    rng = rng_for(3)
    x = ellipsoid_surface(rng, n)
    f = rng.standard_normal(3 * n)
    return dict(kernel="stokes_vel", m=8, max_pts=200, src=x, trg=x, den=f)


def uniform_poten_1m_32_densities(n=1048576, n_dens=32):
    """laplace_poten on one uniform tree of 2^20 points with 32 density vectors; m=8, max_pts=128."""
    rng = rng_for(4)
    x = uniform(rng, n)
    q = rng.standard_normal((n_dens, n))
    return dict(kernel="laplace_poten", m=8, max_pts=128, src=x, trg=x, den=q)


def uniform_clouds_64(n_clouds=64):
    """64 independent uniform clouds with sizes floor(2^U(12,18)); m=8, max_pts=128."""
    rng = rng_for(5)
    sizes = np.floor(2.0 ** rng.uniform(12, 18, n_clouds)).astype(np.int64)
    clouds = [(uniform(rng, int(s)), rng.standard_normal(int(s))) for s in sizes]
    return dict(kernel="laplace_poten", m=8, max_pts=128, clouds=clouds)


def uniform_energy_1m(n=1048576):
    """Energy E = 0.5 sum q phi for laplace_poten, N=2^20 uniform, q ~ N(0,1), m=10, max_pts=256."""
    rng = rng_for(6)
    x = uniform(rng, n)
    q = rng.standard_normal(n)
    return dict(kernel="laplace_poten", m=10, max_pts=256, src=x, trg=x, den=q)


def uniform_periodic_1m(n=1048576):
    """laplace_poten in the periodic unit cube (PXYZ), N=2^20 uniform, q ~ N(0,1) (the net charge
    is neutralised by a uniform background); m=8, max_pts=128."""
    rng = rng_for(7)
    x = uniform(rng, n)
    q = rng.standard_normal(n)
    return dict(kernel="laplace_poten", m=8, max_pts=128, src=x, trg=x, den=q, bndry="pxyz")


WORKLOADS = {
    "uniform-poten-64k": uniform_poten_64k,
    "adaptive-grad-2m": adaptive_grad_2m,
    "ellipsoid-stokes-1m": ellipsoid_stokes_1m,
    "densities-32": uniform_poten_1m_32_densities,
    "clouds-64": uniform_clouds_64,
    "periodic-poten-1m": uniform_periodic_1m,
    "energy-1m": uniform_energy_1m,
}
