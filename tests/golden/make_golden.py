"""Writes the committed golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).  The reference ships no golden vectors and cannot be
built here (its SCTL submodule is not vendored), so the fixtures come from this repository's
oracle and from closed forms, each cross-checked when written:

* kernel_kat.npz    -- direct sums of the three kernels at seeded points, summed with
                       math.fsum (exactly rounded) from the closed forms;
* tree_small.npz    -- a seeded adaptive tree: Morton keys, depths, leaf flags, point counts
                       and every interaction list of the oracle (bit-exact integer data);
* m2l_kat.npz       -- dense equivalent->check M2L of seeded child densities for three parent
                       offsets at m = 6 (the values the FFT M2L must reproduce);
* fd_energy.npz     -- E = 1/2 q^T K q for Laplace, with dE/dq and dE/dx by central finite
                       differences (h = 1e-4) of the closed form;
* fmm_small.npz     -- the oracle's FMM potential (laplace_grad, m = 6) on a seeded cloud.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402


def kernel_fsum(kid, src, den, trg):
    """Closed-form kernels (include/kernel.txx:1276-1364), exactly rounded sums."""
    k0, k1 = O.KDIM[kid]
    out = np.zeros((len(trg), k1))
    for i, t in enumerate(trg):
        terms = [[] for _ in range(k1)]
        for j, s in enumerate(src):
            r = t - s
            rr = float(r @ r)
            if rr == 0.0:
                continue
            rn = math.sqrt(rr)
            if kid == 0:
                terms[0].append(den[j] / (4 * math.pi * rn))
            elif kid == 1:
                for c in range(3):
                    terms[c].append(den[j] * r[c] / (4 * math.pi * rn ** 3))
            else:
                f = den[3 * j:3 * j + 3]
                fr = float(f @ r)
                for c in range(3):
                    terms[c].append((f[c] / rn + r[c] * fr / rn ** 3) / (8 * math.pi))
        out[i] = [math.fsum(tc) for tc in terms]
    return out.ravel()


def main():
    rng = np.random.default_rng(20261202)
    # ---- kernel KATs
    kat = {}
    src = rng.random((200, 3))
    trg = np.concatenate([rng.random((49, 3)), src[:1]])  # one coincident pair
    kat["src"], kat["trg"] = src, trg
    for kid, name in ((0, "laplace_poten"), (1, "laplace_grad"), (2, "stokes_vel")):
        den = rng.standard_normal(O.KDIM[kid][0] * len(src))
        val = kernel_fsum(kid, src, den, trg)
        assert np.allclose(val, O.direct(kid, src, den, trg), rtol=1e-12, atol=0)
        kat[f"{name}_den"], kat[f"{name}_val"] = den, val
    np.savez_compressed(os.path.join(HERE, "kernel_kat.npz"), **kat)

    # ---- tree and lists
    n = 900
    x = np.concatenate([0.5 + 0.35 * (lambda d: d / np.linalg.norm(d, axis=1)[:, None])(rng.standard_normal((n // 2, 3))),
                        rng.random((n - n // 2, 3))])
    t = O.OracleTree(x, x, 16)
    nodes = t.nodes()
    tree = {"points": x, "max_pts": np.int64(16), "nodes": nodes}
    for ty, nm in ((O.U0, "U0"), (O.U1, "U1"), (O.U2, "U2"), (O.V1, "V1"), (O.W, "W"), (O.X, "X")):
        tree[f"list_{nm}"] = t.lists(ty)
    tree["point_order_src"] = t.point_order(0)
    np.savez_compressed(os.path.join(HERE, "tree_small.npz"), **tree)

    # ---- dense M2L at m = 6 (children of two parents at offset R, RAD0 equivalent/check)
    m = 6
    ns = 6 * (m - 1) ** 2 + 2
    m2l = {"m": np.int64(m)}
    for R in (0, 10, 22):
        Rv = np.array([R % 3 - 1, (R // 3) % 3 - 1, R // 9 - 1])
        X = rng.standard_normal((8, ns))
        dense = np.zeros((8, ns))
        for tt in range(8):
            bt = np.array([tt & 1, (tt >> 1) & 1, (tt >> 2) & 1])
            chk = O.surf_at(m, bt * 1.0, 0, O.RAD0)
            for s in range(8):
                bs = np.array([s & 1, (s >> 1) & 1, (s >> 2) & 1]) + 2 * Rv
                if np.abs(bs - bt).max() < 2:
                    continue
                eq = O.surf_at(m, bs * 1.0, 0, O.RAD0)
                dense[tt] += X[s] @ O.kernel_matrix(0, eq, chk)
        m2l[f"R{R}_X"], m2l[f"R{R}_Y"] = X, dense
    np.savez_compressed(os.path.join(HERE, "m2l_kat.npz"), **m2l)

    # ---- finite-difference energy gradients (Laplace)
    ne = 256
    xe = rng.random((ne, 3))
    qe = rng.standard_normal(ne)

    def energy(xx, qq):
        d = xx[:, None, :] - xx[None, :, :]
        r = np.sqrt((d ** 2).sum(-1))
        np.fill_diagonal(r, np.inf)
        return 0.5 * qq @ ((1 / (4 * np.pi * r)) @ qq)

    h = 1e-4
    dq = np.array([(energy(xe, qe + h * e) - energy(xe, qe - h * e)) / (2 * h) for e in np.eye(ne)])
    dx = np.zeros((ne, 3))
    for i in range(ne):
        for c in range(3):
            xp, xm = xe.copy(), xe.copy()
            xp[i, c] += h
            xm[i, c] -= h
            dx[i, c] = (energy(xp, qe) - energy(xm, qe)) / (2 * h)
    phi = O.direct(0, xe, qe, xe)
    g = O.direct(1, xe, qe, xe).reshape(-1, 3)
    # central differences: O(h^2) truncation, largest at the closest pairs
    assert np.abs(dq - phi).max() < 1e-8 * np.abs(phi).max()
    assert np.abs(dx + qe[:, None] * g).max() < 1e-5 * np.abs(dx).max()
    np.savez_compressed(os.path.join(HERE, "fd_energy.npz"), x=xe, q=qe, energy=energy(xe, qe), dE_dq=dq, dE_dx=dx)

    # ---- oracle FMM
    nf = 3000
    xf = rng.random((nf, 3))
    qf = rng.standard_normal(nf)
    uf = O.fmm("laplace_grad", 6, 48, xf, qf)
    ud = O.direct(1, xf, qf, xf)
    assert np.linalg.norm(uf - ud) / np.linalg.norm(ud) < 1e-4
    np.savez_compressed(os.path.join(HERE, "fmm_small.npz"), x=xf, q=qf, m=np.int64(6), max_pts=np.int64(48), u=uf)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
