"""Operator precompute: the precompute file (Precomp_<ker>_m<m>[_mixed].data in the container
of the reference's PrecompMat::Save2File / LoadFile, include/precomp_mat.txx:212-347) and the
device precompute (BuildMatrix, Jacobi SVD and the U2U/D2D products on the GPU) against the
host path."""
import os
import struct

import numpy as np
import pytest

import pvfmm_amd as P

OPS = ("uc2ue0", "uc2ue1", "dc2de0", "dc2de1", "u2u", "d2d", "v1")


def read_container(path):
    """The reference's container: int sizeof(T), int scale_invar, then for each of the 16
    Mat_Type values an int count and that many (int rows, int cols, rows*cols values)."""
    with open(path, "rb") as f:
        data = f.read()
    sz, inv = struct.unpack_from("<ii", data, 0)
    off = 8
    mats = []
    for _ in range(16):
        (n,) = struct.unpack_from("<i", data, off)
        off += 4
        lst = []
        for _ in range(n):
            r, c = struct.unpack_from("<ii", data, off)
            off += 8
            lst.append(np.frombuffer(data, dtype=np.float64, count=r * c, offset=off).reshape(r, c))
            off += 8 * r * c
        mats.append(lst)
    assert off == len(data)
    return sz, inv, mats


@pytest.fixture
def cache_dir(tmp_path, monkeypatch):
    monkeypatch.setenv("PVFMM_DIR", str(tmp_path))
    monkeypatch.setenv("PVFMM_PRECOMP", "host")
    return tmp_path


@pytest.mark.parametrize("kernel,name,prec,suffix", [("laplace_poten", "laplace", "fp64", ""),
                                                     ("stokes_vel", "stokes_vel", "mixed", "_mixed")])
def test_precomp_file_round_trip(cache_dir, kernel, name, prec, suffix):
    m = 4
    f1 = P.PtFMM(kernel, m, prec)
    path = cache_dir / f"Precomp_{name}_m{m}{suffix}.data"
    assert path.exists()
    sz, inv, mats = read_container(path)
    assert (sz, inv) == (8, 1)
    ns = 6 * (m - 1) ** 2 + 2
    e = 3 if kernel == "stokes_vel" else 1
    K = ns * e
    # Mat_Type slots of include/precomp_mat.hpp:21-39: UC2UE0 0, UC2UE1 1, DC2DE0 2, DC2DE1 3,
    # U2U 5, D2D 6, V1 14
    assert mats[0][0].shape[0] == K and mats[1][0].shape == (mats[0][0].shape[1], K)
    assert len(mats[5]) == 8 and all(a.shape == (K, K) for a in mats[5])
    assert len(mats[6]) == 8 and all(a.shape == (K, K) for a in mats[6])
    assert len(mats[14]) == 27 and mats[14][13].size == 0
    assert all(len(mats[t]) == 0 for t in (4, 7, 8, 9, 10, 11, 12, 13))
    # BC_Type: one periodic boundary operator per BoundaryType, FreeSpace as zeros
    assert len(mats[15]) == 4 and all(a.shape == (K, K) for a in mats[15]) and not mats[15][0].any()
    np.testing.assert_array_equal(mats[0][0].ravel(), f1.operator("uc2ue0"))
    for b, name_ in ((1, "bc_px"), (2, "bc_pxy"), (3, "bc_pxyz")):
        np.testing.assert_array_equal(mats[15][b].ravel(), f1.operator(name_))
    mtime = path.stat().st_mtime_ns
    f2 = P.PtFMM(kernel, m, prec)  # read back
    for op in OPS + ("bc_px", "bc_pxy", "bc_pxyz"):
        np.testing.assert_array_equal(f1.operator(op), f2.operator(op))
    assert path.stat().st_mtime_ns == mtime  # an existing file is not rewritten
    # the same evaluation from computed and from read operators
    x = np.random.default_rng(2).random((800, 3))
    k0 = 3 if kernel == "stokes_vel" else 1
    q = np.random.default_rng(3).standard_normal(k0 * 800)
    if P.device_count() > 0:
        a = P.PtFMMTree(x, x, 32).setup(f1).evaluate(q)
        b = P.PtFMMTree(x, x, 32).setup(f2).evaluate(q)
        np.testing.assert_array_equal(a, b)


def test_precomp_file_with_other_contents_is_kept(cache_dir):
    """A file of the same name whose matrices have other shapes (the reference keeps full
    square factors and stores symmetry-class representatives only) is read where it fits,
    the rest is computed, and the file is left as it was."""
    m = 4
    ns = 6 * (m - 1) ** 2 + 2
    ref = P.PtFMM("laplace_poten", m)
    os.remove(cache_dir / f"Precomp_laplace_m{m}.data")
    F0 = ref.operator("uc2ue0").reshape(ns, -1)
    F1 = ref.operator("uc2ue1").reshape(-1, ns)
    r = F0.shape[1]
    # full square factors with zero columns / rows past the rank, one U2U matrix only
    F0f = np.zeros((ns, ns)); F0f[:, :r] = F0
    F1f = np.zeros((ns, ns)); F1f[:r] = F1
    D0 = ref.operator("dc2de0").reshape(ns, -1)
    D1 = ref.operator("dc2de1").reshape(-1, ns)
    mats = [[] for _ in range(16)]
    mats[0], mats[1], mats[2], mats[3] = [F0f], [F1f], [D0], [D1]
    mats[5] = [np.eye(ns)]
    blob = struct.pack("<ii", 8, 1)
    for lst in mats:
        blob += struct.pack("<i", len(lst))
        for a in lst:
            blob += struct.pack("<ii", *a.shape) + np.ascontiguousarray(a, dtype=np.float64).tobytes()
    path = cache_dir / f"Precomp_laplace_m{m}.data"
    path.write_bytes(blob)
    f = P.PtFMM("laplace_poten", m)
    assert path.read_bytes() == blob
    np.testing.assert_array_equal(f.operator("uc2ue0"), F0f.ravel())
    # U2U recomputed from the read factors: the zero columns change nothing
    np.testing.assert_allclose(f.operator("u2u"), ref.operator("u2u"), rtol=0, atol=1e-12 * np.abs(ref.operator("u2u")).max())
    np.testing.assert_array_equal(f.operator("v1"), ref.operator("v1"))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,m,prec", [("laplace_poten", 10, "fp64"), ("laplace_poten", 10, "mixed"),
                                           ("stokes_vel", 6, "fp64")])
def test_device_precompute_matches_host(gpu, monkeypatch, kernel, m, prec):
    """BuildMatrix, the Jacobi SVD and the U2U/D2D products on the device against the host
    path: the translations act alike on the densities sources produce (the matrices
    themselves may differ in directions the cutoff leaves undetermined), and FMM evaluations
    with either operator set agree to rounding."""
    import time

    monkeypatch.delenv("PVFMM_DIR", raising=False)
    monkeypatch.setenv("PVFMM_PRECOMP", "host")
    t0 = time.perf_counter()
    fh = P.PtFMM(kernel, m, prec)
    th = time.perf_counter() - t0
    monkeypatch.delenv("PVFMM_PRECOMP")
    t0 = time.perf_counter()
    fd = P.PtFMM(kernel, m, prec)
    td = time.perf_counter() - t0
    ns = 6 * (m - 1) ** 2 + 2
    K = ns * (3 if kernel == "stokes_vel" else 1)
    e_u, e_d = _translation_action(fh, fd, kernel, m)
    rng = np.random.default_rng(9)
    x = rng.random((6000, 3))
    k0 = 3 if kernel == "stokes_vel" else 1
    q = rng.standard_normal(k0 * 6000)
    uh = P.PtFMMTree(x, x, 64).setup(fh).evaluate(q)
    ud = P.PtFMMTree(x, x, 64).setup(fd).evaluate(q)
    e_fmm = float(np.linalg.norm(ud - uh) / np.linalg.norm(uh))
    print(f"{kernel} m={m} {prec}: host {th:.2f} s, device {td:.2f} s; U2U action {e_u:.2e}, D2D action {e_d:.2e}, "
          f"FMM {e_fmm:.2e}")
    assert e_fmm < (1e-10 if prec == "fp64" else 1e-6)
    assert e_u < 1e-9 and e_d < 1e-9


def _translation_action(fh, fd, kernel, m):
    """The U2U and D2D translations of two operator sets applied to the data they meet in an
    FMM, compared where it matters.  At m = 10 in FP64 the check-to-equivalent systems have
    condition numbers near 1e16, so the singular vectors kept just above the 4 eps max_S
    cutoff are not determined by the matrix: two correct SVDs (a host and a device Jacobi
    sweep, whose sums round differently) give factors, and translation matrices, that differ by
    1e-2 in norm.  What is fixed is their action on densities produced by sources:
      U2U: random sources in each child box -> child equivalent density (child check potential
           times the depth-1 check-to-equivalent map, half the depth-0 one for these kernels of
           degree -1) -> U2U -> potential of the parent equivalent density at 200 points at
           distance 3 from the parent centre;
      D2D: sources at distance 3 from the parent -> potential on the parent's downward check
           surface -> D2D -> child check potential.
    Returns the relative differences of the two sets' results (host as reference)."""
    import oracle as O

    kid = {"laplace_poten": 0, "laplace_grad": 1, "stokes_vel": 2}[kernel]
    ke = O.expansion_kernel(kid)
    k0 = 3 if ke == 2 else 1
    ns = 6 * (m - 1) ** 2 + 2
    K = ns * k0
    rng = np.random.default_rng(31)
    d = rng.standard_normal((200, 3))
    far = 0.5 + 3.0 * d / np.linalg.norm(d, axis=1)[:, None]
    ue = O.surf_at(m, [0, 0, 0], 0, O.RAD0)
    dc = O.surf_at(m, [0, 0, 0], 0, O.RAD0)
    ku = np.zeros(2)
    kd = np.zeros(2)
    for c in range(8):
        bits = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], dtype=np.float64)
        y = bits * 0.5 + 0.5 * rng.random((40, 3))
        qy = rng.standard_normal(40 * k0)
        child_check = qy @ O.kernel_matrix(ke, y, O.surf_at(m, bits * 0.5, 1, O.RAD1))
        # the two factors applied in turn (their explicit product would lose digits at m = 10)
        q_child = 0.5 * ((child_check @ fh.operator("uc2ue0").reshape(K, -1)) @ fh.operator("uc2ue1").reshape(-1, K))
        kf = O.kernel_matrix(ke, ue, far)
        ph = q_child @ fh.operator("u2u").reshape(8, K, K)[c] @ kf
        pd = q_child @ fd.operator("u2u").reshape(8, K, K)[c] @ kf
        ku += [np.sum((pd - ph) ** 2), np.sum(ph ** 2)]
        parent_check = qy @ O.kernel_matrix(ke, far[:40], dc)
        dh = parent_check @ fh.operator("d2d").reshape(K, 8, K)[:, c, :]
        dd = parent_check @ fd.operator("d2d").reshape(K, 8, K)[:, c, :]
        kd += [np.sum((dd - dh) ** 2), np.sum(dh ** 2)]
    return float(np.sqrt(ku[0] / ku[1])), float(np.sqrt(kd[0] / kd[1]))
