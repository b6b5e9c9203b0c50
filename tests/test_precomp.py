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
    assert all(len(mats[t]) == 0 for t in (4, 7, 8, 9, 10, 11, 12, 13, 15))
    np.testing.assert_array_equal(mats[0][0].ravel(), f1.operator("uc2ue0"))
    mtime = path.stat().st_mtime_ns
    f2 = P.PtFMM(kernel, m, prec)  # read back
    for op in OPS:
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
    path: the pseudo-inverses (F0 F1) and translations agree to the conditioning of the
    check-to-equivalent systems, and FMM evaluations with either agree to that level."""
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
    ph = fh.operator("uc2ue0").reshape(K, -1) @ fh.operator("uc2ue1").reshape(-1, K)
    pd = fd.operator("uc2ue0").reshape(K, -1) @ fd.operator("uc2ue1").reshape(-1, K)
    e_p = float(np.linalg.norm(pd - ph) / np.linalg.norm(ph))
    e_u = float(np.linalg.norm(fd.operator("u2u") - fh.operator("u2u")) / np.linalg.norm(fh.operator("u2u")))
    e_d = float(np.linalg.norm(fd.operator("d2d") - fh.operator("d2d")) / np.linalg.norm(fh.operator("d2d")))
    rng = np.random.default_rng(9)
    x = rng.random((6000, 3))
    k0 = 3 if kernel == "stokes_vel" else 1
    q = rng.standard_normal(k0 * 6000)
    uh = P.PtFMMTree(x, x, 64).setup(fh).evaluate(q)
    ud = P.PtFMMTree(x, x, 64).setup(fd).evaluate(q)
    e_fmm = float(np.linalg.norm(ud - uh) / np.linalg.norm(uh))
    print(f"{kernel} m={m} {prec}: host {th:.2f} s, device {td:.2f} s; pinv {e_p:.2e}, u2u {e_u:.2e}, "
          f"d2d {e_d:.2e}, FMM {e_fmm:.2e}")
    assert e_fmm < (1e-10 if prec == "fp64" else 1e-6)
    assert e_u < 1e-6 and e_d < 1e-6
