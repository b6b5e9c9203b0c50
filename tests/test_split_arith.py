"""The arithmetic the FP32 scalar V-list Hadamard relies on, checked on the CPU.

The kernel (k_hadamard_split in pvfmm_amd/csrc/kernels.hip) splits each staged FP32 spectrum
value x by truncation into three bf16 parts, x = x0 + x1 + x2, and the host splits each FP64
transfer-block value t into three bf16 parts by round-to-nearest of the remainders.  It keeps
the six products x_i * t_j with i + j <= 2, each exact in FP32, summed in FP32.  These tests
pin the three facts that make the result FP32-accurate:
  * the truncation split is exact (x0 + x1 + x2 == x for every finite normal FP32 x);
  * a product of two bf16 values is exact in FP32;
  * the six-product sum of a dot product of length 32 (one MFMA's k) is within a few FP32
    roundings of the exact value, while dropping any one of the three parts of x is not.
A mistake such as keeping x0*t0 + x0*t1 + x1*t0 only, or truncating t instead of carrying its
remainder, fails the last test.
"""
import numpy as np


def trunc_bf16(x):
    """bf16 truncation of float32 values (the high 16 bits), as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32)
    return (u & np.uint32(0xFFFF0000)).view(np.float32)


def rn_bf16(x):
    """bf16 round-to-nearest-even of float32 values, as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split_trunc(x):
    x = np.asarray(x, dtype=np.float32)
    x0 = trunc_bf16(x)
    r1 = (x - x0).astype(np.float32)
    x1 = trunc_bf16(r1)
    r2 = (r1 - x1).astype(np.float32)
    x2 = trunc_bf16(r2)
    return x0, x1, x2, r2


def split_rn(t):
    t = np.asarray(t, dtype=np.float64)
    parts = []
    r = t.copy()
    for _ in range(3):
        h = rn_bf16(r.astype(np.float32))
        parts.append(h)
        r = r - h.astype(np.float64)
    return parts


def random_f32(rng, n):
    mant = rng.uniform(1.0, 2.0, n)
    expo = rng.integers(-60, 60, n)
    sign = rng.choice([-1.0, 1.0], n)
    return (sign * mant * np.exp2(expo)).astype(np.float32)


def test_truncation_split_is_exact():
    rng = np.random.default_rng(7)
    x = np.concatenate([random_f32(rng, 200000),
                        np.array([1.0, -1.0, 3.0, np.float32(1 / 3), np.finfo(np.float32).max,
                                  np.finfo(np.float32).tiny * 2 ** 20], dtype=np.float32)])
    x0, x1, x2, r2 = split_trunc(x)
    # every part is a bf16 value
    for p in (x0, x1, x2):
        assert np.all((p.view(np.uint32) & 0xFFFF) == 0)
    # the last remainder has at most 8 significant bits, so its truncation loses nothing
    assert np.array_equal(x2, r2)
    total = x0.astype(np.float64) + x1.astype(np.float64) + x2.astype(np.float64)
    assert np.array_equal(total, x.astype(np.float64))
    # magnitudes: |x1| <= 2^-7 |x|, |x2| <= 2^-15 |x|
    ax = np.abs(x.astype(np.float64))
    assert np.all(np.abs(x1) <= ax * 2.0 ** -7)
    assert np.all(np.abs(x2) <= ax * 2.0 ** -15)


def test_bf16_products_are_exact_in_fp32():
    rng = np.random.default_rng(8)
    # exponents within +-60 keep the products normal
    a = rn_bf16(random_f32(rng, 100000) * np.float32(2.0 ** -10))
    b = rn_bf16(random_f32(rng, 100000) * np.float32(2.0 ** 10))
    p32 = (a * b).astype(np.float32)
    p64 = a.astype(np.float64) * b.astype(np.float64)
    assert np.array_equal(p32.astype(np.float64), p64)


def _six_product_dot(x, t, keep):
    xs = split_trunc(x)[:3]
    ts = split_rn(t)
    acc = np.zeros(x.shape[0], dtype=np.float32)
    for (i, j) in keep:
        prod = (xs[i].astype(np.float32) * ts[j].astype(np.float32)).astype(np.float32)
        for k in range(x.shape[1]):
            acc = (acc + prod[:, k]).astype(np.float32)
    return acc


def test_six_products_give_fp32_accuracy():
    rng = np.random.default_rng(9)
    n, k = 4000, 32
    x = rng.standard_normal((n, k)).astype(np.float32)
    t = rng.standard_normal((n, k))
    exact = np.sum(x.astype(np.float64) * t, axis=1)
    scale = np.sum(np.abs(x.astype(np.float64) * t), axis=1)
    six = [(2, 0), (1, 1), (0, 2), (1, 0), (0, 1), (0, 0)]
    err = np.abs(_six_product_dot(x, t, six) - exact) / scale
    # k FP32 additions: a few units of 2^-24 relative to the sum of |terms|
    assert err.max() < 8 * 2.0 ** -24 * k ** 0.5
    # the FP32 dot product itself is no better
    fp32 = np.zeros(n, dtype=np.float32)
    for j in range(k):
        fp32 = (fp32 + (x[:, j] * t[:, j].astype(np.float32)).astype(np.float32)).astype(np.float32)
    err32 = np.abs(fp32 - exact) / scale
    assert err.mean() < 2 * err32.mean()
    # dropping the third part of either operand costs 8 bits
    for keep in ([(1, 1), (0, 2), (1, 0), (0, 1), (0, 0)],     # no x2
                 [(2, 0), (1, 1), (1, 0), (0, 1), (0, 0)],     # no t2
                 [(0, 0), (0, 1), (1, 0)]):                    # two parts each
        e = np.abs(_six_product_dot(x, t, keep) - exact) / scale
        assert e.mean() > 20 * err.mean()
