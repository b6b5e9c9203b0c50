"""The arithmetic the FP32 scalar V-list Hadamards rely on, checked on the CPU.

The 16-vector kernel (k_hadamard_vec16 in pvfmm_amd/csrc/kernels.hip) splits each FP32
spectrum value x by truncation into three bf16 parts, x = x0 + x1 + x2, and the host splits
each FP64 transfer-block value t into three bf16 parts by round-to-nearest of the remainders.
It keeps the six products x_i * t_j with i + j <= 2, each exact in FP32, summed in FP32.

The tiled kernel (k_hadamard_split) scales x by its run's power of two s (s times the run's
largest spectrum bound below 2^14) and t by tau_f (tau_f times the frequency's largest entry
below 2^14), splits both into two fp16 parts by round-to-nearest (x0 = RN(x), x1 = RN(x - x0)),
keeps x0 t0 + x0 t1 + x1 t0, each exact in FP32, summed in FP32, and divides by s tau_f.

The bf16 tests pin the three facts that make the first result FP32-accurate:
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


# ---------------------------------------------------------------- fp16 parts (k_hadamard_split)
def split_f16(x):
    """x0 = RN_fp16(x), x1 = RN_fp16(x - x0) (x - x0 formed exactly, as the kernel's fused
    multiply-add does), as float64."""
    x = np.asarray(x, dtype=np.float64)
    x0 = x.astype(np.float16).astype(np.float64)
    x1 = (x - x0).astype(np.float16).astype(np.float64)
    return x0, x1


def test_fp16_split_keeps_22_bits_and_products_are_exact():
    rng = np.random.default_rng(10)
    # scaled spectra: at most 2^14, down to 2^-3 (x1 stays a normal fp16 value)
    x = (rng.choice([-1.0, 1.0], 200000) * rng.uniform(1, 2, 200000) *
         np.exp2(rng.integers(-3, 14, 200000))).astype(np.float32).astype(np.float64)
    x0, x1 = split_f16(x)
    assert np.all(np.abs(x0) <= 2.0 ** 14) and np.all(np.abs(x0) < 65504)
    # x - x0 is exact in FP32 (what the kernel subtracts)
    r = (x.astype(np.float32) - x0.astype(np.float32)).astype(np.float64)
    assert np.array_equal(r, x - x0)
    assert np.all(np.abs(x - x0 - x1) <= np.abs(x) * 2.0 ** -22)
    # dropping the second part leaves 11 bits
    assert np.quantile(np.abs(x - x0) / np.abs(x), 0.5) > 2.0 ** -14
    # every product of fp16 values (subnormals included) is exact in FP32
    a = np.concatenate([x0, x1])
    b = rng.permutation(a)
    p32 = (a.astype(np.float32) * b.astype(np.float32)).astype(np.float64)
    assert np.array_equal(p32, a * b)


def _f16_dot(x, t, keep, s, tau):
    xs = split_f16(x.astype(np.float64) * s)
    ts = split_f16(t * tau)
    acc = np.zeros(x.shape[0], dtype=np.float32)
    for (i, j) in keep:
        prod = (xs[i] * ts[j]).astype(np.float32)  # exact
        for k in range(x.shape[1]):
            acc = (acc + prod[:, k]).astype(np.float32)
    return acc.astype(np.float64) / (s * tau)


def test_three_fp16_products_give_fp32_accuracy():
    rng = np.random.default_rng(11)
    n, k = 4000, 32
    # a run's spectra spread over 2^20 in magnitude below its bound b
    x = (rng.standard_normal((n, k)) * np.exp2(rng.uniform(-20, 0, (n, k)))).astype(np.float32)
    t = rng.standard_normal((n, k)) * 1e-3
    b = np.abs(x.astype(np.float64)).max()
    s = 2.0 ** (14 - np.frexp(b)[1])  # s b < 2^14 <= 2 s b
    tau = 2.0 ** (14 - np.frexp(np.abs(t).max())[1])
    assert s * b < 2.0 ** 14 <= 2 * s * b
    exact = np.sum(x.astype(np.float64) * t, axis=1)
    scale = np.sum(np.abs(x.astype(np.float64) * t), axis=1)
    three = [(1, 0), (0, 1), (0, 0)]
    err = np.abs(_f16_dot(x, t, three, s, tau) - exact) / scale
    fp32 = np.zeros(n, dtype=np.float32)
    for j in range(k):
        fp32 = (fp32 + (x[:, j] * t[:, j].astype(np.float32)).astype(np.float32)).astype(np.float32)
    err32 = np.abs(fp32.astype(np.float64) - exact) / scale
    # FP32-level: within a small factor of the FP32 dot product's own error
    assert err.mean() < 2 * err32.mean()
    assert err.max() < 16 * 2.0 ** -24 * k ** 0.5
    # dropping the second part of either operand costs 11 bits
    for keep in ([(0, 1), (0, 0)], [(1, 0), (0, 0)]):
        e = np.abs(_f16_dot(x, t, keep, s, tau) - exact) / scale
        assert e.mean() > 50 * err.mean()
