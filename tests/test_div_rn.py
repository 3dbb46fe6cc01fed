"""The ray-driven kernels divide by the voxel size as float(double(x) * (1 / double(d))).
For float x and d that rounds to the correctly rounded float quotient: the double product
is within 2^-52 (relative) of x / d, and a quotient of two floats is never closer than
2^-50 to a float rounding midpoint.  Checked here against numpy's IEEE float32 division on
random pairs over wide exponent ranges, on voxel sizes like the kernels', and on quotients
that land exactly on floats."""
import numpy as np


def div_rn(x, d):
    return (x.astype(np.float64) * (1.0 / d.astype(np.float64))).astype(np.float32)


def test_div_rn_equals_float_division_on_random_pairs():
    rng = np.random.default_rng(7)
    n = 4_000_000
    x = (rng.uniform(-1, 1, n) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)
    d = (rng.uniform(0.5, 1, n) * np.exp2(rng.integers(-10, 10, n))).astype(np.float32)
    want = x / d
    assert want.dtype == np.float32
    assert np.array_equal(div_rn(x, d), want)


def test_div_rn_on_voxel_sizes_and_exact_quotients():
    rng = np.random.default_rng(8)
    sizes = np.array([0.1, 0.25, 0.3, 0.4, 0.5, 0.6, 0.7, 0.9, 1.0, 1.1, 1.3, 2.0, 0.390625, 0.7734375],
                     np.float32)
    x = (rng.uniform(-300, 300, 2_000_000)).astype(np.float32)
    for d in sizes:
        dd = np.full_like(x, d)
        assert np.array_equal(div_rn(x, dd), x / dd), float(d)
        # x = k * d exactly representable: the quotient is k
        k = rng.integers(-4000, 4000, 100_000).astype(np.float32)
        xe = (k * d).astype(np.float32)
        de = np.full_like(xe, d)
        assert np.array_equal(div_rn(xe, de), xe / de), float(d)
    z = np.array([0.0, -0.0], np.float32)
    assert np.array_equal(np.signbit(div_rn(z, np.ones(2, np.float32))), np.signbit(z))
