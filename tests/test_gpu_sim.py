"""GPU data side: Poisson counts (distribution checks), the log attenuation and the PWLS
weights (exact against numpy restatements of the reference formulas)."""
import math

import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu


def _counts(p, i0, seed, dtype="f32"):
    import ctypes as C
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    d_p = DeviceArray.from_numpy(p, dtype)
    d_c = DeviceArray(p.size, "f32")
    call("ctsim_poisson_counts", d_p.vp, d_p.dtype, p.size, float(i0), C.c_uint64(seed), d_c.vp, None)
    return d_c.to_numpy()


@pytest.mark.parametrize("lam", [0.3, 3.0, 9.5, 10.0, 47.0, 2e4, 1e5])
def test_poisson_mean_variance_and_shape(lam):
    n = 400000
    p = np.full(n, -math.log(lam / 1e5), np.float64)  # i0 exp(-p) = lam
    c = _counts(p, 1e5, 7, "f64").astype(np.float64)
    assert np.all(c >= 0) and np.all(c == np.round(c))
    se = math.sqrt(lam / n)
    assert abs(c.mean() - lam) < 5 * se, (c.mean(), lam)
    assert abs(c.var() / lam - 1) < 5 * math.sqrt(2.0 / n) + 1e-3
    if lam < 60:  # chi-square against the Poisson pmf on the bulk of the support
        ks = np.arange(int(max(0, lam - 5 * math.sqrt(lam))), int(lam + 5 * math.sqrt(lam)) + 2)
        obs = np.array([(c == k).sum() for k in ks], np.float64)
        exp = n * stats.poisson.pmf(ks, lam)
        keep = exp > 20
        chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
        assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, chi2
    else:  # standardised skewness of a Poisson variate is 1/sqrt(lam)
        z = (c - lam) / math.sqrt(lam)
        assert abs((z ** 3).mean() - 1 / math.sqrt(lam)) < 0.03


def test_poisson_depends_on_seed_and_index_only():
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 6, 100001).astype(np.float32)
    a = _counts(p, 2e4, 11)
    b = _counts(p, 2e4, 11)
    np.testing.assert_array_equal(a, b)
    part = _counts(p[:5000], 2e4, 11)
    np.testing.assert_array_equal(part, a[:5000])
    assert not np.array_equal(a, _counts(p, 2e4, 12))


def test_log_attenuation_and_weights_match_reference_formulas():
    import ctypes as C
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    rng = np.random.default_rng(2)
    counts = rng.poisson(50, 50001).astype(np.float32)
    counts[:100] = 0
    d_c = DeviceArray.from_numpy(counts, "f32")
    y = DeviceArray(counts.size, "f64")
    call("ctsim_log_attenuation", d_c.vp, counts.size, 2e4, 10.0, y.vp, y.dtype, None)
    want = -np.log(np.maximum(counts.astype(np.float64), 10.0) / 2e4)
    np.testing.assert_allclose(y.to_numpy(), want, rtol=1e-14)
    w = DeviceArray(counts.size, "f32")
    call("ctsim_pwls_weights", d_c.vp, counts.size, 2.0, 5.0, w.vp, w.dtype, None)
    s = counts
    ww = ((s - np.float32(2.0)) ** 2 / (s + np.float32(5.0))).astype(np.float32)
    ww[s <= 2.0] = 0
    np.testing.assert_allclose(w.to_numpy(), ww, rtol=2e-7)


@pytest.mark.parametrize("compound", [False, True])
@pytest.mark.parametrize("with_ff,with_prof", [(False, False), (True, False), (False, True), (True, True)])
def test_log_attenuation_ex_matches_the_raw_data_restatement(compound, with_ff, with_prof):
    from oracle import sim_ref
    from ctsim_amd import sim
    from ctsim_amd.device import DeviceArray
    rng = np.random.default_rng(4)
    nv, nc, i0 = 37, 211, 3e4
    prof = rng.uniform(0.5, 2.0, nv).astype(np.float32)
    ff = rng.uniform(0.2, 1.0, nc).astype(np.float32)
    raw = rng.poisson(i0 * 0.05, (nv, nc)).astype(np.float32)
    raw[:, :20] = rng.integers(0, 150, (nv, 20)).astype(np.float32)  # the clamp region
    low = 100.0 if compound else 10.0
    d_raw = DeviceArray.from_numpy(raw, "f32")
    d_ff = DeviceArray.from_numpy(ff, "f32") if with_ff else None
    d_prof = DeviceArray.from_numpy(prof, "f32") if with_prof else None
    want = sim_ref.process_raw(raw, i0, prof if with_prof else None, ff if with_ff else None, compound, low)
    y = sim.log_attenuation(d_raw, nc, i0, compound, None, d_ff, d_prof, "f64").to_numpy()
    np.testing.assert_allclose(y, want, rtol=1e-14, atol=1e-14)
    y32 = sim.log_attenuation(d_raw, nc, i0, compound, None, d_ff, d_prof, "f32").to_numpy()
    np.testing.assert_allclose(y32, want.astype(np.float32), rtol=2e-7, atol=1e-7)
