"""GPU data side: Poisson counts (distribution checks), the log attenuation, the PWLS
weights and the water beam-hardening polynomial (exact against numpy restatements of the
reference formulas)."""
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


def _water_f32(p, coeffs, nv, nu):
    """The correction in float32, one rounding per operation, in the reference's order
    (beamHardeningWarterCorrection.m:56-60: p c_1, then + p^k c_k for k = 2 .. order)."""
    p = np.asarray(p, np.float32).reshape(-1, nu, nv)
    c = np.asarray(coeffs, np.float32)
    c = np.broadcast_to(c[:, None, :], (nu, nv, c.shape[1])) if c.shape[0] > 1 else np.broadcast_to(
        c[0], (nu, nv, c.shape[1]))
    acc = (p * c[..., 0]).astype(np.float32)
    pk = p.copy()
    for k in range(1, c.shape[-1]):
        pk = (pk * p).astype(np.float32)
        acc = (acc + (pk * c[..., k]).astype(np.float32)).astype(np.float32)
    return acc.reshape(-1)


@pytest.mark.parametrize("nv,nu,nviews,order,per_channel", [(1, 367, 29, 3, True), (16, 89, 7, 3, True),
                                                            (16, 89, 7, 4, False), (1, 5, 3, 1, True)])
@pytest.mark.parametrize("in_dt", ["f32", "f16", "bf16"])
def test_water_correction_matches_float32_evaluation(nv, nu, nviews, order, per_channel, in_dt):
    from ctsim_amd.beamhardening import water_correction
    from ctsim_amd.device import DeviceArray, bf16_from_f32, f32_from_bf16
    rng = np.random.default_rng(nv * 1000 + nu + order)
    p = rng.uniform(0.0, 9.0, nv * nu * nviews).astype(np.float32)
    p[:7] = [0.0, 1e-30, 0.5, 1.0, 8.999, 3.25, 2.0][:min(7, p.size)]
    if in_dt == "f16":
        p = p.astype(np.float16).astype(np.float32)
    elif in_dt == "bf16":
        p = f32_from_bf16(bf16_from_f32(p))
    coeffs = np.stack([np.r_[1.0 + 0.01 * u, rng.uniform(-0.02, 0.05, order - 1)] for u in range(nu)]).astype(
        np.float32)
    if not per_channel:
        coeffs = coeffs[:1]
    want = _water_f32(p, coeffs, nv, nu)
    d = DeviceArray.from_numpy(p, in_dt)
    got = water_correction(d, nv, nu, nviews, coeffs, out=DeviceArray(p.size, "f32")).to_numpy()
    np.testing.assert_array_equal(got, want)
    # narrower output: the float32 result rounded once
    got16 = water_correction(d, nv, nu, nviews, coeffs, out=DeviceArray(p.size, "f16")).to_numpy()
    np.testing.assert_array_equal(got16, want.astype(np.float16))
    gotb = water_correction(d, nv, nu, nviews, coeffs, out=DeviceArray(p.size, "bf16")).to_numpy()
    np.testing.assert_array_equal(gotb, f32_from_bf16(bf16_from_f32(want)))
    # in place over the input
    water_correction(d, nv, nu, nviews, coeffs, out=d)
    stored = {"f32": want, "f16": want.astype(np.float16), "bf16": f32_from_bf16(bf16_from_f32(want))}[in_dt]
    np.testing.assert_array_equal(d.to_numpy(), stored)


def test_water_correction_with_fitted_coefficients_linearises_water():
    """The fitted polynomial of a filtered 120 kVp spectrum, applied on the device to the
    polychromatic line integrals of water slabs, returns L mu_eff to the fit's accuracy."""
    import os
    from conftest import GOLDEN
    from ctsim_amd.beamhardening import (interp_geometric, spectral_attenuation, water_coefficient_table,
                                         water_correction)
    from ctsim_amd.device import DeviceArray
    with np.load(os.path.join(GOLDEN, "material_tables.npz"), allow_pickle=False) as z:
        t = {k: z[k] for k in z.files}
    e = np.arange(20.0, 121.0)
    ph = (120.0 - e) / e * (1 - np.exp(-((e - 15.0) / 12.0) ** 2)) * 1e6
    eeff = float(np.sum(e * ph) / np.sum(ph))
    nu, nviews = 24, 5
    thick = np.linspace(0.0, 30.0, nu)
    tab = water_coefficient_table(e, ph, t["water_E"], t["water_mu"], eeff,
                                  bowtie=(t["aluminum_E"], t["aluminum_mu"], thick))
    m = spectral_attenuation(e, t["water_E"], t["water_mu"])
    meff = float(interp_geometric(e, m, [eeff])[0])
    L = np.linspace(1.0, 40.0, nviews)
    sino = np.empty((nviews, nu), np.float32)
    for iv in range(nviews):
        for u in range(nu):
            pc = ph * e * np.exp(-spectral_attenuation(e, t["aluminum_E"], t["aluminum_mu"]) * thick[u] / 10.0)
            sino[iv, u] = -np.log(np.sum(pc * np.exp(-L[iv] * m)) / np.sum(pc))
    got = water_correction(DeviceArray.from_numpy(sino, "f32"), 1, nu, nviews, tab).to_numpy().reshape(nviews, nu)
    np.testing.assert_allclose(got, np.broadcast_to((L * meff)[:, None], got.shape), atol=5e-3 * 50.0 * meff)
    # the device result is the float32 polynomial of those coefficients
    np.testing.assert_array_equal(got.reshape(-1), _water_f32(sino.reshape(-1), tab.astype(np.float32), 1, nu))


def test_process_ct_raw_data_chains_log_and_water():
    from ctsim_amd import sim
    from ctsim_amd.device import DeviceArray
    rng = np.random.default_rng(9)
    nviews, nu, nv, i0 = 6, 41, 4, 2e4
    raw = rng.poisson(i0 * 0.1, (nviews, nu, nv)).astype(np.float32)
    prof = rng.uniform(0.5, 2.0, nviews).astype(np.float32)
    coeffs = np.stack([[1.0 + 0.002 * u, 0.03, -0.001] for u in range(nu)]).astype(np.float32)
    d_raw, d_prof = DeviceArray.from_numpy(raw, "f32"), DeviceArray.from_numpy(prof, "f32")
    y = sim.log_attenuation(d_raw, nu * nv, i0, tube_profile=d_prof).to_numpy()
    got = sim.process_ct_raw_data(d_raw, nv, nu, i0, tube_profile=d_prof, water_coeffs=coeffs).to_numpy()
    assert got.shape == raw.shape
    np.testing.assert_array_equal(got.reshape(-1), _water_f32(y, coeffs, nv, nu))
    plain = sim.process_ct_raw_data(d_raw, nv, nu, i0, tube_profile=d_prof).to_numpy()
    np.testing.assert_array_equal(plain, y)
