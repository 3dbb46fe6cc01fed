"""Host side of the raw-data log: the oracle pinned by closed forms, and source_intensity
against the formulas of the two reference log functions."""
import math

import numpy as np

from oracle import sim_ref
from ctsim_amd import sim


def _raw(nv=7, nc=33, i0=2e4, seed=3):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.0, 4.0, (nv, nc))
    prof = rng.uniform(0.5, 2.0, nv)
    ff = rng.uniform(0.2, 1.0, nc)
    return p, prof, ff, i0


def test_noise_free_counts_give_back_the_line_integrals():
    # counts = profile * i0 * ff * exp(-p): both models recover p while no clamp applies
    p, prof, ff, i0 = _raw()
    raw = prof[:, None] * i0 * ff[None, :] * np.exp(-p)
    for compound in (False, True):
        y = sim_ref.process_raw(raw, i0, prof, ff, compound, low=1e-3)
        np.testing.assert_allclose(y, p, rtol=0, atol=1e-12)


def test_clamp_order_differs_between_the_models():
    # simple clamps the counts before the bowtie division, compound clamps the ratio after it
    _, _, ff, i0 = _raw()
    raw = np.zeros((2, ff.size))
    ys = sim_ref.process_raw(raw, i0, None, ff, False, low=10.0)
    yc = sim_ref.process_raw(raw, i0, None, ff, True, low=100.0)
    np.testing.assert_allclose(ys, np.broadcast_to(-np.log(10.0 / i0) + np.log(ff), ys.shape), rtol=1e-14)
    np.testing.assert_allclose(yc, np.full(yc.shape, -math.log(100.0 / i0)), rtol=1e-14)


def test_profile_applies_before_the_clamp():
    i0 = 1e4
    raw = np.array([[40.0], [40.0]])
    y = sim_ref.process_raw(raw, i0, np.array([1.0, 8.0]), None, False, low=10.0)
    assert math.isclose(y[0, 0], -math.log(40.0 / i0)) and math.isclose(y[1, 0], -math.log(10.0 / i0))


def test_source_intensity_branches():
    e = np.array([30.0, 50.0, 70.0])
    ph = np.array([1e3, 4e3, 2e3])
    assert sim.source_intensity([5e4], [60.0], dqe=0.8, photons_total=7e4) == 0.8 * 7e4
    assert sim.source_intensity([5e4], [60.0], dqe=0.8) == 0.8 * 5e4
    assert math.isclose(sim.source_intensity(ph, e, dqe=0.9, gain=0.1), 0.9 * (3e4 + 2e5 + 1.4e5) * 0.1)
    assert math.isclose(sim.source_intensity(ph, e, dqe=0.9, gain=0.1, energy_integrating=False), 0.9 * 7e3 * 0.1)
    mu = np.array([20.0, 6.0, 2.5])
    want = sum(0.9 * ph[i] * (1 - math.exp(-1.5 * mu[i])) * e[i] * 0.1 for i in range(3))
    assert math.isclose(sim.source_intensity(ph, e, gain=0.1, compound=True, scintillator_mu=mu), want)
