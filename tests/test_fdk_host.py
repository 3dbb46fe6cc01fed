"""CPU tests pinning the FDK pre-processing (ctsim_amd/fdk.py) against closed forms.

designFilter2 samples the Ram-Lak kernel h(0) = 1/(4 g^2), h(odd n) = -1/(n pi g)^2; the
infinite sequence has the response |w| / (2 pi g^2) on [-pi, pi], so the truncated one
must follow that ramp; filterFreq is a circular convolution of the zero-padded row
with ifft(H); Parker weights of conjugate rays add up to one."""
import math

import numpy as np
import pytest

from ctsim_amd import fdk
from oracle import fdk_ref
from ctsim_amd.geometry import Geometry, circular_betas


def test_nextpow2_matches_matlab():
    assert [fdk.nextpow2(n) for n in (1, 2, 3, 1023, 1024, 1025, 2047)] == [0, 1, 2, 10, 10, 11, 11]


@pytest.mark.parametrize("g", [0.388, 1.0, 0.75])
def test_ramlak_response_is_the_ramp(g):
    n = 1024
    H = fdk.design_filter2(g, "ram-lak", n)
    order = H.size
    assert order == 2048
    w = 2 * math.pi * np.arange(order // 2 + 1) / order
    ramp = w / (2 * math.pi * g * g)
    # truncation of the 1/n^2 tail moves the response by ~1/(order g^2); compare away from DC
    k = np.arange(order // 16, order // 2 + 1)
    np.testing.assert_allclose(H[k], ramp[k], rtol=2e-3)
    assert H[order // 2] == pytest.approx(1 / (2 * g * g), rel=1e-3)
    np.testing.assert_allclose(H[1:], H[1:][::-1])  # symmetric spectrum


def test_windows_and_crop():
    n, g = 512, 0.5
    ram = fdk.design_filter2(g, "ram-lak", n)
    order = ram.size
    w = 2 * math.pi * np.arange(order) / order
    hann = fdk.design_filter2(g, "hann", n)
    half = order // 2 + 1
    np.testing.assert_allclose(hann[1:half], ram[1:half] * (1 + np.cos(w[1:half])) / 2, rtol=1e-12)
    ham = fdk.design_filter2(g, "hamming", n)
    np.testing.assert_allclose(ham[1:half], ram[1:half] * (0.54 + 0.46 * np.cos(w[1:half])), rtol=1e-12)
    cropped = fdk.design_filter2(g, "ram-lak", n, crop=0.5)
    assert np.all(cropped[:half][w[:half] > math.pi * 0.5] == 0)
    with pytest.raises(ValueError):
        fdk.design_filter2(g, "nonsense", n)


def test_filter_rows_is_circular_convolution():
    rng = np.random.default_rng(0)
    n = 100
    H = fdk.design_filter2(1.0, "hann", n)
    x = rng.standard_normal((3, n))
    got = fdk_ref.filter_rows(x, H)
    kern = np.real(np.fft.ifft(H))
    order = H.size
    for r in range(3):
        pad = np.zeros(order)
        pad[:n] = x[r]
        conv = np.array([np.sum(pad * np.roll(kern[::-1], k + 1)) for k in range(order)])
        np.testing.assert_allclose(got[r], conv[:n], atol=1e-10)


def test_parker_weights_of_conjugate_rays_sum_to_one():
    """For a ray of fan angle g (the weight's own sign convention, g = -gamma) scanned over
    exactly pi - 2g, the weight of (beta, g) and of its conjugate (beta + pi + 2g, -g) add up
    to one; a sign slip in either use of g breaks the identity."""
    sdd = 1000.0
    u = np.linspace(-150, 150, 13)
    gam = np.arctan(-u / sdd)
    for gm in gam:
        g = -gm
        tau = math.pi - 2 * g
        for beta in np.linspace(-2 * g, math.pi - 2 * g, 9):
            w1 = fdk.parker_weight(np.array([gm]), beta, tau)[0]
            w2 = fdk.parker_weight(np.array([-gm]), beta + math.pi + 2 * g, tau)[0]
            assert 0.0 <= w1 <= 1.0
            assert w1 + w2 == pytest.approx(1.0, abs=1e-12)
    w0 = fdk.parker_weight(gam, 0.0, math.pi + 2 * np.max(np.abs(gam)))
    assert np.all(w0 == 0.0)


def test_fdk_scale_closed_form():
    g = Geometry(reconSize=(4, 4, 4), reconSpacing=(1, 1, 1), reconOffset=(0, 0, 0), detSize=(8, 8),
                 detSpacing=(1, 1), detOffset=(0, 0), SAD=500, ADD=500, betas=circular_betas(360))
    dbeta = abs(g.betas[-1] - g.betas[0]) / 359
    assert fdk.fdk_scale(g) == pytest.approx(dbeta / 2 * 10)
    g.shortScan = 1
    assert fdk.fdk_scale(g) == pytest.approx(dbeta * 10)


def test_equiangular_kernel_tends_to_the_flat_one():
    """designEquiangularFilter2's kernel -1/(SDD pi sin(n g / SDD))^2 becomes designFilter2's
    -1/(n pi g)^2 as SDD grows; at a finite SDD it differs at the high indices"""
    g, n = 0.5, 256
    flat = fdk.design_filter2(g, "hann", n)
    far = fdk.design_filter2(g, "hann", n, sdd=1e7)
    np.testing.assert_allclose(far, flat, rtol=1e-8, atol=1e-12)
    near = fdk.design_filter2(g, "hann", n, sdd=300.0)
    assert np.max(np.abs(near - flat)) > 1e-4 * np.max(np.abs(flat))


def test_sliver_weights_of_conjugate_rays_sum_to_one():
    """shortScanSliverWeight: the ray (beta, gamma) and its conjugate (beta + pi - 2 gamma,
    -gamma) carry weights that add to one over the short scan"""
    from ctsim_amd.fdk import sliver_weight
    rng = np.random.default_rng(3)
    tau = 0.4
    for _ in range(200):
        g = rng.uniform(-0.15, 0.15)
        beta = rng.uniform(0, math.pi + 2 * g)
        w1 = sliver_weight(np.array([g]), beta, tau)[0]
        w2 = sliver_weight(np.array([-g]), beta + math.pi - 2 * g, tau)[0]
        assert w1 + w2 == pytest.approx(1.0, abs=1e-12), (g, beta, w1, w2)


@pytest.mark.parametrize("flat", [1, 0])
def test_fdk_cosine_weight_special_cases(flat):
    """With an all-pass filter (H = 1) fdk_filter returns the cosine weight of reconFDK.m:39
    (flat) / :44 (arc).  On the mid row (v = 0) the flat weight reduces to SAD/SDD for every u
    (the row-wise form, not the textbook SAD/sqrt(u^2+v^2+SDD^2)), and on the mid column
    (u = 0) it is SAD/sqrt(v^2 + SDD^2); the arc weight on the mid row is SAD/SDD cos(u/SDD)
    and on the mid column SAD/SDD."""
    nu, nv = 41, 33
    g = Geometry(reconSize=(8, 8, 8), reconSpacing=(1, 1, 1), reconOffset=(0, 0, 0), detSize=(nu, nv),
                 detSpacing=(3.0, 4.0), detOffset=(0, 0), SAD=400.0, ADD=250.0, betas=circular_betas(4),
                 flatPanel=flat)
    one = np.ones((nv, nu, 1))
    wgt = fdk_ref.fdk_filter(one, g, np.ones(64))[:, :, 0]
    u = (np.arange(nu) - (nu - 1) / 2) * 3.0
    v = (np.arange(nv) - (nv - 1) / 2) * 4.0
    mid_v, mid_u = (nv - 1) // 2, (nu - 1) // 2
    if flat:
        np.testing.assert_allclose(wgt[mid_v, :], 400.0 / 650.0, rtol=1e-13)
        np.testing.assert_allclose(wgt[:, mid_u], 400.0 / np.sqrt(v ** 2 + 650.0 ** 2), rtol=1e-13)
        assert wgt[0, mid_u] < wgt[mid_v, mid_u]
    else:
        np.testing.assert_allclose(wgt[mid_v, :], 400.0 / 650.0 * np.cos(u / 650.0), rtol=1e-13)
        # cos(u / sqrt(SDD^2 + v^2)) is one on the mid column whatever v is
        np.testing.assert_allclose(wgt[:, mid_u], 400.0 / 650.0, rtol=1e-13)
        assert wgt[0, 0] > wgt[mid_v, 0]
