"""GPU FDK pre-processing and reconstruction against the FP64 numpy restatement
(oracle/fdk_ref.py) and the FP64 oracle backprojector.

Tolerances: FP32 filter within 1e-5 relative L2 of the FP64 filter (FFT rounding grows
like log2 of the 2048-point length); FP16 storage of the filtered views within 1e-3
(11-bit significand); the FP64 path within 1e-11."""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _geom(flat=1, short=False, nviews=24, nuv=(100, 30), n=(40, 40, 20)):
    from ctsim_amd.geometry import Geometry
    span = (np.pi + 2 * np.arctan(nuv[0] * 0.4 / 2 / 1500.0) + 0.05) if short else 2 * np.pi
    betas = np.linspace(0, span, nviews, endpoint=short) + 0.2
    return Geometry(reconSize=n, reconSpacing=(0.5, 0.5, 0.5), reconOffset=(0, 0, 0), detSize=nuv,
                    detSpacing=(0.4 if flat else 0.4, 0.5), detOffset=(0.7, -1.3), SAD=1000.0, ADD=500.0,
                    SDD=1500.0, betas=betas, flatPanel=flat, shortScan=int(short))


@pytest.mark.parametrize("flat,short", [(1, False), (1, True), (0, False), (0, True)])
def test_fdk_filter_matches_fp64_restatement(flat, short):
    from oracle import fdk_ref
    from ctsim_amd import fdk
    from ctsim_amd.device import DeviceArray
    g = _geom(flat, short)
    rng = np.random.default_rng(3 + flat + 2 * short)
    sino = rng.uniform(0, 3, (g.detSize[1], g.detSize[0], g.noViews)).astype(np.float32)
    H = fdk.fdk_filter_response(g, "hann")
    want = fdk_ref.fdk_filter(sino, g, H, fdk.fdk_view_weights(g))
    d_in = DeviceArray.from_numpy(sino.ravel(order="F"), "f32")
    f32 = fdk.FdkFilter(g, "hann")
    got = f32(d_in, DeviceArray(d_in.count, "f32")).to_numpy().reshape(sino.shape, order="F")
    assert rel_l2(got, want) < 1e-5, rel_l2(got, want)
    got16 = f32(d_in, DeviceArray(d_in.count, "f16")).to_numpy().reshape(sino.shape, order="F")
    assert rel_l2(got16, want) < 1e-3
    # a view range: views 5..12 only, weights indexed by the plan's view number
    d_part = DeviceArray.from_numpy(np.asfortranarray(sino[:, :, 5:13]).ravel(order="F"), "f32")
    part = f32(d_part, DeviceArray(d_part.count, "f32"), view0=5, nsub=8).to_numpy()
    assert rel_l2(part.reshape((g.detSize[1], g.detSize[0], 8), order="F"), want[:, :, 5:13]) < 1e-5
    f64 = fdk.FdkFilter(g, "hann", compute="f64")
    d64 = DeviceArray.from_numpy(sino.astype(np.float64).ravel(order="F"), "f64")
    got64 = f64(d64, DeviceArray(d64.count, "f64")).to_numpy().reshape(sino.shape, order="F")
    assert rel_l2(got64, want) < 1e-11


@pytest.mark.parametrize("nu", [100, 200, 500, 1000, 60])
def test_fdk_filter_lengths_and_ragged_rows(nu):
    """the fused filter at each filter length it takes (256 .. 2048 taps for 100 .. 1000 cells)
    and the hipFFT chain below it (60 cells: 128 taps), with a row count that is not a
    multiple of the fused form's eight-row tiles, against the FP64 restatement"""
    from oracle import fdk_ref
    from ctsim_amd import fdk
    from ctsim_amd.device import DeviceArray
    g = _geom(flat=nu % 2, short=True, nviews=5, nuv=(nu, 13))
    rng = np.random.default_rng(nu)
    sino = rng.uniform(0, 3, (13, nu, 5)).astype(np.float32)
    H = fdk.fdk_filter_response(g, "hann")
    want = fdk_ref.fdk_filter(sino, g, H, fdk.fdk_view_weights(g))
    f32 = fdk.FdkFilter(g, "hann")
    for dt in ("f32", "f16", "bf16"):
        d_in = DeviceArray.from_numpy(sino.ravel(order="F"), dt)
        x = sino if dt == "f32" else d_in.to_numpy().reshape(sino.shape, order="F")
        ref = want if dt == "f32" else fdk_ref.fdk_filter(x.astype(np.float32), g, H, fdk.fdk_view_weights(g))
        got = f32(d_in, DeviceArray(d_in.count, "f32")).to_numpy().reshape(sino.shape, order="F")
        assert rel_l2(got, ref) < 1e-5, (dt, rel_l2(got, ref))


def test_fdk_filter_view_ranges_agree_with_whole():
    """filtering views in two ranges gives the bits of filtering them all at once"""
    from ctsim_amd import fdk
    from ctsim_amd.device import DeviceArray
    g = _geom(nviews=40, nuv=(64, 16))
    rng = np.random.default_rng(8)
    sino = rng.uniform(0, 1, (16, 64, 40)).astype(np.float32)
    f = fdk.FdkFilter(g, "ram-lak")
    d = DeviceArray.from_numpy(sino.ravel(order="F"), "f32")
    whole = f(d, DeviceArray(d.count, "f32")).to_numpy()
    for v0, ns in ((0, 17), (17, 23)):
        dp = DeviceArray.from_numpy(np.asfortranarray(sino[:, :, v0:v0 + ns]).ravel(order="F"), "f32")
        part = f(dp, DeviceArray(dp.count, "f32"), view0=v0, nsub=ns).to_numpy()
        np.testing.assert_array_equal(part, whole.reshape((16, 64, 40), order="F")[:, :, v0:v0 + ns].ravel(order="F"))


def test_recon_fdk_matches_oracle_chain():
    from oracle import fdk_ref
    from oracle import oracle as O
    from ctsim_amd import fdk
    g = _geom(nviews=30)
    rng = np.random.default_rng(5)
    sino = rng.uniform(0, 2, (g.detSize[1], g.detSize[0], g.noViews)).astype(np.float32)
    filt = fdk_ref.fdk_filter(sino, g, fdk.fdk_filter_response(g, "hamming"))
    g3 = O.G3(g.reconSize, g.reconSpacing, g.reconOffset, g.detSize, g.detSpacing, g.detOffset, g.SAD, g.ADD,
              g.SDD, g.betas, dettype=g.flatPanel)
    want = O.cbct("back,hpd", g3, filt, None, prec=64) * fdk.fdk_scale(g)
    got = fdk.recon_fdk(sino, g, "hamming")
    assert rel_l2(got, want) < 1e-5, rel_l2(got, want)
    got16 = fdk.recon_fdk(sino, g, "hamming", storage="f16")
    assert rel_l2(got16, want) < 1e-3
    got64 = fdk.recon_fdk(sino.astype(np.float64), g, "hamming", compute="f64", storage="f64")
    assert rel_l2(got64, want) < 1e-11


def test_fdk_of_a_uniform_cylinder_recovers_its_attenuation():
    """closed-form check of the whole chain (cosine weight, ramp filter scaled by du, hpd
    backprojection, dbeta * 10): the projectors take attenuation in 1/cm (path lengths
    dx/10), FDK returns 1/cm, so a uniform cylinder of attenuation mu reconstructs to mu in
    its interior central slices, within the few percent that discretisation and the cone
    angle leave, and to about zero outside it"""
    from oracle import oracle as O
    from ctsim_amd import fdk
    from ctsim_amd.geometry import Geometry, circular_betas
    n, nz = 64, 8
    g = Geometry(reconSize=(n, n, nz), reconSpacing=(1.0, 1.0, 1.0), reconOffset=(0, 0, 0), detSize=(128, 16),
                 detSpacing=(1.0, 1.0), detOffset=(0, 0), SAD=500.0, ADD=250.0, SDD=750.0,
                 betas=circular_betas(180), flatPanel=1)
    x = (np.arange(n) - (n - 1) / 2)
    X, Y = np.meshgrid(x, x, indexing="ij")
    mu = 0.02
    vol = np.repeat(((X ** 2 + Y ** 2) <= 20.0 ** 2)[:, :, None] * mu, nz, axis=2).astype(np.float32)
    g3 = O.G3(g.reconSize, g.reconSpacing, g.reconOffset, g.detSize, g.detSpacing, g.detOffset, g.SAD, g.ADD, g.SDD,
              g.betas)
    sino = O.cbct("proj,dd", g3, vol, None, prec=64).astype(np.float32)
    img = fdk.recon_fdk(sino, g, "hann")
    inner = (X ** 2 + Y ** 2) <= 12.0 ** 2
    val = img[:, :, nz // 2 - 1:nz // 2 + 1][inner].mean()
    assert val == pytest.approx(mu, rel=0.03), val
    outside = (X ** 2 + Y ** 2) >= 26.0 ** 2
    assert abs(img[:, :, nz // 2][outside].mean()) < 0.05 * mu


def _fan(nviews=120, short=False, flat=1):
    from ctsim_amd.geometry import Geometry
    span = (math.pi + 2 * 0.2) if short else 2 * math.pi
    betas = np.linspace(0, span, nviews, endpoint=not short) if short else np.arange(nviews) * (span / nviews)
    return Geometry(reconSize=(48, 40), reconSpacing=(1.0, 1.1), reconOffset=(0.5, -0.5), detSize=(96,),
                    detSpacing=(1.1,), detOffset=(0.3,), SAD=500.0, ADD=250.0, SDD=800.0, betas=betas,
                    flatPanel=flat, shortScan=short)


@pytest.mark.parametrize("flat", [1, 0])
@pytest.mark.parametrize("short", [False, True])
def test_recon_fbp2d_matches_oracle_chain(flat, short):
    from oracle import fdk_ref
    from oracle import oracle as O
    from ctsim_amd import fdk
    g = _fan(short=short, flat=flat)
    rng = np.random.default_rng(6)
    sino = rng.uniform(0, 2, (g.detSize[0], g.noViews)).astype(np.float32)
    filt = fdk_ref.fbp2d_filter(sino, g, fdk.fdk_filter_response(g, "hamming"), fdk.fdk_view_weights(g))
    g2 = O.G2(g.reconSize[0], g.reconSize[1], *g.reconSpacing, *g.reconOffset, g.detSize[0], g.detSpacing[0],
              g.detOffset[0], g.SAD, g.ADD, g.SDD, g.betas, flat)
    want = O.fbct("back,pd", g2, filt, prec=64) * fdk.fdk_scale(g)
    got = fdk.recon_fbp2d(sino, g, "hamming")
    assert rel_l2(got, want) < 1e-5, rel_l2(got, want)
    got64 = fdk.recon_fbp2d(sino.astype(np.float64), g, "hamming", compute="f64", storage="f64")
    # FP64 against FP64: the ramp-filtered views change sign and the 120-view sum cancels,
    # so the agreement sits a little above the per-operation 1e-16 (3e-11 measured)
    assert rel_l2(got64, want) < 1e-10, rel_l2(got64, want)
    got16 = fdk.recon_fbp2d(sino, g, "hamming", storage="f16")
    assert rel_l2(got16, want) < 1e-3


def test_fbp2d_of_a_uniform_disc_recovers_its_attenuation():
    from oracle import oracle as O
    from ctsim_amd import fdk
    from ctsim_amd.geometry import Geometry, circular_betas
    n = 96
    g = Geometry(reconSize=(n, n), reconSpacing=(1.0, 1.0), reconOffset=(0, 0), detSize=(192,), detSpacing=(1.0,),
                 detOffset=(0,), SAD=500.0, ADD=250.0, SDD=750.0, betas=circular_betas(360), flatPanel=1)
    x = (np.arange(n) - (n - 1) / 2)
    X, Y = np.meshgrid(x, x, indexing="ij")
    mu = 0.02
    img0 = (((X ** 2 + Y ** 2) <= 30.0 ** 2) * mu).astype(np.float32)
    g2 = O.G2(n, n, 1.0, 1.0, 0, 0, 192, 1.0, 0, 500.0, 250.0, 750.0, g.betas, 1)
    sino = O.fbct("proj,dd", g2, img0, prec=64).astype(np.float32)
    img = fdk.recon_fbp2d(sino, g, "hann")
    inner = (X ** 2 + Y ** 2) <= 20.0 ** 2
    assert img[inner].mean() == pytest.approx(mu, rel=0.02), img[inner].mean()
    outside = (X ** 2 + Y ** 2) >= 36.0 ** 2
    assert abs(img[outside].mean()) < 0.05 * mu
