"""GPU tests of the PWLS-OS building blocks and loop against the FP64 oracle
(oracle/pwls_ref.py)."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _problem(seed=3, n=(24, 20, 12), nuv=(48, 16), nviews=24):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False) + 0.1
    g = O.G3(n, (1.0, 1.0, 1.0), (0, 0, 0), nuv, (0.9, 1.1), (0, 0), 500.0, 0, 850.0, betas)
    xt = np.zeros(n)
    xt[5:18, 4:15, 2:10] = 0.2
    xt[8:12, 7:10, 4:7] = 0.35
    p = O.cbct("proj,tf", g, xt, prec=64)
    counts = rng.poisson(2e4 * np.exp(-p)).astype(np.float64)
    y = -np.log(np.maximum(counts, 10) / 2e4)
    w = counts
    x0 = np.clip(xt + rng.normal(0, 0.02, n), 0, None)
    return g, y, w, x0


def _geometry(g):
    from helpers import geometry_from_g3
    return geometry_from_g3(g)


def test_huber_kernels_match_oracle():
    from oracle import pwls_ref as R
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, y, w, x0 = _problem()
    P = Projector(_geometry(g))
    rng = np.random.default_rng(9)
    x = (rng.uniform(0, 0.01, g.img_shape())).astype(np.float32)
    dx = P.upload_volume(x)
    nx, ny, nz = g.img_shape()
    delta = 0.003
    for mode in (1, 2):
        out = DeviceArray(P.n_img, "f32")
        call("ctsim_huber", dx.vp, nx, ny, nz, delta, mode, out.vp, None, None)
        want = R.huber(x.astype(np.float64), mode, delta)
        assert rel_l2(P.download_volume(out), want) < 1e-6, mode
    cost = DeviceArray(1, "f64")
    call("ctsim_huber", dx.vp, nx, ny, nz, delta, 0, None, cost.vp, None)
    assert cost.to_numpy()[0] == pytest.approx(R.huber(x.astype(np.float64), 0, delta), rel=1e-6)


def test_residual_kernel_and_fp64_cost():
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    rng = np.random.default_rng(4)
    n = 300001
    p, yy, ww = (rng.uniform(0, 2, n).astype(np.float32) for _ in range(3))
    dp, dy, dw = (DeviceArray.from_numpy(a, "f32") for a in (p, yy, ww))
    r = DeviceArray(n, "f32")
    phi = DeviceArray(1, "f64").fill(0.0)
    call("ctsim_pwls_residual", dp.vp, 0, dy.vp, 0, dw.vp, 0, r.vp, 0, n, 1.0, phi.vp, None)
    call("ctsim_pwls_residual", dp.vp, 0, dy.vp, 0, dw.vp, 0, r.vp, 0, n, 3.0, phi.vp, None)  # accumulates
    d = p.astype(np.float64) - yy
    np.testing.assert_allclose(r.to_numpy(), (ww * (p - yy)).astype(np.float32), rtol=1e-6)
    assert phi.to_numpy()[0] == pytest.approx(4 * np.sum(ww * d * d) / 2, rel=1e-9)


def test_pwls_plain_matches_fp64_oracle():
    from oracle import pwls_ref as R
    from ctsim_amd.pwls import PwlsOS
    g, y, w, x0 = _problem()
    beta, delta, M, its = 50.0, 1e-3, 4, 3
    want, phis, rmsds = R.nesterov_sqs(y, w, g, beta, delta, its, M, x0)
    sol = PwlsOS(_geometry(g), y, w, M, beta, delta, mixed=False)
    x, ph, rm = sol.run(x0.astype(np.float32), its)
    print("plain vs fp64 oracle:", rel_l2(x, want), ph, phis)
    assert rel_l2(x, want) < 1e-4
    np.testing.assert_allclose(ph, phis, rtol=1e-4)
    np.testing.assert_allclose(rm, rmsds, rtol=1e-3)


def test_pwls_mixed_graph_equals_stream_and_is_close_to_plain():
    from ctsim_amd.pwls import PwlsOS
    g, y, w, x0 = _problem(seed=5)
    beta, delta, M, its = 50.0, 1e-3, 4, 3
    a = PwlsOS(_geometry(g), y, w, M, beta, delta, mixed=True, graph=True)
    xa, pa, _ = a.run(x0.astype(np.float32), its)
    b = PwlsOS(_geometry(g), y, w, M, beta, delta, mixed=True, graph=False)
    xb, pb, _ = b.run(x0.astype(np.float32), its)
    # the graph replays the same kernels with the same momentum scalars
    assert np.array_equal(xa, xb)
    np.testing.assert_array_equal(pa, pb)
    c = PwlsOS(_geometry(g), y, w, M, beta, delta, mixed=False)
    xc, pc, _ = c.run(x0.astype(np.float32), its)
    # the stored data are y - A x0 and the projector reads z - x0: the FP16 rounding is
    # taken on the small difference, so the iterate stays close to the plain FP32 one
    d = PwlsOS(_geometry(g), y, w, M, beta, delta, mixed=True, graph=True, offset=False)
    xd, _, _ = d.run(x0.astype(np.float32), its)
    # the offset taken at the projector input (the update kernel writes z - x0) instead of
    # at the residual (A x0 stored and subtracted): the same data, other FP32 roundings
    e = PwlsOS(_geometry(g), y, w, M, beta, delta, mixed=True, graph=True, offset_mode="input")
    assert e.zc is not None and a.zc is None and a.ax0 is not None
    xe, _, _ = e.run(x0.astype(np.float32), its)
    print("mixed vs plain:", rel_l2(xa, xc), "without the offset:", rel_l2(xd, xc),
          "offset at the projector input:", rel_l2(xe, xc), "residual vs input offset:", rel_l2(xa, xe))
    assert rel_l2(xa, xc) < 5e-4
    assert rel_l2(xe, xc) < 5e-4
    assert rel_l2(xa, xe) < 5e-4
    assert rel_l2(xd, xc) < 2e-3
    np.testing.assert_allclose(pa, pc, rtol=1e-3)


def test_pwls_fp64_matches_fp64_oracle():
    """the all-double form (FP64 plan, FP64 elementwise kernels) is the accuracy reference
    of the mixed form, so it must reproduce the FP64 oracle loop to rounding"""
    from oracle import pwls_ref as R
    from ctsim_amd.pwls import PwlsOS
    g, y, w, x0 = _problem(seed=6)
    beta, delta, M, its = 50.0, 1e-3, 4, 3
    want, phis, rmsds = R.nesterov_sqs(y, w, g, beta, delta, its, M, x0)
    sol = PwlsOS(_geometry(g), y, w, M, beta, delta, precision="f64", graph=False)
    x, ph, rm = sol.run(x0, its)
    print("fp64 vs fp64 oracle:", rel_l2(x, want), ph, phis)
    assert rel_l2(x, want) < 1e-9
    np.testing.assert_allclose(ph, phis, rtol=1e-9)
    np.testing.assert_allclose(rm, rmsds, rtol=1e-7)


@pytest.mark.parametrize("pfun", ["quad", "hyper", "aniso", "isohuber", "tv", "isotv"])
def test_penalty_family_kernels_match_oracle(pfun):
    from oracle import pwls_ref as R
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    from ctsim_amd.pwls import penalty_code
    g, y, w, x0 = _problem()
    P = Projector(_geometry(g))
    rng = np.random.default_rng(10)
    x = rng.uniform(0, 0.01, g.img_shape()).astype(np.float32)
    dx = P.upload_volume(x)
    nx, ny, nz = g.img_shape()
    delta = 0.003
    for mode in (1, 2):
        out = DeviceArray(P.n_img, "f32")
        call("ctsim_penalty", dx.vp, nx, ny, nz, penalty_code(pfun), delta, mode, out.vp, None, None)
        want = R.penalty(pfun, x.astype(np.float64), mode, delta)
        assert rel_l2(P.download_volume(out), want) < 1e-6, (pfun, mode)
    cost = DeviceArray(1, "f64")
    call("ctsim_penalty", dx.vp, nx, ny, nz, penalty_code(pfun), delta, 0, None, cost.vp, None)
    assert cost.to_numpy()[0] == pytest.approx(R.penalty(pfun, x.astype(np.float64), 0, delta), rel=1e-6)


def test_pwls_nesterov_hyperbola_matches_fp64_oracle():
    from oracle import pwls_ref as R
    from ctsim_amd.pwls import PwlsOS
    g, y, w, x0 = _problem(seed=8)
    beta, delta, M, its = 50.0, 1e-2, 4, 3
    want, phis, _ = R.nesterov_sqs(y, w, g, beta, delta, its, M, x0, pfun="hyper")
    sol = PwlsOS(_geometry(g), y, w, M, beta, delta, mixed=False, pfun="hyper")
    x, ph, _ = sol.run(x0.astype(np.float32), its)
    assert rel_l2(x, want) < 1e-4
    np.testing.assert_allclose(ph, phis, rtol=1e-4)


def test_lalm_os_matches_fp64_oracle():
    from oracle import pwls_ref as R
    from ctsim_amd.pwls import PwlsLALM
    g, y, w, x0 = _problem(seed=9)
    beta, delta, M, its = 50.0, 1e-3, 4, 2
    want, phis, rmsds = R.lalm_os(y, w, g, beta, delta, its, M, x0, rho=0.1)
    sol = PwlsLALM(_geometry(g), y, w, M, beta, delta, rho=0.1)
    x, ph, rm = sol.run(x0.astype(np.float32), its)
    print("lalm vs fp64 oracle:", rel_l2(x, want), ph, phis)
    assert rel_l2(x, want) < 1e-4
    np.testing.assert_allclose(ph, phis, rtol=1e-4)
    np.testing.assert_allclose(rm, rmsds, rtol=1e-3)


@pytest.mark.parametrize("pfun", ["huber", "quad", "hyper", "aniso", "isohuber", "tv", "isotv"])
@pytest.mark.parametrize("shape", [(9, 7, 5), (9, 7, 1)])
def test_penalty_family_fp32_fp64_small_volumes(pfun, shape):
    """every penalty and mode on a 3D and a 2D (nz = 1) volume, FP32 and FP64 kernels"""
    from oracle import pwls_ref as R
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.pwls import penalty_code
    rng = np.random.default_rng(11)
    x = rng.uniform(0.001, 0.011, shape)
    nx, ny, nz = shape
    zxy = np.ascontiguousarray(np.transpose(x, (1, 0, 2)))  # device order: z fastest, then x, then y
    delta = 0.003
    modes = (1, 2, 3) if pfun == "huber" else (1, 2)
    for dt, fn, tol in (("f32", "ctsim_penalty", 2e-6), ("f64", "ctsim_penalty_f64", 1e-12)):
        dx = DeviceArray.from_numpy(zxy.ravel(), dt)
        for mode in modes:
            out = DeviceArray(x.size, dt)
            call(fn, dx.vp, nx, ny, nz, penalty_code(pfun), delta, mode, out.vp, None, None)
            got = np.transpose(out.to_numpy().reshape(ny, nx, nz), (1, 0, 2))
            want = R.penalty(pfun, x, mode, delta)
            assert rel_l2(got, want) < tol, (pfun, shape, dt, mode, rel_l2(got, want))
        cost = DeviceArray(1, "f64")
        call(fn, dx.vp, nx, ny, nz, penalty_code(pfun), delta, 0, None, cost.vp, None)
        assert cost.to_numpy()[0] == pytest.approx(R.penalty(pfun, x, 0, delta), rel=max(tol, 1e-9))


def test_nusqs_seed_and_dynamic_range_match_oracle():
    from oracle import pwls_ref as R
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    rng = np.random.default_rng(12)
    shape = (23, 17, 6)
    nx, ny, nz = shape
    x = rng.uniform(0, 0.05, shape)
    x[4:9, 3:7, :] += 0.2
    to_dev = lambda a: DeviceArray.from_numpy(np.ascontiguousarray(np.transpose(a, (1, 0, 2))).ravel(), "f32")  # noqa: E731
    from_dev = lambda d: np.transpose(d.to_numpy().reshape(ny, nx, nz), (1, 0, 2))  # noqa: E731
    dx = to_dev(x)
    v = DeviceArray(x.size, "f32")
    call("ctsim_nusqs_seed", dx.vp, v.vp, nx, ny, nz, 1e-3, None)
    want = np.maximum(2 * R.sobel_edges(x) + np.abs(x), 1e-3)
    assert rel_l2(from_dev(v), want) < 1e-6
    xr = np.clip(x + rng.normal(0, 0.01, shape), 0, None)
    dxr = to_dev(xr)
    d = DeviceArray(x.size, "f32")
    call("ctsim_absdiff_floor", dx.vp, dxr.vp, d.vp, x.size, 1e-3, None)
    # the same FP32 operands and operations: bit-exact
    x32, xr32 = x.astype(np.float32), xr.astype(np.float32)
    dref = np.maximum(np.abs(x32 - xr32), np.float32(1e-3))
    np.testing.assert_array_equal(from_dev(d), dref)
    z = DeviceArray(x.size, "f32")
    for _ in range(2):  # the statistics block is reused: a second call must give the same map
        call("ctsim_dynamic_range", d.vp, x.size, 40.0, 0.05, z.vp, None)
        np.testing.assert_allclose(from_dev(z), R.dynamic_range(dref.astype(np.float64), 40, 0.05), rtol=2e-4,
                                   atol=1e-6)
    q = DeviceArray(x.size, "f32")
    call("ctsim_div", dx.vp, d.vp, q.vp, x.size, None)
    np.testing.assert_array_equal(from_dev(q), x32 / dref)


def test_nusqs_matches_fp64_oracle():
    """four iterations cover every branch of the NU-SQS schedule (seed curvature, xref, the
    dynamic-range update of v and dLtilde, and the return to the seed dL with H(v_new))"""
    from oracle import pwls_ref as R
    from ctsim_amd.pwls import PwlsNUSQS
    g, y, w, x0 = _problem(seed=13)
    beta, delta, M, its = 50.0, 1e-3, 4, 4
    want, phis, rmsds = R.nusqs(y, w, g, beta, delta, its, M, x0)
    sol = PwlsNUSQS(_geometry(g), y, w, M, beta, delta)
    x, ph, rm = sol.run(x0.astype(np.float32), its)
    print("nusqs vs fp64 oracle:", rel_l2(x, want), ph, phis)
    assert rel_l2(x, want) < 1e-4
    np.testing.assert_allclose(ph, phis, rtol=1e-4)


@pytest.mark.parametrize("nz", [12, 13])
@pytest.mark.parametrize("pfun", ["huber", "hyper", "isohuber"])
def test_nesterov_update_step_both_kernel_forms(nz, pfun):
    """one fused update against an FP64 numpy step: nz = 12 runs the four-voxel form (for the
    13-neighbour penalties), nz = 13 and the iso-Huber penalty the one-voxel form"""
    from oracle import pwls_ref as R
    from ctsim_amd._lib import call
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.pwls import penalty_code
    rng = np.random.default_rng(14)
    nx, ny = 9, 7
    shp = (nx, ny, nz)
    x, z, s0 = (rng.uniform(0.0, 0.03, shp) for _ in range(3))
    g = rng.normal(0, 1e-3, shp)
    pre = rng.uniform(5, 20, shp)
    beta, delta, t, tn = 40.0, 4e-3, 1.7, 2.3
    dev = lambda a: DeviceArray.from_numpy(np.ascontiguousarray(np.transpose(a, (1, 0, 2))).ravel(), "f32")  # noqa: E731
    back = lambda d: np.transpose(d.to_numpy().reshape(ny, nx, nz), (1, 0, 2)).astype(np.float64)  # noqa: E731
    dx, dz, ds, dg, dp = dev(x), dev(z), dev(s0), dev(g), dev(pre)
    xo, zo = DeviceArray(x.size, "f32"), DeviceArray(x.size, "f32")
    zc = DeviceArray(x.size, "f16")
    call("ctsim_pwls_nesterov_update", dx.vp, dz.vp, ds.vp, dg.vp, dp.vp, xo.vp, zo.vp, zc.vp, zc.dtype, None, nx, ny, nz,
         beta, delta, penalty_code(pfun), t, tn, None)
    f = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731  the device's FP32 inputs
    x, z, s0, g, pre = f(x), f(z), f(s0), f(g), f(pre)
    du = (g + beta * R.penalty(pfun, z, 1, delta)) / (pre + beta * R.penalty(pfun, x, 2, delta))
    # z' is formed from the unclamped x' and s' and clamped after (reconPwlsSeNesterovSqs.m:72-84)
    xn = np.maximum(z - du, 0)
    sn = np.maximum(s0 - t * du, 0)
    zn = np.maximum((1 - 1 / tn) * (z - du) + (1 / tn) * (s0 - t * du), 0)
    np.testing.assert_allclose(back(xo), xn, rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(back(ds), sn, rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(back(zo), zn, rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(back(zc), zn, rtol=1e-3, atol=1e-6)


@pytest.mark.parametrize("mixed,graph,with_map", [(True, True, False), (False, False, False), (True, True, True)])
def test_fused_backprojection_update_equals_separate_kernels(mixed, graph, with_map):
    """ctsim_plan_back_nesterov (the backprojection with the update in its epilogue) agrees with
    ctsim_plan_back followed by the update kernel, masked columns included. Not bit for bit: the
    separate update kernel's four-voxel form sums the penalty's neighbour terms in another
    order than the fused epilogue's one-voxel form (a last-ulp difference per step).  In the
    mixed form such a difference can also flip the BF16 rounding of a residual element (a
    2^-9 relative step of that element), so its bound over the twelve subset updates is 5e-6"""
    from helpers import geometry_from_g3
    from ctsim_amd.pwls import PwlsOS
    g, y, w, x0 = _problem(seed=7)
    mp = (np.random.default_rng(8).uniform(size=g.n[:2]) < 0.85) if with_map else None
    beta, delta, M, its = 50.0, 1e-3, 4, 3
    a = PwlsOS(geometry_from_g3(g, mp), y, w, M, beta, delta, mixed=mixed, graph=graph, fused=True)
    b = PwlsOS(geometry_from_g3(g, mp), y, w, M, beta, delta, mixed=mixed, graph=graph, fused=False)
    assert a.fused and not b.fused
    xa, pa, _ = a.run(x0.astype(np.float32), its)
    xb, pb, _ = b.run(x0.astype(np.float32), its)
    assert rel_l2(xa, xb) < (5e-6 if mixed else 1e-6), rel_l2(xa, xb)
    np.testing.assert_allclose(pa, pb, rtol=1e-6)


def test_sqs_matches_fp64_oracle():
    """the plain SQS loop (reconPwlsSeSqs.m:51-76) on the device against its FP64 restatement"""
    from oracle import pwls_ref as R
    from ctsim_amd.pwls import PwlsSQS
    g, y, w, x0 = _problem(seed=15)
    beta, delta, M, its = 50.0, 1e-3, 4, 3
    want, phis, rmsds = R.sqs(y, w, g, beta, delta, its, M, x0)
    sol = PwlsSQS(_geometry(g), y, w, M, beta, delta)
    x, ph, rm = sol.run(x0.astype(np.float32), its)
    print("sqs vs fp64 oracle:", rel_l2(x, want), ph, phis)
    assert rel_l2(x, want) < 1e-4
    np.testing.assert_allclose(ph, phis, rtol=1e-4)
    np.testing.assert_allclose(rm, rmsds, rtol=1e-3)


@pytest.mark.parametrize("loop", ["sqs", "lalm", "nusqs"])
def test_graph_replay_equals_stream_execution(loop):
    """each iteration captured once as a hipGraph and replayed: the same kernels with the same
    arguments, so every replay gives the bits of the launch-by-launch loop (four NU-SQS
    iterations cover the schedule steps that run between replays)"""
    from ctsim_amd.pwls import PwlsLALM, PwlsNUSQS, PwlsSQS
    cls = {"sqs": PwlsSQS, "lalm": PwlsLALM, "nusqs": PwlsNUSQS}[loop]
    g, y, w, x0 = _problem(seed=16)
    beta, delta, M, its = 50.0, 1e-3, 4, 4 if loop == "nusqs" else 3
    a = cls(_geometry(g), y, w, M, beta, delta, graph=True)
    xa, pa, ra = a.run(x0.astype(np.float32), its)
    b = cls(_geometry(g), y, w, M, beta, delta, graph=False)
    xb, pb, rb = b.run(x0.astype(np.float32), its)
    assert np.array_equal(xa, xb)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ra, rb)


def test_graph_replay_rejects_odd_subset_counts():
    from ctsim_amd.pwls import PwlsSQS
    g, y, w, x0 = _problem(seed=16)
    with pytest.raises(ValueError, match="even number of subsets"):
        PwlsSQS(_geometry(g), y, w, 3, 50.0, 1e-3, graph=True)


def test_fused_request_that_cannot_be_honoured_raises():
    from ctsim_amd.pwls import PwlsOS
    g, y, w, x0 = _problem(seed=16)
    with pytest.raises(ValueError, match="fused=True"):
        PwlsOS(_geometry(g), y, w, 4, 50.0, 1e-3, precision="f32", pfun="hyper", fused=True)
