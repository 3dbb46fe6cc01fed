"""CPU tests that pin the oracle (oracle/ctsim_oracle.cpp).

1. Golden vectors written by the reference's own compiled kernels (tests/golden/): the
   float instantiation must reproduce them bit for bit.
2. Live comparison with oracle/_ref on random geometries when the reference is present.
3. Closed forms from independent vector geometry (line/detector intersections, ray
   directions), adjoint identities and mirror symmetry for the fp64 instantiation.
"""
import math
import os

import numpy as np
import pytest

from conftest import golden_files, load_golden, rel_l2
from helpers import g2_from, g3_from

OPS2 = ("proj,dd", "proj,tf", "back,dd", "back,tf", "back,pd")
OPS3 = ("proj,dd", "proj,tf", "back,dd", "back,tf", "back,pd", "back,hpd")
# The ray-driven projector ("proj,rd") is checked separately: its sample positions come
# from float accumulation of the ray parameter, so its FP64 form samples elsewhere and is
# no accuracy reference; it is pinned to the compiled reference bit for bit.


@pytest.mark.parametrize("path", golden_files("fan_"), ids=os.path.basename)
def test_fbct_oracle_matches_reference_golden(oracle_mod, path):
    rec = load_golden(path)
    g = g2_from(rec)
    for op in OPS2:
        x = rec["img"] if op.startswith("proj") else rec["sino"]
        got = oracle_mod.fbct(op, g, x)
        want = rec[op.replace(",", "_")]
        assert np.array_equal(got, want), (op, np.abs(got - want).max())
    got = oracle_mod.fbct("proj,rd", g, rec["img"])
    assert np.array_equal(got, rec["proj_rd"]), np.abs(got - rec["proj_rd"]).max()


@pytest.mark.parametrize("path", golden_files("cone_"), ids=os.path.basename)
def test_cbct_oracle_matches_reference_golden(oracle_mod, path):
    rec = load_golden(path)
    g = g3_from(rec)
    mp = rec["map"] if bool(rec["has_map"]) else None
    for op in OPS3:
        x = rec["img"] if op.startswith("proj") else rec["sino"]
        got = oracle_mod.cbct(op, g, x, mp)
        want = rec[op.replace(",", "_")]
        assert np.array_equal(got, want), (op, np.abs(got - want).max())
    got = oracle_mod.cbct("proj,rd", g, rec["img"], None)
    assert np.array_equal(got, rec["proj_rd"]), np.abs(got - rec["proj_rd"]).max()


@pytest.mark.skipif(not os.path.isdir("/root/reference/geom"), reason="reference tree not present")
def test_oracle_matches_live_reference_random(oracle_mod):
    oracle_mod.build(ref=True)
    rng = np.random.default_rng(7)
    for trial in range(8):
        nx, ny = (int(v) for v in rng.integers(6, 20, 2))
        nu, nv, nz, nviews = int(rng.integers(12, 40)), int(rng.integers(4, 16)), int(rng.integers(3, 9)), 4
        betas = rng.uniform(0, 2 * np.pi, nviews)
        dettype = 1 if trial % 3 else 0
        du = float(rng.uniform(0.3, 1.2))
        g2 = oracle_mod.G2(nx, ny, 0.9, 0.7, 0.3, -0.2, nu, du, 0.7, 600, 0, 950, betas, dettype)
        img = rng.uniform(size=(nx, ny)).astype(np.float32)
        sino = rng.uniform(size=(nu, nviews)).astype(np.float32)
        for op in OPS2 + ("proj,rd",):
            x = img if op.startswith("proj") else sino
            assert np.array_equal(oracle_mod.fbct(op, g2, x), oracle_mod.ref_fbct(op, g2, x)), op
        g3 = oracle_mod.G3([nx, ny, nz], [0.9, 0.7, 1.1], [0.3, -0.2, 0.4], [nu, nv], [du, 0.8], [0.7, -0.6], 600, 0,
                           950, betas, couchz=rng.uniform(-1, 1, nviews), dettype=dettype)
        vol = rng.uniform(size=(nx, ny, nz)).astype(np.float32)
        s3 = rng.uniform(size=(nv, nu, nviews)).astype(np.float32)
        mp = rng.uniform(size=(nx, ny)) < 0.7
        for op in OPS3:
            x = vol if op.startswith("proj") else s3
            assert np.array_equal(oracle_mod.cbct(op, g3, x, mp), oracle_mod.ref_cbct(op, g3, x, mp)), op
        assert np.array_equal(oracle_mod.cbct("proj,rd", g3, vol), oracle_mod.ref_cbct("proj,rd", g3, vol))


@pytest.mark.parametrize("path", golden_files("fan_") + golden_files("cone_"), ids=os.path.basename)
def test_fp64_oracle_close_to_fp32(oracle_mod, path):
    rec = load_golden(path)
    if "nxy" in rec:
        g, ops, run = g2_from(rec), OPS2, lambda op, x, p: oracle_mod.fbct(op, g, x, prec=p)
    else:
        g = g3_from(rec)
        mp = rec["map"] if bool(rec["has_map"]) else None
        ops, run = OPS3, lambda op, x, p: oracle_mod.cbct(op, g, x, mp, prec=p)
    for op in ops:
        x = rec["img"] if op.startswith("proj") else rec["sino"]
        assert rel_l2(run(op, x, 32), run(op, x, 64)) < 2e-6, op


# ------------------------------------------------------------------------------------
# independent geometry: source, detector line and projections by vector algebra
# ------------------------------------------------------------------------------------
def det_u(beta, px, py, sad, sdd, du, wu, flat=True):
    """Detector cell coordinate (unshifted) of the ray from the source through (px, py)."""
    s, c = math.sin(beta), math.cos(beta)
    S = np.array([-sad * s, -sad * c])
    n = np.array([s, c])  # central ray direction
    eu = np.array([c, -s])  # detector axis
    q = np.array([px, py]) - S
    if flat:
        lam = sdd / q.dot(n)
        X = S + lam * q
        C = S + sdd * n
        return (X - C).dot(eu) / du - wu
    ang = math.atan2(q.dot(eu), q.dot(n))
    return ang * sdd / du - wu


def overlap(a, b, lo, hi):
    return max(0.0, min(b, hi) - max(a, lo))


def f32(*v):
    """the MEX passes geometry as single: the closed forms use the same rounded values"""
    r = tuple(float(np.float32(x)) for x in v)
    return r if len(r) > 1 else r[0]


@pytest.mark.parametrize("beta", [0.0, 0.3, 1.2, 2.5, 4.0])
def test_dd_single_pixel_closed_form(oracle_mod, beta):
    nx = ny = 9
    dx = f32(0.8)
    nu, (du, sad, sdd) = 64, f32(0.5, 500.0, 800.0)
    g = oracle_mod.G2(nx, ny, dx, dx, 0, 0, nu, du, 0, sad, 0, sdd, [beta], 1)
    ix, iy = 6, 2
    img = np.zeros((nx, ny))
    img[ix, iy] = 1.0
    got = oracle_mod.fbct("proj,dd", g, img, prec=64)[:, 0]
    wu = -(nu - 1) / 2
    xc, yc = (ix - (nx - 1) / 2) * dx, (iy - (ny - 1) / 2) * dx
    s, c = math.sin(beta), math.cos(beta)
    S = np.array([-sad * s, -sad * c])
    d = np.array([xc, yc]) - S
    d /= np.linalg.norm(d)
    xdom = abs(d[1]) > abs(d[0])  # ray closer to the y axis: footprint of the x-edges
    amp = dx / (abs(d[1]) if xdom else abs(d[0])) / 10
    if xdom:
        e = [det_u(beta, xc + dx / 2, yc, sad, sdd, du, wu), det_u(beta, xc - dx / 2, yc, sad, sdd, du, wu)]
    else:
        e = [det_u(beta, xc, yc + dx / 2, sad, sdd, du, wu), det_u(beta, xc, yc - dx / 2, sad, sdd, du, wu)]
    lo, hi = min(e) + 0.5, max(e) + 0.5
    want = np.array([amp * overlap(lo, hi, k, k + 1) for k in range(nu)])
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


def trapezoid_cdf(x, t):
    """integral from -inf to x of the unit-height trapezoid with sorted corners t"""
    t0, t1, t2, t3 = t
    if x <= t0:
        return 0.0
    if x <= t1:
        return (x - t0) ** 2 / (2 * (t1 - t0))
    a = (t1 - t0) / 2
    if x <= t2:
        return a + (x - t1)
    b = a + (t2 - t1)
    if x <= t3:
        return b + (t3 - t2) / 2 - (t3 - x) ** 2 / (2 * (t3 - t2))
    return b + (t3 - t2) / 2


SF_CASES = [(b, True) for b in np.linspace(0.05, 2 * np.pi, 13)] + [(0.4, False), (3.9, False), (5.1, False)]


@pytest.mark.parametrize("beta,flat", SF_CASES)
def test_sf_single_pixel_trapezoid(oracle_mod, beta, flat):
    nx, ny = 11, 7
    dx, dy = f32(0.9, 0.6)
    nu, (sad, sdd) = 80, f32(450.0, 750.0)
    du = f32(0.3)
    g = oracle_mod.G2(nx, ny, dx, dy, 0, 0, nu, du, 0, sad, 0, sdd, [beta], 1 if flat else 0)
    ix, iy = 3, 5
    img = np.zeros((nx, ny))
    img[ix, iy] = 2.0
    got = oracle_mod.fbct("proj,tf", g, img, prec=64)[:, 0]
    wu = -(nu - 1) / 2
    xc, yc = (ix - (nx - 1) / 2) * dx, (iy - (ny - 1) / 2) * dy
    corners = [det_u(beta, xc + a * dx / 2, yc + b * dy / 2, sad, sdd, du, wu, flat) for a in (-1, 1) for b in (-1, 1)]
    t = sorted(corners)
    s, c = math.sin(beta), math.cos(beta)
    d = np.array([xc + sad * s, yc + sad * c])
    d /= np.linalg.norm(d)
    amp = dx / max(abs(d[0]), abs(d[1])) / 10
    # cell iu covers [iu - 1/2, iu + 1/2] in unshifted detector coordinates
    want = np.array([2.0 * amp * (trapezoid_cdf(k + 0.5, t) - trapezoid_cdf(k - 0.5, t)) for k in range(nu)])
    np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-12)
    assert got.sum() == pytest.approx(2.0 * amp * (t[3] + t[2] - t[1] - t[0]) / 2, rel=1e-9)


def test_cbct_single_voxel_closed_form(oracle_mod):
    n = (9, 9, 7)
    dx, dz = f32(0.5, 0.75)
    nu, nv, (du, dv, sad, sdd) = 40, 30, f32(0.4, 0.35, 1000.0, 1500.0)
    beta = 0.9
    g = oracle_mod.G3(n, (dx, dx, dz), (0, 0, 0), (nu, nv), (du, dv), (0, 0), sad, 0, sdd, [beta])
    ix, iy, iz = 4, 4, 5
    vol = np.zeros(n)
    vol[ix, iy, iz] = 1.0
    got = oracle_mod.cbct("proj,dd", g, vol, prec=64)[:, :, 0]
    # centre column: t = 0, D = sad, ray along the central ray: cos phi = max(|c|,|s|) of beta
    s, c = math.sin(beta), math.cos(beta)
    amp_u = dx / max(abs(s), abs(c)) / 10
    zc = (iz - (n[2] - 1) / 2) * dz
    amp_v = math.sqrt(sad ** 2 + zc ** 2) / sad
    mag = sdd / sad
    v_lo = (zc - dz / 2) * mag / dv + (nv - 1) / 2 + 0.5
    v_hi = (zc + dz / 2) * mag / dv + (nv - 1) / 2 + 0.5
    wu = -(nu - 1) / 2
    e = sorted(det_u(beta, a, b, sad, sdd, du, wu) for a, b in
               (((dx / 2, 0), (-dx / 2, 0)) if abs(c) > abs(s) else ((0, dx / 2), (0, -dx / 2))))
    u_lo, u_hi = e[0] + 0.5, e[1] + 0.5
    want = np.zeros((nv, nu))
    for iu in range(nu):
        for iv in range(nv):
            want[iv, iu] = amp_u * amp_v * overlap(u_lo, u_hi, iu, iu + 1) * overlap(v_lo, v_hi, iv, iv + 1)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("op,flat", [("back,hpd", True), ("back,pd", True), ("back,hpd", False)])
def test_pd_backprojection_is_exact_on_linear_data(oracle_mod, op, flat):
    """Bilinear interpolation is exact on data linear in (iu, iv): the FDK backprojection of
    one view equals w2s * (a + b u + c v) at the voxel's projection (vector geometry)."""
    n = (7, 8, 5)
    d = f32(0.7, 0.7, 0.9)
    nu, nv, (sad, sdd) = 50, 40, f32(600.0, 1000.0)
    du = f32(0.45)
    dv = 0.5
    beta = 2.1
    g = oracle_mod.G3(n, d, (0, 0, 0), (nu, nv), (du, dv), (0, 0), sad, 0, sdd, [beta], dettype=1 if flat else 0)
    a, b, cc = 0.3, 0.02, -0.01
    iu, iv = np.meshgrid(np.arange(nu), np.arange(nv))
    sino = (a + b * iu + cc * iv)[:, :, None]
    got = oracle_mod.cbct(op, g, sino, prec=64)
    s, c = math.sin(beta), math.cos(beta)
    wu, wv = -(nu - 1) / 2, -(nv - 1) / 2
    for ix in range(n[0]):
        for iy in range(n[1]):
            xc, yc = (ix - (n[0] - 1) / 2) * d[0], (iy - (n[1] - 1) / 2) * d[1]
            depth = np.array([xc + sad * s, yc + sad * c]).dot([s, c])  # distance along the central ray
            u = det_u(beta, xc, yc, sad, sdd, du, wu, flat)
            if flat:
                w2s = (sdd / depth) ** 2
            else:
                lat = np.array([xc + sad * s, yc + sad * c]).dot([c, -s])
                w2s = sdd ** 2 / (depth ** 2 + lat ** 2)
            for iz in range(n[2]):
                v = (iz - (n[2] - 1) / 2) * d[2] * sdd / depth / dv - wv
                assert 1 <= u <= nu - 2 and 1 <= v <= nv - 2
                assert got[ix, iy, iz] == pytest.approx(w2s * (a + b * u + cc * v), rel=1e-10)


def _adjoint(oracle_mod, fwd, back, x, y):
    ax = fwd(x)
    aty = back(y)
    return abs(np.vdot(ax, y) - np.vdot(x, aty)) / (np.linalg.norm(ax) * np.linalg.norm(y))


@pytest.mark.parametrize("kind", ["dd", "tf"])
def test_adjoint_identity_2d_and_3d(oracle_mod, kind):
    rng = np.random.default_rng(11)
    for path in golden_files("fan_"):
        g = g2_from(load_golden(path))
        x = rng.standard_normal((int(g.nxy[0]), int(g.nxy[1])))
        y = rng.standard_normal((g.nu, g.nviews))
        m = _adjoint(oracle_mod, lambda a: oracle_mod.fbct(f"proj,{kind}", g, a, prec=64),
                     lambda b: oracle_mod.fbct(f"back,{kind}", g, b, prec=64), x, y)
        assert m < 1e-13, (path, m)
    for path in golden_files("cone_"):
        rec = load_golden(path)
        g = g3_from(rec)
        mp = rec["map"] if bool(rec["has_map"]) else None
        x = rng.standard_normal(g.img_shape())
        y = rng.standard_normal(g.sino_shape())
        m = _adjoint(oracle_mod, lambda a: oracle_mod.cbct(f"proj,{kind}", g, a, mp, prec=64),
                     lambda b: oracle_mod.cbct(f"back,{kind}", g, b, mp, prec=64), x, y)
        assert m < 1e-13, (path, m)


def test_mirror_symmetry_2d(oracle_mod):
    """Reflecting the image in x and the angle beta -> -beta reflects the sinogram in u."""
    rng = np.random.default_rng(3)
    betas = rng.uniform(0, 2 * np.pi, 6)
    g = oracle_mod.G2(16, 12, 0.8, 0.8, 0, 0, 40, 0.9, 0, 500, 0, 800, betas, 1)
    gm = oracle_mod.G2(16, 12, 0.8, 0.8, 0, 0, 40, 0.9, 0, 500, 0, 800, -betas, 1)
    img = rng.uniform(size=(16, 12))
    for kind in ("dd", "tf"):
        p = oracle_mod.fbct(f"proj,{kind}", g, img, prec=64)
        pm = oracle_mod.fbct(f"proj,{kind}", gm, img[::-1, :], prec=64)
        np.testing.assert_allclose(pm, p[::-1, :], rtol=1e-9, atol=1e-12)


def test_work_counts_are_data_independent(oracle_mod):
    g = g3_from(load_golden(golden_files("cone_flat_map")[0]))
    U, Cp, K = oracle_mod.cbct_count("tf", g)
    assert U > 0 and Cp >= U and K > 0
    U2, C2, K2 = oracle_mod.cbct_count("tf", g, np.zeros(g.img_shape()[:2], bool))
    assert (U2, C2, K2) == (0.0, 0.0, 0.0)
