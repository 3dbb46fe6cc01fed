"""CPU pins of the projection-matrix (PMat / C-arm) oracle, oracle/pmat_oracle.cpp.

The reference (geom/pcbct) cannot be compiled here: its kernels include ct_def.h, which
needs MATLAB's mex.h (mxArray, mxCalloc).  The restatement is pinned instead by:
  * the matched pair: <A x, y> = <x, A^T y> for proj,dd / back,dd (the cell amplitude is
    applied after the forward sum and before the backward one);
  * a circular orbit written as projection matrices (oracle.circular_pmat) against the
    already-pinned circular oracle (geom/cbct_os, bit-exact to the compiled reference):
    the DD footprints are the same projections of the voxel edges, and the per-cell
    amplitude of the PMat code and the per-voxel one of cbct_os agree to O(rounding) on
    these detectors, so a transposed matrix, a wrong edge or a lost +1/2 shift shows;
  * back,hpd of a sinogram linear in the cell indices: bilinear interpolation is exact
    there, so every voxel projecting inside the detector receives
    w2s (a u + b v + c) with (u, v) from numpy's own matrix product;
  * FP32 against FP64, the map, and tf = dd (cbct_geom_pmat_mex.c:115-118, 245-248).
"""
import numpy as np
import pytest

from conftest import rel_l2


@pytest.fixture(scope="module")
def geo(oracle_mod):
    O = oracle_mod
    n, d, o = (24, 20, 12), (0.8, 0.8, 1.0), (0.5, -0.25, 0.5)
    nuv, duv = (48, 20), (1.0, 1.2)
    betas = np.linspace(0, 2 * np.pi, 8, endpoint=False) + 0.1
    P, cams, sdds, ouv = O.circular_pmat(betas, 1000.0, 1500.0, nuv, duv)
    m = np.ones(n[:2], bool)
    m[:2, :] = False
    g = O.GP(n, d, o, nuv, duv[0], duv[1], P, cams, sdds, ouv, map_=m)
    g3 = O.G3(n, d, o, nuv, duv, (0, 0), 1000.0, 500.0, 1500.0, betas)
    return O, g, g3, m


def test_pmat_pair_is_adjoint(geo):
    O, g, _, _ = geo
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, g.img_shape())
    y = rng.uniform(0, 1, g.sino_shape())
    for prec, tol in ((64, 1e-13), (32, 1e-6)):
        ax, aty = O.pmat("proj,dd", g, x, prec), O.pmat("back,dd", g, y, prec)
        mis = abs(np.vdot(ax, y) - np.vdot(x, aty)) / (np.linalg.norm(ax) * np.linalg.norm(y))
        assert mis < tol, (prec, mis)
        assert np.linalg.norm(ax) > 0


def test_circular_pmat_equals_circular_oracle(geo):
    O, g, g3, m = geo
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, g.img_shape())
    y = rng.uniform(0, 1, g.sino_shape())
    # (the cell amplitude alone moves the result by about 1e-5 here, so 2e-6 also pins it)
    assert rel_l2(O.pmat("proj,dd", g, x, 64), O.cbct("proj,dd", g3, x, map_=m, prec=64)) < 2e-6
    assert rel_l2(O.pmat("back,dd", g, y, 64), O.cbct("back,dd", g3, y, map_=m, prec=64)) < 2e-6
    # the pd weight here is SDD^2 / |camera - voxel|^2 (in-plane) against cbct_os's
    # (SDD / depth)^2: the same to O(fan angle^2) on this 20 mm field
    assert rel_l2(O.pmat("back,pd", g, y, 64), O.cbct("back,pd", g3, y, map_=m, prec=64)) < 2e-4


def test_back_hpd_exact_on_linear_data(oracle_mod):
    O = oracle_mod
    n, d, o = (16, 16, 10), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    nuv = (64, 40)
    P, cams, sdds, ouv = O.circular_pmat([0.3], 800.0, 1200.0, nuv, (1.0, 1.0))
    g = O.GP(n, d, o, nuv, 1.0, 1.0, P, cams, sdds, ouv)
    a, b, c = 0.25, -0.5, 3.0
    iu, iv = np.meshgrid(np.arange(nuv[0]), np.arange(nuv[1]))  # [nv, nu]
    sino = (a * iu + b * iv + c)[:, :, None]
    got = O.pmat("back,hpd", g, sino, 64)
    wx = -(n[0] - 1) / 2.0 + o[0]
    wy = -(n[1] - 1) / 2.0 + o[1]
    wz = -(n[2] - 1) / 2.0 + o[2]
    X, Y, Z = np.meshgrid((np.arange(n[0]) + wx) * d[0], (np.arange(n[1]) + wy) * d[1],
                          (np.arange(n[2]) + wz) * d[2], indexing="ij")
    # the gateway takes single-precision matrices and cameras: the same rounded values here
    P32 = g.pmat.astype(np.float64)
    c32 = g.cameras.astype(np.float64)
    H = np.stack([X, Y, Z, np.ones_like(X)], -1) @ P32[:, :, 0].T
    u, v = H[..., 0] / H[..., 2], H[..., 1] / H[..., 2]
    w2s = 1200.0 ** 2 / ((c32[0, 0] - X) ** 2 + (c32[1, 0] - Y) ** 2)
    inside = (u > 1e-6) & (u < nuv[0] - 1 - 1e-6) & (v >= 0.5) & (v <= nuv[1] - 1.5)
    assert inside.sum() > 0.9 * inside.size
    want = w2s * (a * u + b * v + c)
    assert np.max(np.abs(got[inside] - want[inside]) / np.abs(want[inside])) < 1e-10


def test_pmat_fp32_tf_and_map(geo):
    O, g, _, m = geo
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, g.img_shape()).astype(np.float32)
    y = rng.uniform(0, 1, g.sino_shape()).astype(np.float32)
    for op in ("proj,dd", "back,dd", "back,pd", "back,hpd"):
        data = x if op.startswith("proj") else y
        assert rel_l2(O.pmat(op, g, data, 32), O.pmat(op, g, data, 64)) < 5e-6, op
    assert np.array_equal(O.pmat("proj,tf", g, x), O.pmat("proj,dd", g, x))
    assert np.array_equal(O.pmat("back,tf", g, y), O.pmat("back,dd", g, y))
    back = O.pmat("back,dd", g, y)
    assert not back[~m].any() and back[m].any()
    x2 = x.copy()
    x2[~m] = 7.0  # voxels outside the map do not project
    assert np.array_equal(O.pmat("proj,dd", g, x2), O.pmat("proj,dd", g, x))
