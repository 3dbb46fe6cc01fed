"""CPU tests pinning the PWLS oracle (oracle/pwls_ref.py): the Huber gradient is the
derivative of the Huber cost (central finite differences), the curvature counts the
neighbour pairs inside the quadratic zone, and the loop decreases the cost."""
import numpy as np
import pytest

from oracle import pwls_ref as R


def test_huber_gradient_is_derivative_of_cost():
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 0.02, (6, 5, 4))
    delta = 0.004
    S = R.huber(x, 1, delta)
    h = 1e-7
    for idx in [(0, 0, 0), (2, 3, 1), (5, 4, 3), (3, 0, 2), (1, 4, 0)]:
        e = np.zeros_like(x)
        e[idx] = h
        fd = (R.huber(x + e, 0, delta) - R.huber(x - e, 0, delta)) / (2 * h)
        assert S[idx] == pytest.approx(fd, rel=1e-6, abs=1e-12)


def test_huber_curvature_counts_quadratic_neighbours():
    x = np.zeros((4, 4, 3))
    T = R.huber(x, 2, 1e-3)
    # interior voxel: 4 axial neighbours (1), 4 diagonals (1/sqrt2), 2 z neighbours (1/2)
    assert T[1, 1, 1] == pytest.approx(4 + 4 / np.sqrt(2) + 1)
    # corner voxel (0,0,0): x+, y+, one diagonal, z+
    assert T[0, 0, 0] == pytest.approx(2 + 1 / np.sqrt(2) + 0.5)
    x[1, 1, 1] = 1.0  # far outside delta: its 10 pairs leave the quadratic zone
    T2 = R.huber(x, 2, 1e-3)
    assert T2[1, 1, 1] == 0.0


def test_huber_cost_closed_form_two_voxels():
    x = np.zeros((2, 1, 1))
    x[0, 0, 0] = 0.01
    d = 0.002
    assert R.huber(x, 0, d) == pytest.approx(d * 0.01 - d * d / 2)
    x[0, 0, 0] = 0.001
    assert R.huber(x, 0, d) == pytest.approx(0.001 ** 2 / 2)


def test_nesterov_sqs_decreases_cost_on_noiseless_data():
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    n, nuv, nviews = (10, 10, 4), (24, 8), 12
    betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    g = O.G3(n, (1.0, 1.0, 1.0), (0, 0, 0), nuv, (1.0, 1.0), (0, 0), 500.0, 0, 800.0, betas)
    xt = np.zeros(n)
    xt[3:7, 3:7, 1:3] = 0.2
    y = O.cbct("proj,tf", g, xt, prec=64)
    w = np.full(y.shape, 100.0)
    x, phis, rmsds = R.nesterov_sqs(y, w, g, 1e-2, 1e-3, 6, 3, np.zeros(n))
    assert phis[-1] < phis[0]
    assert np.linalg.norm(x - xt) < np.linalg.norm(xt)


@pytest.mark.parametrize("pfun", ["quad", "hyper"])
def test_penalty_gradient_and_curvature_by_finite_differences(pfun):
    """S = dR/dx and, for these smooth potentials, T = the diagonal of the Hessian of R
    (quadPenalty.m's constant T equals the interior voxels' diagonal)"""
    from oracle import pwls_ref as R
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 0.02, (6, 5, 4))
    delta = 0.01
    S = R.penalty(pfun, x, 1, delta)
    T = R.penalty(pfun, x, 2, delta)
    h = 1e-6
    for idx in [(0, 0, 0), (2, 3, 1), (5, 4, 3), (3, 2, 2)]:
        e = np.zeros_like(x)
        e[idx] = h
        num = (R.penalty(pfun, x + e, 0, delta) - R.penalty(pfun, x - e, 0, delta)) / (2 * h)
        assert S[idx] == pytest.approx(num, rel=1e-6, abs=1e-10)
        curv = (R.penalty(pfun, x + e, 1, delta)[idx] - R.penalty(pfun, x - e, 1, delta)[idx]) / (2 * h)
        if pfun == "hyper" or idx == (3, 2, 2):  # interior voxel for the quadratic constant
            assert T[idx] == pytest.approx(curv, rel=1e-5)


# ---- the rest of loadPenaltyOperator's family, Huber mode 3, and NU-SQS -------------------

def _offsets(x, axis_only):
    """(di, dj, dk, w) neighbour offsets of a voxel, written out independently of the oracle's
    slice pairs: the 13-neighbourhood of huberPenalty.m, or the axis pairs of huberIsoPenalty.m"""
    a = 1 / np.sqrt(2)
    three = x.shape[2] > 1
    if axis_only:
        o = [(1, 0, 0, 1.0), (0, 1, 0, 1.0)] + ([(0, 0, 1, 1.0)] if three else [])
    else:
        o = [(1, 0, 0, 1.0), (0, 1, 0, 1.0), (1, 1, 0, a), (-1, 1, 0, a)] + ([(0, 0, 1, 0.5)] if three else [])
    return o + [(-i, -j, -k, w) for i, j, k, w in o]


def _brute(x, f, axis_only=False, divide=False):
    """gather form: y(v) = sum over neighbours n of v of w f(x_v, x_n), by loops"""
    y = np.zeros_like(x)
    nx, ny, nz = x.shape
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for di, dj, dk, w in _offsets(x, axis_only):
                    a, b, c = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        y[i, j, k] += w * f(x[i, j, k], x[a, b, c])
                if divide:
                    y[i, j, k] /= x[i, j, k]
    return y


def test_aniso_by_brute_force_loops():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 0.05, (5, 4, 3))
    dl = 0.02
    c = lambda d: np.exp(-d * d / dl ** 2)  # noqa: E731
    np.testing.assert_allclose(R.aniso(x, 1, dl), _brute(x, lambda p, q: (p - q) * c(p - q)), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(R.aniso(x, 2, dl), _brute(x, lambda p, q: c(p - q)), rtol=1e-12)
    # every pair counted once: half the gather sum of the pair cost
    half = 0.5 * _brute(x, lambda p, q: (p - q) ** 2 * c(p - q) / 2).sum()
    assert R.aniso(x, 0, dl) == pytest.approx(half, rel=1e-12)


def test_huber_mode3_brute_force_and_constant_image():
    rng = np.random.default_rng(4)
    x = rng.uniform(0.01, 0.03, (5, 4, 3))
    dl = 0.008
    ref = _brute(x, lambda p, q: float(abs(p - q) <= dl) * (p + q), divide=True)
    np.testing.assert_allclose(R.huber(x, 3, dl), ref, rtol=1e-12)
    H = R.huber(np.full((4, 4, 3), 0.02), 3, dl)
    assert H[1, 1, 1] == pytest.approx(2 * (4 + 4 / np.sqrt(2) + 1))  # all pairs quadratic: 2 sum w


@pytest.mark.parametrize("pfun", ["isohuber", "tv", "isotv"])
@pytest.mark.parametrize("shape", [(6, 5, 4), (6, 5, 1)])
def test_gradient_penalties_by_finite_differences(pfun, shape):
    """S is the derivative of R for the iso-Huber and both TV forms (3D and 2D volumes)"""
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 0.02, shape)
    S = R.penalty(pfun, x, 1, 0.004)
    h = 1e-7
    for idx in [(0, 0, 0), (2, 3, 0), (5, 4, shape[2] - 1), (3, 0, shape[2] // 2), (4, 2, 0)]:
        e = np.zeros_like(x)
        e[idx] = h
        fd = (R.penalty(pfun, x + e, 0, 0.004) - R.penalty(pfun, x - e, 0, 0.004)) / (2 * h)
        assert S[idx] == pytest.approx(fd, rel=2e-5, abs=1e-9)


def test_isohuber_curvature_and_tv_curvature_closed_forms():
    x = np.zeros((4, 4, 3))
    T = R.isohuber(x, 2, 1e-3)
    assert T[1, 1, 1] == 6 and T[0, 0, 0] == 3  # axis neighbours, weight 1
    # TV on a constant volume: c = 1/sqrt(eps); an interior voxel gets c once per direction
    # from its own cell and once from each lower neighbour's cell
    Tt = R.tv(x, 2)
    c = 1 / np.sqrt(R.TV_EPS)
    assert Tt[1, 1, 1] == pytest.approx(4 * c)
    assert Tt[1, 1, 2] == 0.0  # the last z slice is outside the crop and no x/y cell reaches it
    Ti = R.isotv(x, 2)
    assert Ti[1, 1, 1] == pytest.approx(6 * c)
    assert Ti[1, 1, 2] == pytest.approx(c)  # reached by the z cell of (1, 1, 1)


def test_sobel_edges_match_scipy_correlate_nearest():
    from scipy import ndimage
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (7, 6, 3))
    k1 = np.array([[1, 2, 1], [0, 0, 0], [-1, -2, -1]], float)
    k2 = np.array([[1, 0, -1], [2, 0, -2], [1, 0, -1]], float)
    ref = np.stack([np.abs(ndimage.correlate(x[:, :, k], k1, mode="nearest"))
                    + np.abs(ndimage.correlate(x[:, :, k], k2, mode="nearest")) for k in range(3)], axis=2)
    np.testing.assert_allclose(R.sobel_edges(x), ref, rtol=1e-12, atol=1e-12)


def test_dynamic_range_properties():
    rng = np.random.default_rng(7)
    x = rng.gamma(2.0, 1.0, (9, 8, 5))
    z = R.dynamic_range(x, 40, 0.05)
    order = np.argsort(x, axis=None)
    assert np.all(np.diff(z.ravel()[order]) >= 0)  # monotone in x
    assert z.ravel()[order[0]] == 0.05 and z.ravel()[order[-1]] == pytest.approx(1.0)
    z1 = R.dynamic_range(x, 1, 0.0)  # the CDF itself passes through the empirical fractions
    t2 = x.mean()
    near = np.argmin(np.abs(np.where(x <= t2, x, -np.inf) - t2))
    assert z1.ravel()[near] <= (x <= t2).mean() + 1e-12


def test_nusqs_majorizes_without_penalty():
    """one subset, beta = 0: dLtilde = A'WA v / v majorizes A'WA, so every step lowers the
    weighted least-squares cost"""
    from oracle import oracle as O
    n, nuv, nviews = (10, 10, 4), (24, 8), 12
    betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    g = O.G3(n, (1.0, 1.0, 1.0), (0, 0, 0), nuv, (1.0, 1.0), (0, 0), 500.0, 0, 800.0, betas)
    xt = np.zeros(n)
    xt[3:7, 3:7, 1:3] = 0.2
    y = O.cbct("proj,tf", g, xt, prec=64)
    w = np.full(y.shape, 100.0)
    x0 = np.full(n, 0.05)
    x, phis, _ = R.nusqs(y, w, g, 0.0, 1e-3, 9, 1, x0)
    assert np.all(np.diff(phis) <= 1e-9 * phis[0])
    assert np.linalg.norm(x - xt) < np.linalg.norm(x0 - xt)
