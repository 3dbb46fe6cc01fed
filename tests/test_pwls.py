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
