"""CPU pins of the FP64 PWLS loops (oracle/pwls_ref.py: sqs, nesterov_sqs, lalm_os) and of
the oracle's work counts (oracle_cbct_count), by mathematics rather than by retyping.

On a tiny cone-beam problem the system matrix A is built densely, column by column, from
the already pinned FP64 projector oracle, and the penalty is evaluated by a brute-force
loop over voxel pairs (x, y, two in-plane diagonals, z with weights 1, 1, 1/sqrt2, 1/sqrt2,
1/2, the neighbourhood of penalty/huberPenalty.m:101-161), independent of the oracle's
slice arithmetic.  Then:

  * with one subset (M = 1) every loop (NU-SQS included) converges to a point that satisfies the KKT
    conditions of  min_{x >= 0}  Phi(x) = 1/2 ||A x - y||^2_W + beta R(x)  and agrees with
    scipy's L-BFGS-B minimiser of the same Phi to 1e-6;
  * with beta = 0 the SQS loop never raises the cost: precom = A'W A 1 majorises A'W A
    for a non-negative A (De Pierro), so each step minimises a majoriser;
  * with M = 2 subsets the SQS loop settles next to the true minimiser, not next to the
    minimiser of the cost whose data term lacks the x M factor of backProjectMex.m:97-99;
  * two iterations of two subsets of each loop equal the reference's update sequence
    (reconPwlsSeNesterovSqs.m:59-87, reconPwlsSeSqs.m:55-66, reconPwlsLALMOs.m:63-86)
    evaluated with the dense subset matrices and the brute-force penalty; the same
    sequence with a plausible mistake planted (T(z) for T(x), clamping before z is
    formed, the momentum taken after its update, no x M, swapped LALM averaging weights)
    lands far outside that tolerance, so each such mistake in the oracle would fail.

The work counts U (voxel-views with a non-empty footprint), C (voxel-cell weight pairs)
and K (column-views) are recounted by a loop over every voxel, view and slice from vector
geometry (source, detector line, the voxel's edges or corners projected onto it), with no
iz window: a window that dropped a voxel would change U and C.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import pwls_ref as R

OFFSETS = (((1, 0, 0), 1.0), ((0, 1, 0), 1.0), ((1, 1, 0), 1 / math.sqrt(2)), ((-1, 1, 0), 1 / math.sqrt(2)),
           ((0, 0, 1), 0.5))


def pairs(shape):
    """every unordered neighbour pair (a, b, weight) of the penalty, as flat MATLAB indices"""
    nx, ny, nz = shape
    out = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for (di, dj, dk), wt in OFFSETS:
                    i2, j2, k2 = i + di, j + dj, k + dk
                    if 0 <= i2 < nx and 0 <= j2 < ny and 0 <= k2 < nz:
                        out.append((i + nx * (j + ny * k), i2 + nx * (j2 + ny * k2), wt))
    a, b, w = (np.array(v) for v in zip(*out))
    return a, b, w.astype(np.float64)


class Huber:
    """R, S = grad R and the curvature count T of the Huber pair penalty, by pair loops"""

    def __init__(self, shape, delta):
        self.a, self.b, self.w = pairs(shape)
        self.delta = delta

    def R(self, x):
        d = np.abs(x[self.a] - x[self.b])
        return float(np.sum(self.w * np.where(d > self.delta, self.delta * d - self.delta ** 2 / 2, d * d / 2)))

    def S(self, x):
        v = self.w * np.clip(x[self.a] - x[self.b], -self.delta, self.delta)
        return np.bincount(self.a, v, x.size) - np.bincount(self.b, v, x.size)

    def T(self, x):
        v = self.w * (np.abs(x[self.a] - x[self.b]) <= self.delta)
        return np.bincount(self.a, v, x.size) + np.bincount(self.b, v, x.size)


class Problem:
    def __init__(self, seed=0, n=(6, 5, 3), nuv=(14, 6), nviews=10, beta_frac=0.2, delta=0.02):
        rng = np.random.default_rng(seed)
        betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False) + 0.3
        self.g = O.G3(n, (1.0, 1.0, 1.0), (0, 0, 0), nuv, (1.0, 1.0), (0, 0), 200.0, 0, 340.0, betas)
        self.n = n
        nvox = int(np.prod(n))
        self.A = np.empty((nuv[0] * nuv[1] * nviews, nvox))
        for j in range(nvox):
            e = np.zeros(nvox)
            e[j] = 1.0
            self.A[:, j] = O.cbct("proj,tf", self.g, e.reshape(n, order="F"), prec=64).ravel(order="F")
        xt = np.zeros(n)
        xt[1:5, 1:4, :] = 0.2
        xt[2, 2, 1] = 0.45
        self.xt = xt
        p = self.A @ xt.ravel(order="F")
        self.yv = p + rng.normal(0, 0.03 * p.max(), p.size)
        self.wv = rng.uniform(50, 150, p.size)
        self.sshape = (nuv[1], nuv[0], nviews)
        self.y = self.yv.reshape(self.sshape, order="F")
        self.w = self.wv.reshape(self.sshape, order="F")
        self.precom = self.A.T @ (self.wv * (self.A @ np.ones(nvox)))
        self.delta = delta
        self.pen = Huber(n, delta)
        tmax = 2 * (4 + 4 / math.sqrt(2) + 1)
        self.beta = beta_frac * float(np.median(self.precom)) / tmax
        self.cell = nuv[0] * nuv[1]
        self.nviews = nviews

    def rows(self, M, isub):
        """sinogram rows of ordered subset isub (1-based) of M: views 1+mod(isub,M):M:end"""
        views = np.arange(isub % M, self.nviews, M) if M > 1 else np.arange(self.nviews)
        return (views[:, None] * self.cell + np.arange(self.cell)[None, :]).ravel()

    def phi(self, x, scale=1.0):
        r = self.A @ x - self.yv
        return scale * 0.5 * float(np.sum(self.wv * r * r)) + self.beta * self.pen.R(x)

    def grad(self, x, scale=1.0):
        return scale * (self.A.T @ (self.wv * (self.A @ x - self.yv))) + self.beta * self.pen.S(x)

    def minimise(self, scale=1.0):
        from scipy.optimize import minimize
        x0 = np.full(self.A.shape[1], 0.1)
        res = minimize(lambda v: (self.phi(v, scale), self.grad(v, scale)), x0, jac=True, method="L-BFGS-B",
                       bounds=[(0, None)] * x0.size,
                       options=dict(maxiter=50000, maxfun=100000, ftol=1e-18, gtol=1e-14, maxcor=50))
        return res.x

    def vol(self, v):
        return v.reshape(self.n, order="F")

    def flat(self, x):
        return np.asarray(x).ravel(order="F")


@pytest.fixture(scope="module")
def prob():
    return Problem()


@pytest.fixture(scope="module")
def xstar(prob):
    return prob.minimise()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def kkt_residual(p, x):
    """max violation of grad_i = 0 (x_i > 0) and grad_i >= 0 (x_i = 0), relative to the
    gradient at the start image"""
    g = p.grad(x)
    scale = np.abs(p.grad(np.full_like(x, 0.1))).max()
    free = x > 1e-12
    return max(np.abs(g[free]).max(initial=0.0), (-g[~free]).max(initial=0.0)) / scale


def test_lbfgsb_reference_point_satisfies_kkt(prob, xstar):
    assert kkt_residual(prob, xstar) < 1e-8
    # the bound is active somewhere and both Huber zones are populated, so the checks below
    # exercise the clamp and both branches of the penalty
    assert (xstar < 1e-12).sum() > 0
    d = np.abs(xstar[prob.pen.a] - xstar[prob.pen.b])
    assert (d > prob.delta).sum() > 0 and (d <= prob.delta).sum() > 0


@pytest.mark.parametrize("loop,its", [("sqs", 4000), ("nesterov_sqs", 1500), ("lalm_os", 800), ("nusqs", 3000)])
def test_single_subset_loops_converge_to_the_constrained_minimiser(prob, xstar, loop, its):
    x0 = prob.vol(np.full(xstar.size, 0.1))
    f = getattr(R, loop)
    x, phis, rmsds = f(prob.y, prob.w, prob.g, prob.beta, prob.delta, its, 1, x0)
    xv = prob.flat(x)
    print(loop, "rel to L-BFGS-B", rel(xv, xstar), "KKT", kkt_residual(prob, xv), "last rmsd", rmsds[-1])
    assert rel(xv, xstar) < 1e-6
    assert kkt_residual(prob, xv) < 1e-6
    # the loop's own cost is the restated Phi (its data term of the last pass is taken at the
    # image before the pass, so compare once the iterate has stopped moving)
    assert phis[-1] == pytest.approx(prob.phi(xv), rel=1e-9)


def test_sqs_without_penalty_is_monotone(prob):
    """beta = 0: precom = A'W A 1 >= A'W A, every step minimises a majoriser of Phi"""
    x = prob.vol(np.full(prob.A.shape[1], 0.1))
    prev = math.inf
    for it in range(60):
        x, phis, _ = R.sqs(prob.y, prob.w, prob.g, 0.0, prob.delta, 1, 1, x)
        cur = 0.5 * float(np.sum(prob.wv * (prob.A @ prob.flat(x) - prob.yv) ** 2))
        assert cur <= prev * (1 + 1e-12), it
        prev = cur


def test_subset_scaling_targets_the_full_cost(prob, xstar):
    """two ordered subsets: the OS-SQS limit lies next to the true minimiser; the minimiser
    of the cost with the data term not scaled by M lies far away"""
    x0 = prob.vol(np.full(xstar.size, 0.1))
    x, _, _ = R.sqs(prob.y, prob.w, prob.g, prob.beta, prob.delta, 1500, 2, x0)
    x_wrong = prob.minimise(scale=0.5)
    near, far = rel(prob.flat(x), xstar), rel(x_wrong, xstar)
    print("OS limit to minimiser", near, "unscaled-data minimiser to minimiser", far)
    assert near < 0.25 * far
    assert rel(prob.flat(x), x_wrong) > 3 * near


# -------------------------------------------------------------------------------------
# explicit update sequences with the dense subset matrices
# -------------------------------------------------------------------------------------
def nesterov_steps(p, M, itnlim, x0, mutation=None):
    """reconPwlsSeNesterovSqs.m:55-87"""
    x = x0.copy()
    t = 1.0
    z = x.copy()
    s = x.copy()
    clamped = 0
    for _ in range(itnlim):
        for isub in range(1, M + 1):
            rows = p.rows(M, isub)
            As = p.A[rows]
            d = As @ z - p.yv[rows]
            k = 1.0 if mutation == "no_xM" else float(M)
            grad = k * (As.T @ (p.wv[rows] * d)) + p.beta * p.pen.S(z)
            du = grad / (p.precom + p.beta * p.pen.T(z if mutation == "T_of_z" else x))
            x = z - du
            t_new = (1 + math.sqrt(1 + 4 * t * t)) / 2
            s = s - (t_new if mutation == "t_after" else t) * du
            t = t_new
            clamped += int((x < 0).sum())
            if mutation == "clamp_first":
                x[x < 0] = 0
                s[s < 0] = 0
            z = (1 - 1 / t) * x + (1 / t) * s
            for v in (x, s, z):
                v[v < 0] = 0
    return x, clamped


def sqs_steps(p, M, itnlim, x0, mutation=None):
    """reconPwlsSeSqs.m:51-66"""
    x = x0.copy()
    for _ in range(itnlim):
        for isub in range(1, M + 1):
            rows = p.rows(M, isub)
            As = p.A[rows]
            d = As @ x - p.yv[rows]
            k = 1.0 if mutation == "no_xM" else float(M)
            num = k * (As.T @ (p.wv[rows] * d)) + p.beta * p.pen.S(x)
            x = x - num / (p.precom + p.beta * p.pen.T(x))
            x[x < 0] = 0
    return x


def lalm_steps(p, M, itnlim, x0, rho=0.1, inner=10, mutation=None):
    """reconPwlsLALMOs.m:60-86"""
    x = x0.copy()
    g = np.zeros_like(x)
    for _ in range(itnlim):
        for isub in range(1, M + 1):
            rows = p.rows(M, isub)
            As = p.A[rows]
            xold = x.copy()
            r = As @ x - p.yv[rows]
            dL = (1.0 if mutation == "no_xM" else float(M)) * (As.T @ (p.wv[rows] * r))
            if mutation == "swap_weights":
                g = 1 / (rho + 1) * dL + rho / (rho + 1) * g
            else:
                g = rho / (rho + 1) * dL + 1 / (rho + 1) * g
            s = rho * dL + (1 - rho) * g
            for _ in range(inner - 1 if mutation == "inner_short" else inner):
                x = x - (p.beta * p.pen.S(x) + rho * p.precom * (x - xold) + s) / (
                    p.beta * p.pen.T(x) + rho * p.precom)
                x[x < 0] = 0
    return x


def start_image(p):
    """a start image with a negative-bound pull: the first steps clamp voxels"""
    rng = np.random.default_rng(1)
    return np.clip(p.flat(p.xt) + rng.normal(0, 0.05, p.A.shape[1]), 0, None)


STEP_TOL = 1e-10


def test_nesterov_sqs_steps_equal_the_reference_sequence(prob):
    x0 = start_image(prob)
    got, _, _ = R.nesterov_sqs(prob.y, prob.w, prob.g, prob.beta, prob.delta, 2, 2, prob.vol(x0))
    want, clamped = nesterov_steps(prob, 2, 2, x0)
    assert clamped > 0
    assert rel(prob.flat(got), want) < STEP_TOL
    for m in ("T_of_z", "clamp_first", "t_after", "no_xM"):
        bad, _ = nesterov_steps(prob, 2, 2, x0, mutation=m)
        assert rel(bad, want) > 1e3 * STEP_TOL, m


def test_sqs_steps_equal_the_reference_sequence(prob):
    x0 = start_image(prob)
    got, _, _ = R.sqs(prob.y, prob.w, prob.g, prob.beta, prob.delta, 2, 2, prob.vol(x0))
    want = sqs_steps(prob, 2, 2, x0)
    assert rel(prob.flat(got), want) < STEP_TOL
    assert rel(sqs_steps(prob, 2, 2, x0, mutation="no_xM"), want) > 1e3 * STEP_TOL


def test_lalm_os_steps_equal_the_reference_sequence(prob):
    x0 = start_image(prob)
    got, _, _ = R.lalm_os(prob.y, prob.w, prob.g, prob.beta, prob.delta, 2, 2, prob.vol(x0))
    want = lalm_steps(prob, 2, 2, x0)
    assert rel(prob.flat(got), want) < STEP_TOL
    for m in ("no_xM", "swap_weights", "inner_short"):
        assert rel(lalm_steps(prob, 2, 2, x0, mutation=m), want) > 1e3 * STEP_TOL, m


def test_brute_force_penalty_agrees_with_oracle_and_finite_differences(prob):
    """the brute-force pair loop is itself pinned: S is the gradient of R (central
    differences), and it equals the oracle's huber modes 0-2"""
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 0.06, prob.A.shape[1])
    h = 1e-7
    for j in (0, 17, 44, 89):
        e = np.zeros_like(x)
        e[j] = h
        fd = (prob.pen.R(x + e) - prob.pen.R(x - e)) / (2 * h)
        assert prob.pen.S(x)[j] == pytest.approx(fd, rel=1e-6, abs=1e-12)
    xv = prob.vol(x)
    assert prob.pen.R(x) == pytest.approx(R.huber(xv, 0, prob.delta), rel=1e-12)
    np.testing.assert_allclose(prob.pen.S(x), prob.flat(R.huber(xv, 1, prob.delta)), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(prob.pen.T(x), prob.flat(R.huber(xv, 2, prob.delta)), rtol=1e-12, atol=1e-15)


# -------------------------------------------------------------------------------------
# work counts U, C, K of the 3D DD / SF operators
# -------------------------------------------------------------------------------------
def brute_counts(kind, g, mp):
    nx, ny, nz = (int(v) for v in g.n)
    dx, dy, dz = (float(v) for v in g.d)
    ox, oy, oz = (float(v) for v in g.o)
    nu, nv = (int(v) for v in g.nuv)
    du, dv = (float(v) for v in g.duv)
    ou, ov = (float(v) for v in g.ouv)
    sad, sdd = float(np.float32(g.sad)), float(np.float32(g.sdd))
    wx, wy = -(nx - 1) / 2 + ox, -(ny - 1) / 2 + oy
    wu, wv = -(nu - 1) / 2 + ou, -(nv - 1) / 2 + ov

    def cells(lo, hi, n):
        clo, chi = max(lo + 0.5, 0.0), min(hi + 0.5, float(n))
        if clo > chi:
            return 0
        ilo = math.floor(clo)
        ihi = min(math.floor(chi), n - 1, ilo + 15)
        return ihi - ilo + 1

    U = C = K = 0
    for ib, beta in enumerate(g.betas):
        s, c = float(np.float32(math.sin(beta))), float(np.float32(math.cos(beta)))
        S = np.array([-sad * s, -sad * c])
        nrm = np.array([s, c])
        eu = np.array([c, -s])

        def u_of(px, py):
            q = np.array([px, py]) - S
            if g.dettype == 1:
                return sdd * q.dot(eu) / q.dot(nrm) / du - wu
            return math.atan(q.dot(eu) / q.dot(nrm)) * sdd / du - wu

        wz = -(nz - 1) / 2 + oz + float(g.couchz[ib]) / dz
        for iy in range(ny):
            for ix in range(nx):
                if mp is not None and not mp[ix, iy]:
                    continue
                xc, yc = (ix + wx) * dx, (iy + wy) * dy
                ray = np.array([xc, yc]) - S
                if kind == "dd":
                    if abs(ray[1]) > abs(ray[0]):
                        us = [u_of(xc + dx / 2, yc), u_of(xc - dx / 2, yc)]
                    else:
                        us = [u_of(xc, yc + dy / 2), u_of(xc, yc - dy / 2)]
                else:
                    us = [u_of(xc + a * dx / 2, yc + b * dy / 2) for a in (-1, 1) for b in (-1, 1)]
                nuc = cells(min(us), max(us), nu)
                if nuc == 0:
                    continue
                K += 1
                mag = sdd / ray.dot(nrm)
                for iz in range(nz):
                    vlo = (iz + wz - 0.5) * dz * mag / dv - wv
                    nvc = cells(vlo, vlo + dz * mag / dv, nv)
                    if nvc:
                        U += 1
                        C += nuc * nvc
    return U, C, K


@pytest.mark.parametrize("seed", range(6))
def test_work_counts_brute_force(seed):
    rng = np.random.default_rng(900 + seed)
    n = (int(rng.integers(5, 9)), int(rng.integers(5, 9)), int(rng.integers(3, 7)))
    nuv = (int(rng.integers(12, 30)), int(rng.integers(8, 14)))
    nviews = 5
    dettype = seed % 2
    g = O.G3(n, tuple(rng.uniform(0.6, 1.3, 3)), tuple(rng.uniform(-0.8, 0.8, 3)), nuv,
             (float(rng.uniform(0.5, 1.2)), float(rng.uniform(1.4, 2.0))), tuple(rng.uniform(-1, 1, 2)), 300.0, 0,
             520.0, rng.uniform(0, 2 * np.pi, nviews), couchz=rng.uniform(-0.5, 0.5, nviews), dettype=dettype)
    mp = (rng.uniform(size=n[:2]) < 0.8) if seed % 3 else None
    for kind in ("dd", "tf"):
        want = brute_counts(kind, g, mp)
        got = O.cbct_count(kind, g, mp, nthreads=1)
        assert tuple(int(v) for v in got) == want, (kind, got, want)
