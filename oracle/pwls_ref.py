"""FP64 numpy restatement of the PWLS ordered-subset loop and the Huber penalty.

TEST INFRASTRUCTURE ONLY (imported by tests/): the checker of ctsim_amd.pwls.

  huber(x, mode, delta)          penalty/huberPenalty.m:101-161 (3D), helpers :222-238
  quad(x, mode)                  penalty/quadPenalty.m (mode 2 is the constant 4 + 4/sqrt2 + 1)
  hyper(x, mode, delta)          penalty/hyperbolaPenalty.m (sumHyper/dHyper/cHyper :147-157)
  nesterov_sqs(...)              statsrecon/pwls/reconPwlsSeNesterovSqs.m:37-137
  lalm_os(...)                   statsrecon/pwls/reconPwlsLALMOs.m:37-110

The projectors are the oracle's (oracle.cbct, FP64), called per ordered subset with the
subset's views as geom/shell/forwardProjectMex.m:35-39 selects them, and the subset
backprojection scaled by M (geom/shell/backProjectMex.m:191-193).
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as O

AXY = 1.0 / math.sqrt(2.0)
AZ = 0.5


def _pairs(x):
    """(first, second, weight) slices of the neighbour pairs, in huberPenalty.m order"""
    return [
        ((slice(0, -1), slice(None), slice(None)), (slice(1, None), slice(None), slice(None)), 1.0),
        ((slice(None), slice(0, -1), slice(None)), (slice(None), slice(1, None), slice(None)), 1.0),
        ((slice(0, -1), slice(0, -1), slice(None)), (slice(1, None), slice(1, None), slice(None)), AXY),
        ((slice(1, None), slice(0, -1), slice(None)), (slice(0, -1), slice(1, None), slice(None)), AXY),
        ((slice(None), slice(None), slice(0, -1)), (slice(None), slice(None), slice(1, None)), AZ),
    ]


def huber(x, mode, delta):
    x = np.asarray(x, np.float64)
    if mode == 0:
        y = 0.0
        for a, b, wt in _pairs(x):
            d = np.abs(x[a] - x[b])
            h = np.where(d > delta, delta * d - delta ** 2 / 2, d ** 2 / 2)
            y += wt * h.sum()
        return y
    y = np.zeros_like(x)
    for a, b, wt in _pairs(x):
        d = x[a] - x[b]
        if mode == 1:
            v = np.clip(d, -delta, delta)
            y[a] += wt * v
            y[b] -= wt * v
        elif mode == 2:
            v = (np.abs(d) <= delta).astype(np.float64)
            y[a] += wt * v
            y[b] += wt * v
        else:
            raise ValueError("wrong penalty mode")
    return y


def _pair_penalty(x, mode, dpsi, cpsi, psi):
    x = np.asarray(x, np.float64)
    if mode == 0:
        return sum(wt * psi(x[a] - x[b]).sum() for a, b, wt in _pairs(x))
    y = np.zeros_like(x)
    for a, b, wt in _pairs(x):
        d = x[a] - x[b]
        if mode == 1:
            v = dpsi(d)
            y[a] += wt * v
            y[b] -= wt * v
        elif mode == 2:
            v = cpsi(d)
            y[a] += wt * v
            y[b] += wt * v
        else:
            raise ValueError("wrong penalty mode")
    return y


def quad(x, mode):
    if mode == 2:
        x = np.asarray(x)
        return np.full(x.shape, 4 + 4 * AXY + (2 * AZ if x.ndim == 3 and x.shape[2] > 1 else 0.0))
    return _pair_penalty(x, mode, lambda d: d, None, lambda d: d ** 2 / 2)


def hyper(x, mode, delta):
    return _pair_penalty(x, mode, lambda d: d / np.sqrt(1 + (d / delta) ** 2),
                         lambda d: 1 / np.sqrt(1 + (d / delta) ** 2) ** 3,
                         lambda d: delta ** 2 * (np.sqrt(1 + (d / delta) ** 2) - 1))


def penalty(pfun, x, mode, delta):
    """loadPenaltyOperator.m:22-53 for 'huber', 'quad', 'hyper'"""
    if pfun == "huber":
        return huber(x, mode, delta)
    if pfun == "quad":
        return quad(x, mode)
    if pfun == "hyper":
        return hyper(x, mode, delta)
    raise ValueError("unknow penalty function")


def subset_geom(g: O.G3, M: int, isub: int):
    idx = np.arange(g.nviews) if M == 1 else np.arange(isub % M, g.nviews, M)
    return O.G3(g.n, g.d, g.o, g.nuv, g.duv, g.ouv, g.sad, g.add, g.sdd, g.betas[idx], couchz=g.couchz[idx],
                dettype=g.dettype), idx


def nesterov_sqs(y, w, g: O.G3, beta, delta, itnlim, numos, x0, map_=None, kind="tf", pfun="huber"):
    """returns x, phis, rmsds (FP64)"""
    A = lambda v: O.cbct(f"proj,{kind}", g, v, map_, prec=64)  # noqa: E731
    At = lambda s: O.cbct(f"back,{kind}", g, s, map_, prec=64)  # noqa: E731
    subs = [subset_geom(g, numos, isub) for isub in range(1, numos + 1)]
    x = np.asarray(x0, np.float64)
    precom = At(w * A(np.ones_like(x)))
    t = 1.0
    z = x.copy()
    s = x.copy()
    phis, rmsds = [], []
    for itn in range(1, itnlim + 1):
        xold = x.copy()
        phi = 0.0
        for isub in range(1, numos + 1):
            gs, idx = subs[isub - 1]
            d = O.cbct(f"proj,{kind}", gs, z, map_, prec=64) - y[:, :, idx]
            grad = numos * O.cbct(f"back,{kind}", gs, w[:, :, idx] * d, map_, prec=64) + beta * penalty(
                pfun, z, 1, delta)
            du = grad / (precom + beta * penalty(pfun, x, 2, delta))
            x = z - du
            s = s - t * du
            t = (1 + math.sqrt(1 + 4 * t * t)) / 2
            z = (1 - 1 / t) * x + (1 / t) * s
            for v in (x, s, z):
                v[(v < 0) | ~np.isfinite(v)] = 0
            phi += np.sum(w[:, :, idx] * d ** 2) / 2
        phi += beta * penalty(pfun, x, 0, delta)
        phis.append(phi)
        rmsds.append(math.sqrt(np.mean((x - xold) ** 2)))
    return x, np.array(phis), np.array(rmsds)


def lalm_os(y, w, g: O.G3, beta, delta, itnlim, numos, x0, rho=0.1, pfun="huber", map_=None, kind="tf", inner=10):
    """reconPwlsLALMOs.m:37-110 (extendVoi with k = 1 leaves the volume as is); returns
    x, phis, rmsds (FP64).  RMSD compares with the image before the last subset, as the
    reference's xold is updated per subset."""
    A = lambda v: O.cbct(f"proj,{kind}", g, v, map_, prec=64)  # noqa: E731
    At = lambda s: O.cbct(f"back,{kind}", g, s, map_, prec=64)  # noqa: E731
    subs = [subset_geom(g, numos, isub) for isub in range(1, numos + 1)]
    x = np.asarray(x0, np.float64).copy()
    precom = At(w * A(np.ones_like(x)))
    gacc = np.zeros_like(x)
    phis, rmsds = [], []
    for itn in range(1, itnlim + 1):
        phi = 0.0
        for isub in range(1, numos + 1):
            gs, idx = subs[isub - 1]
            xold = x.copy()
            r = O.cbct(f"proj,{kind}", gs, x, map_, prec=64) - y[:, :, idx]
            dL = numos * O.cbct(f"back,{kind}", gs, w[:, :, idx] * r, map_, prec=64)
            gacc = rho / (rho + 1) * dL + 1 / (rho + 1) * gacc
            s = rho * dL + (1 - rho) * gacc
            s[~np.isfinite(s)] = 0
            for _ in range(inner):
                x = x - (beta * penalty(pfun, x, 1, delta) + rho * precom * (x - xold) + s) / (
                    beta * penalty(pfun, x, 2, delta) + rho * precom)
                x[(x < 0) | ~np.isfinite(x)] = 0
            phi += np.sum(w[:, :, idx] * r ** 2) / 2
        phi += beta * penalty(pfun, x, 0, delta)
        phis.append(phi)
        rmsds.append(math.sqrt(np.mean((x - xold) ** 2)))
    return x, np.array(phis), np.array(rmsds)
