"""FP64 numpy restatement of CTSim's FDK pre-processing, the checker of the device FDK
filter (test infrastructure; the product filters on the GPU).

fdk_filter follows fbp/reconFDK.m:19-76 (and fbp2d_filter fbp/reconFBP2d.m:26-55) view by view: the cosine weight (flat :39, arc
:44), the short-scan weight (:55-59), and filterFreq.m:42-51 per detector row
(zero pad to length(H), multiply the FFT by H, real part of the inverse FFT, truncate)."""
from __future__ import annotations

import math

import numpy as np


def filter_rows(x: np.ndarray, H: np.ndarray) -> np.ndarray:
    """filterFreq along the last axis: zero-pad to len(H), multiply in frequency, real part, truncate."""
    n = x.shape[-1]
    order = H.size
    X = np.fft.fft(x, order, axis=-1)
    X = X * H
    return np.real(np.fft.ifft(X, axis=-1))[..., :n]


def fdk_filter(sino: np.ndarray, geom, H: np.ndarray, view_weights=None) -> np.ndarray:
    """sino MATLAB [nv nu nviews] -> weighted, filtered views, same order, float64"""
    nu, nv = geom.detSize
    du, dv = geom.detSpacing
    ou, ov = geom.detOffset
    SDD, SAD = float(geom.SDD), float(geom.SAD)
    u = (np.arange(nu) - (nu - 1) / 2 + ou) * du
    v = (np.arange(nv) - (nv - 1) / 2 + ov) * dv
    uu, vv = np.meshgrid(u, v)  # [nv nu]
    if geom.flatPanel:
        weight = SAD * np.sqrt(1.0 + uu ** 2 / SDD ** 2) / np.sqrt(uu ** 2 + vv ** 2 + SDD ** 2)
    else:
        weight = SAD / SDD * np.cos(uu / np.sqrt(SDD ** 2 + vv ** 2))
    out = np.empty(sino.shape, np.float64, order="F")
    for view in range(sino.shape[2]):
        proj = weight * np.asarray(sino[:, :, view], np.float64)
        if view_weights is not None:
            proj = proj * np.asarray(view_weights[view], np.float64)[None, :]
        out[:, :, view] = filter_rows(proj, H)
    return out


def fbp2d_filter(sino: np.ndarray, geom, H: np.ndarray, view_weights=None) -> np.ndarray:
    """fbp/reconFBP2d.m:26-55: sino MATLAB [nu nviews] -> cosine (SAD / sqrt(u^2 + SDD^2) flat,
    SAD/SDD cos(u/SDD) arc) and short-scan weighted, filterFreq per view, float64"""
    nu = int(np.atleast_1d(geom.detSize)[0])
    du = float(np.atleast_1d(geom.detSpacing)[0])
    ou = float(np.atleast_1d(geom.detOffset)[0])
    SDD, SAD = float(geom.SDD), float(geom.SAD)
    u = (np.arange(nu) - (nu - 1) / 2 + ou) * du
    weight = SAD / np.sqrt(u ** 2 + SDD ** 2) if geom.flatPanel else SAD / SDD * np.cos(u / SDD)
    out = np.empty(sino.shape, np.float64, order="F")
    for view in range(sino.shape[1]):
        proj = weight * np.asarray(sino[:, view], np.float64)
        if view_weights is not None:
            proj = proj * np.asarray(view_weights[view], np.float64)
        out[:, view] = filter_rows(proj, H)
    return out
