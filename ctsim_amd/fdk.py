"""FDK reconstruction (fbp/reconFDK.m:1-85) around the GPU voxel-driven backprojector.

The pre-processing restates the reference's MATLAB host code: the flat-panel cosine
weight (reconFDK.m:39), the ramp filter of fbp/filter/designFilter2.m:18-78 scaled by du,
per-row frequency-domain filtering (fbp/filter/filterFreq.m:42-51), Parker weights for
short scans (fbp/weights/shortScanParkerWeight.m), and the final
backProjectMex(sino, geom, 1, 0, 'back,hpd') * dbeta * 10 (reconFDK.m:78-84), which
runs on the GPU with the filtered views stored in FP32, FP16 or BF16.
"""
from __future__ import annotations

import math

import numpy as np

from .device import DeviceArray
from .geometry import Geometry
from .projector import Projector


def nextpow2(n: float) -> int:
    return int(math.ceil(math.log2(abs(n)))) if n else 0


def design_filter2(gamma: float, window: str, length: int, crop: float = 1.0) -> np.ndarray:
    """Ramp filter in the frequency domain, zero padding included (designFilter2.m)."""
    order = max(64, 2 ** nextpow2(2 * length - 1))
    index = np.arange(-order // 2 + 1, order // 2)
    h = np.zeros(index.size + 1)  # h(order) = 0 pads to length order
    h[:index.size][index == 0] = 1.0 / (4.0 * gamma ** 2)
    odd = np.mod(index, 2) == 1  # MATLAB mod: negative odd indices included
    h[:index.size][odd] = -1.0 / (index[odd] * math.pi * gamma) ** 2
    filt = np.abs(np.fft.fft(h))
    filt = filt[: order // 2 + 1]
    w = 2 * math.pi * np.arange(filt.size) / order
    f = filt.copy()
    if window == "ram-lak":
        pass
    elif window == "shepp-logan":
        f[1:] *= np.sin(w[1:] / (2 * crop)) / (w[1:] / (2 * crop))
    elif window == "cosine":
        f[1:] *= np.cos(w[1:] / (2 * crop))
    elif window == "hamming":
        f[1:] *= 0.54 + 0.46 * np.cos(w[1:] / crop)
    elif window == "hann":
        f[1:] *= (1 + np.cos(w[1:] / crop)) / 2
    elif window in ("hann50", "hann75", "hann80"):
        crop = {"hann50": 0.5, "hann75": 0.75, "hann80": 0.8}[window]
        f[1:] *= (1 + np.cos(w[1:] / crop)) / 2
    elif window in ("B50", "B75", "B80"):
        crop = {"B50": 0.5, "B75": 0.75, "B80": 0.8}[window]
    elif window in ("C50", "C75", "C80"):
        crop = {"C50": 0.5, "C75": 0.75, "C80": 0.8}[window]
        f[1:] *= np.cos(w[1:] / (2 * crop))
    else:
        raise ValueError("Invalid window selected.")
    f[w > math.pi * crop] = 0
    return np.concatenate([f, f[-2:0:-1]])


def filter_rows(x: np.ndarray, H: np.ndarray) -> np.ndarray:
    """filterFreq along the last axis: zero-pad to len(H), multiply in frequency, real part, truncate."""
    n = x.shape[-1]
    order = H.size
    X = np.fft.rfft(x, order, axis=-1)
    X *= H[: order // 2 + 1]
    return np.fft.irfft(X, order, axis=-1)[..., :n]


def parker_weight(gammas: np.ndarray, beta: float, tau: float) -> np.ndarray:
    """shortScanParkerWeight.m:1-40"""

    def S(b):
        if b <= -0.5:
            return 0.0
        if abs(b) < 0.5:
            return 0.5 * (1 + math.sin(math.pi * b))
        return 1.0

    out = np.empty(gammas.size)
    for i, gm in enumerate(gammas):
        g = -gm
        out[i] = S(beta / (tau - 2 * g) - 0.5) - S((beta - math.pi + 2 * g) / (tau + 2 * g) - 0.5)
    return out


def fdk_filter(sino: np.ndarray, geom: Geometry, window: str = "hamming", crop: float = 1.0,
               out_dtype=np.float32) -> np.ndarray:
    """Weighted, ramp-filtered views (MATLAB order [nv nu nviews]) as reconFDK.m:19-76 computes them."""
    if not geom.flatPanel:
        raise NotImplementedError("equiangular (arc) FDK filtering is not provided")
    nu, nv = geom.detSize
    du, dv = geom.detSpacing
    ou, ov = geom.detOffset
    SDD, SAD = geom.SDD, geom.SAD
    u = (np.arange(nu) - (nu - 1) / 2 + ou) * du
    v = (np.arange(nv) - (nv - 1) / 2 + ov) * dv
    uu, vv = np.meshgrid(u, v)  # [nv nu]
    weight = SAD * np.sqrt(1.0 + uu ** 2 / SDD ** 2) / np.sqrt(uu ** 2 + vv ** 2 + SDD ** 2)
    H = du * design_filter2(du, window, nu, crop)
    gamma = np.arctan(-u / SDD)
    tau = abs(geom.betas[-1] - geom.betas[0]) - math.pi
    beta_min = geom.betas.min()
    out = np.empty(sino.shape, out_dtype, order="F")
    for view in range(geom.noViews):
        proj = weight * sino[:, :, view]
        if geom.shortScan:
            proj = proj * parker_weight(gamma, geom.betas[view] - beta_min, tau)[None, :]
        out[:, :, view] = filter_rows(proj, H)
    return out


def fdk_scale(geom: Geometry) -> float:
    """dbeta * 10 of reconFDK.m:78-84"""
    dbeta = abs(geom.betas[-1] - geom.betas[0]) / (geom.noViews - 1)
    if not geom.shortScan:
        dbeta = dbeta / 2
    return dbeta * 10.0


def recon_fdk(sino: np.ndarray, geom: Geometry, window: str = "hamming", crop: float = 1.0,
              storage: str = "f32", projector: Projector | None = None) -> np.ndarray:
    """reconFDK: returns the reconstructed volume (MATLAB order [nx ny nz], 1/cm)."""
    filt = fdk_filter(sino, geom, window, crop)
    P = projector or Projector(geom)
    d_s = P.upload_sino(filt, dtype=storage)
    d_v = DeviceArray(P.n_img, "f32")
    P.back(d_s, d_v, "back,hpd", scale=fdk_scale(geom))
    return P.download_volume(d_v)
