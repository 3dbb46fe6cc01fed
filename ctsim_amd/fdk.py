"""FDK reconstruction (fbp/reconFDK.m:1-85) on the GPU.

Host side (as in the reference, computed once per geometry): the ramp filter of
fbp/filter/designFilter2.m (flat panel) or designEquiangularFilter2.m (arc), scaled by
du, and the short-scan Parker weights of fbp/weights/shortScanParkerWeight.m.  Device
side: the cosine weight and the row filtering of reconFDK.m:33-76 (ctsim_fdk_filter:
weighting, zero padding, batched hipFFT, H, inverse FFT, truncation), the filtered views
stored in FP32, FP16 or BF16, and backProjectMex(sino, geom, 1, 0, 'back,hpd') * dbeta * 10
(reconFDK.m:78-84) with the voxel-driven backprojector.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import F32, F64, call, lib
from .device import DeviceArray, dtype_code
from .geometry import Geometry
from .projector import Projector


def nextpow2(n: float) -> int:
    return int(math.ceil(math.log2(abs(n)))) if n else 0


def design_filter2(gamma: float, window: str, length: int, crop: float = 1.0, sdd: float | None = None) -> np.ndarray:
    """Ramp filter in the frequency domain, zero padding included (designFilter2.m:18-78);
    with sdd, the equiangular (arc detector) kernel of designEquiangularFilter2.m."""
    order = max(64, 2 ** nextpow2(2 * length - 1))
    index = np.arange(-order // 2 + 1, order // 2)
    h = np.zeros(index.size + 1)  # h(order) = 0 pads to length order
    h[:index.size][index == 0] = 1.0 / (4.0 * gamma ** 2)
    odd = np.mod(index, 2) == 1  # MATLAB mod: negative odd indices included
    if sdd is None:
        h[:index.size][odd] = -1.0 / (index[odd] * math.pi * gamma) ** 2
    else:
        h[:index.size][odd] = -1.0 / (sdd * math.pi * np.sin(index[odd] * gamma / sdd)) ** 2
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


def sliver_weight(gammas: np.ndarray, beta: float, tau: float) -> np.ndarray:
    """shortScanSliverWeight.m:8-21"""
    out = np.zeros(gammas.size)
    for i, g in enumerate(gammas):
        if 0 <= beta < tau + 2 * g:
            out[i] = math.sin(math.pi / 4 * beta / (tau / 2 + g)) ** 2
        elif tau + 2 * g <= beta < math.pi + 2 * g:
            out[i] = 1.0
        elif math.pi + 2 * g <= beta < math.pi + tau:
            out[i] = math.sin(math.pi / 4 * (math.pi + tau - beta) / (tau / 2 - g)) ** 2
    return out


def fdk_scale(geom: Geometry) -> float:
    """dbeta * 10 of reconFDK.m:78-84 (and reconFBP2d.m:60-66)"""
    dbeta = abs(geom.betas[-1] - geom.betas[0]) / (geom.noViews - 1)
    if not geom.shortScan:
        dbeta = dbeta / 2
    return dbeta * 10.0


def fdk_filter_response(geom: Geometry, window: str = "hamming", crop: float = 1.0) -> np.ndarray:
    """H of reconFDK.m:38-46 / reconFBP2d.m:31-40: du * designFilter2 (flat) or
    du * designEquiangularFilter2 (arc)."""
    nu = _det(geom.detSize)[0]
    du = _det(geom.detSpacing)[0]
    return du * design_filter2(du, window, nu, crop, sdd=None if geom.flatPanel else geom.SDD)


def _det(x):
    """detector size / spacing / offset as a sequence (a 2D geometry may hold scalars)"""
    return tuple(np.atleast_1d(x))


def fdk_view_weights(geom: Geometry) -> np.ndarray | None:
    """short-scan weights [nviews][nu], or None for full scans: the Parker weight of
    reconFDK.m:40-41,45-46,55-59 for a cone beam, the sliver weight of reconFBP2d.m:33-40,47-51
    for a fan beam"""
    if not geom.shortScan:
        return None
    nu = _det(geom.detSize)[0]
    du, ou = _det(geom.detSpacing)[0], _det(geom.detOffset)[0]
    u = (np.arange(nu) - (nu - 1) / 2 + ou) * du
    if geom.flatPanel:
        gamma = np.arctan(-u / geom.SDD)
        tau = abs(geom.betas[-1] - geom.betas[0]) - math.pi
    else:
        gamma = -u / geom.SDD
        tau = abs(geom.betas.max() - geom.betas.min()) - math.pi
    bmin = geom.betas.min()
    if geom.ndim == 2:
        tau = abs(geom.betas[-1] - geom.betas[0]) - math.pi
        return np.stack([sliver_weight(gamma, b - bmin, tau) for b in geom.betas]).astype(np.float32)
    return np.stack([parker_weight(gamma, b - bmin, tau) for b in geom.betas]).astype(np.float32)


class FdkFilter:
    """Device FDK / fan-beam FBP pre-processing for one geometry (ctsim_fdk_create or, for a
    2D geometry, ctsim_fbp2d_create; ctsim_fdk_filter)."""

    def __init__(self, geom: Geometry, window: str = "hamming", crop: float = 1.0, compute: str = "f32"):
        self.geom = geom
        self.compute = dtype_code(compute)
        H = np.ascontiguousarray(fdk_filter_response(geom, window, crop), np.float64)
        vw = fdk_view_weights(geom)
        self._vw = None if vw is None else np.ascontiguousarray(vw)
        g = geom.to_c()
        p = C.c_void_p()
        call("ctsim_fbp2d_create" if geom.ndim == 2 else "ctsim_fdk_create", C.byref(p), C.byref(g), geom.noViews,
             H.ctypes.data_as(C.POINTER(C.c_double)), H.size,
             None if self._vw is None else self._vw.ctypes.data_as(C.POINTER(C.c_float)), self.compute)
        self.handle = p.value

    def __call__(self, d_sino: DeviceArray, d_out: DeviceArray, view0: int = 0, nsub: int | None = None,
                 stream=None) -> DeviceArray:
        nsub = self.geom.noViews if nsub is None else nsub
        det = _det(self.geom.detSize)
        nu, nv = det[0], (det[1] if len(det) > 1 else 1)
        need = nsub * nu * nv
        if d_sino.count < need or d_out.count < need:
            raise ValueError(f"FDK buffers need {need} elements")
        call("ctsim_fdk_filter", C.c_void_p(self.handle), view0, nsub, d_sino.vp, d_sino.dtype, d_out.vp, d_out.dtype,
             C.c_void_p(stream.handle if stream else 0))
        return d_out

    def __del__(self):
        try:
            if self.handle:
                lib().ctsim_fdk_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


def recon_fdk(sino, geom: Geometry, window: str = "hamming", crop: float = 1.0, storage: str = "f32",
              projector: Projector | None = None, compute: str = "f32") -> np.ndarray:
    """reconFDK: log sinogram (MATLAB [nv nu nviews], numpy or a device array in that
    order) -> volume (MATLAB order [nx ny nz], 1/cm).  The filtered views are kept on the
    device in `storage` (f32, f16, bf16; f64 with compute='f64')."""
    f64 = dtype_code(compute) == F64
    P = projector or Projector(geom, compute="f64" if f64 else "f32")
    if isinstance(sino, DeviceArray):
        d_in = sino
    else:
        d_in = P.upload_sino(np.asarray(sino, np.float64 if f64 else np.float32), dtype=F64 if f64 else F32)
    filt = FdkFilter(geom, window, crop, compute="f64" if f64 else "f32")
    d_f = filt(d_in, DeviceArray(d_in.count, storage))
    d_v = DeviceArray(P.n_img, F64 if f64 else F32)
    P.back(d_f, d_v, "back,hpd", scale=fdk_scale(geom))
    return P.download_volume(d_v)


def recon_fbp2d(sino, geom: Geometry, window: str = "hamming", crop: float = 1.0, storage: str = "f32",
                projector: Projector | None = None, compute: str = "f32") -> np.ndarray:
    """reconFBP2d (fbp/reconFBP2d.m:1-70): fan-beam log sinogram [nu nviews] -> image [nx ny]
    (1/cm): cosine (and short-scan sliver) weight and ramp filter on the device, back,pd
    times dbeta * 10, zero outside geom.map."""
    if geom.ndim != 2:
        raise ValueError("reconFBP2d needs a fan-beam geometry")
    f64 = dtype_code(compute) == F64
    P = projector or Projector(geom, compute="f64" if f64 else "f32")
    if isinstance(sino, DeviceArray):
        d_in = sino
    else:
        d_in = P.upload_sino(np.asarray(sino, np.float64 if f64 else np.float32), dtype=F64 if f64 else F32)
    filt = FdkFilter(geom, window, crop, compute="f64" if f64 else "f32")
    d_f = filt(d_in, DeviceArray(d_in.count, storage))
    d_v = DeviceArray(P.n_img, F64 if f64 else F32)
    P.back(d_f, d_v, "back,pd", scale=fdk_scale(geom))
    img = P.download_volume(d_v)
    if geom.map is not None:
        img = np.where(np.asarray(geom.map, bool), img, 0).astype(img.dtype)
    return img
