"""Batched uneven 2D fan-beam slices in one launch (ctsim_batch2_*), with the launch
schedule picked by the library's autotuner.

Each slice has its own Geometry (reconSize [nx ny], detector, views); images are packed
back to back in slice order ([nx ny] each, x fastest) and so are sinograms ([nu nviews]
each).  This is the batched form of geom/fbct_os's fbct_geom_mex operators
(fbct_proj.c:155-466, fbct_back.c:3-399): one launch projects or backprojects every slice.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import F32, F64, CtsimGeom, call, lib
from .device import DeviceArray, dtype_code
from .geometry import Geometry


class Batch2:
    def __init__(self, geoms: list[Geometry], compute: str = "f32"):
        if not geoms:
            raise ValueError("empty batch")
        for g in geoms:
            if g.ndim != 2:
                raise ValueError("batched slices are 2D fan-beam geometries")
        self.geoms = list(geoms)
        self.compute = dtype_code(compute)
        n = len(geoms)
        arr = (CtsimGeom * n)(*[g.to_c() for g in geoms])
        nvs = np.array([g.noViews for g in geoms], np.int32)
        self._betas = [np.ascontiguousarray(g.betas, np.float64) for g in geoms]
        bptrs = (C.POINTER(C.c_double) * n)(*[b.ctypes.data_as(C.POINTER(C.c_double)) for b in self._betas])
        h = C.c_void_p()
        call("ctsim_batch2_create", C.byref(h), n, arr, nvs.ctypes.data_as(C.POINTER(C.c_int32)), bptrs,
             self.compute)
        self.handle = h.value
        ni, ns = C.c_int64(), C.c_int64()
        call("ctsim_batch2_sizes", C.c_void_p(self.handle), C.byref(ni), C.byref(ns))
        self.img_elems, self.sino_elems = ni.value, ns.value
        self.img_offsets = np.cumsum([0] + [g.reconSize[0] * g.reconSize[1] for g in geoms])
        self.sino_offsets = np.cumsum([0] + [g.detSize[0] * g.noViews for g in geoms])

    # packing ---------------------------------------------------------------------
    def pack_images(self, imgs) -> np.ndarray:
        return np.concatenate([np.asarray(im, np.float32).ravel(order="F") for im in imgs])

    def unpack_sinos(self, flat: np.ndarray):
        return [flat[self.sino_offsets[i]:self.sino_offsets[i + 1]].reshape((g.detSize[0], g.noViews), order="F")
                for i, g in enumerate(self.geoms)]

    def unpack_images(self, flat: np.ndarray):
        return [flat[self.img_offsets[i]:self.img_offsets[i + 1]].reshape(tuple(g.reconSize), order="F")
                for i, g in enumerate(self.geoms)]

    # operators -------------------------------------------------------------------
    def proj(self, d_img: DeviceArray, d_sino: DeviceArray, op: str = "proj,dd", stream=None):
        self._check(d_img, d_sino)
        call("ctsim_batch2_proj", C.c_void_p(self.handle), op.encode(), d_img.vp, d_img.dtype, d_sino.vp,
             d_sino.dtype, C.c_void_p(stream.handle if stream else 0))

    def back(self, d_sino: DeviceArray, d_img: DeviceArray, op: str = "back,dd", stream=None):
        self._check(d_img, d_sino)
        call("ctsim_batch2_back", C.c_void_p(self.handle), op.encode(), d_sino.vp, d_sino.dtype, d_img.vp,
             d_img.dtype, C.c_void_p(stream.handle if stream else 0))

    def _check(self, d_img, d_sino):
        if d_img.count < self.img_elems or d_sino.count < self.sino_elems:
            raise ValueError(f"batch buffers need {self.img_elems} image and {self.sino_elems} sinogram elements")

    # schedule --------------------------------------------------------------------
    @property
    def schedule(self) -> tuple[int, int]:
        a, b = C.c_int32(), C.c_int32()
        call("ctsim_batch2_get_schedule", C.c_void_p(self.handle), C.byref(a), C.byref(b))
        return a.value, b.value

    def set_schedule(self, proj_split: int, back_split: int) -> None:
        call("ctsim_batch2_set_schedule", C.c_void_p(self.handle), int(proj_split), int(back_split))

    def autotune(self, d_img: DeviceArray, d_sino: DeviceArray, d_img_out: DeviceArray, proj_op: str = "proj,dd",
                 back_op: str = "back,dd", reps: int = 2, stream=None) -> dict:
        """time every schedule on these buffers and keep the fastest (overwrites d_sino and d_img_out)"""
        self._check(d_img, d_sino)
        self._check(d_img_out, d_sino)
        ms = (C.c_float * 2)()
        call("ctsim_batch2_autotune", C.c_void_p(self.handle), proj_op.encode(), back_op.encode(), d_img.vp,
             d_img.dtype, d_sino.vp, d_sino.dtype, d_img_out.vp, int(reps), ms,
             C.c_void_p(stream.handle if stream else 0))
        ps, bs = self.schedule
        return dict(proj_split=ps, back_split=bs, proj_ms=float(ms[0]), back_ms=float(ms[1]))

    def __del__(self):
        try:
            if self.handle:
                lib().ctsim_batch2_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass
