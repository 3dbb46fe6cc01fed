"""Device-resident projector: a plan (geometry + per-view table in HBM) and its
operators on DeviceArrays.  3D volumes live in ZXY order (z fastest); sinograms in
[view][u][v] order (MATLAB's [nv nu nviews] memory order)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import F32, F64, call, lib
from .device import DeviceArray, dtype_code
from .geometry import Geometry


class Projector:
    def __init__(self, geom: Geometry, compute="f32", map=None):
        self.geom = geom
        self.ndim = geom.ndim
        self.compute = dtype_code(compute)
        if self.compute not in (F32, F64):
            raise ValueError("compute must be f32 or f64")
        g = geom.to_c()
        self._g = g
        m = None
        mp = map if map is not None else geom.map
        if mp is not None and self.ndim == 3:
            m = np.ascontiguousarray(np.asfortranarray(np.asarray(mp, bool)).ravel(order="F").view(np.uint8))
        self._map = m
        p = C.c_void_p()
        call("ctsim_plan_create", C.byref(p), self.ndim, C.byref(g), geom.noViews,
             geom.betas.ctypes.data_as(C.POINTER(C.c_double)),
             geom.couchZ.ctypes.data_as(C.POINTER(C.c_float)),
             None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8)), self.compute)
        self.handle = p.value

    # sizes -----------------------------------------------------------------------
    @property
    def n_img(self) -> int:
        return int(np.prod(self.geom.reconSize))

    def n_sino(self, nviews=None) -> int:
        return int(np.prod(self.geom.sino_shape(nviews)))

    @staticmethod
    def subset(nviews: int, M: int, isub: int):
        """(view0, vstep, nsub) of ordered subset isub of M (MATLAB 1+mod(i,M):M:end)."""
        if M == 1:
            return 0, 1, nviews
        v0 = isub % M
        return v0, M, len(range(v0, nviews, M))

    # operators ---------------------------------------------------------------------
    def _check(self, img: DeviceArray, sino: DeviceArray, view0, vstep, nsub):
        """host-side shape checks before any launch (the C ABI sees raw pointers)"""
        if img.count != self.n_img:
            raise ValueError(f"image buffer has {img.count} elements, geometry needs {self.n_img}")
        if sino.count < self.n_sino(nsub):
            raise ValueError(f"sinogram buffer has {sino.count} elements, {nsub} views need {self.n_sino(nsub)}")
        if nsub <= 0 or vstep <= 0 or view0 < 0 or view0 + (nsub - 1) * vstep >= self.geom.noViews:
            raise ValueError("view selection outside the plan")

    def proj(self, img: DeviceArray, sino: DeviceArray, op="proj,tf", view0=0, vstep=1, nsub=None, stream=None,
             exhaustive=False):
        nsub = self.geom.noViews if nsub is None else nsub
        self._check(img, sino, view0, vstep, nsub)
        name = "ctsim_plan_proj_exhaustive" if exhaustive else "ctsim_plan_proj"
        call(name, C.c_void_p(self.handle), op.encode(), view0, vstep, nsub, img.vp, img.dtype, sino.vp, sino.dtype,
             C.c_void_p(stream.handle if stream else 0))

    def back(self, sino: DeviceArray, img: DeviceArray, op="back,tf", view0=0, vstep=1, nsub=None, scale=1.0,
             accumulate=False, stream=None):
        nsub = self.geom.noViews if nsub is None else nsub
        self._check(img, sino, view0, vstep, nsub)
        call("ctsim_plan_back", C.c_void_p(self.handle), op.encode(), view0, vstep, nsub, sino.vp, sino.dtype,
             img.vp, img.dtype, float(scale), int(accumulate), C.c_void_p(stream.handle if stream else 0))

    # launch schedules and the autotuner (3D) ------------------------------------------
    def schedule(self) -> tuple:
        """(proj_sched, back_sched) of the plan; see ctsim_plan_set_schedule"""
        a, b = C.c_int32(), C.c_int32()
        call("ctsim_plan_get_schedule", C.c_void_p(self.handle), C.byref(a), C.byref(b))
        return int(a.value), int(b.value)

    def set_schedule(self, proj_sched: int | None = None, back_sched: int | None = None) -> None:
        a, b = self.schedule()
        call("ctsim_plan_set_schedule", C.c_void_p(self.handle), a if proj_sched is None else int(proj_sched),
             b if back_sched is None else int(back_sched))

    def autotune(self, d_in: DeviceArray, d_out: DeviceArray, op: str, view0=0, vstep=1, nsub=None, reps=3,
                 stream=None) -> dict:
        """Time every schedule of op (proj,dd / proj,tf: forward forms and splits; back,dd /
        back,tf: column orders; back,pd / back,hpd: the staged and direct FDK forms) on these
        buffers and keep the fastest in the plan.  Returns
        {"best": code, "candidates": [(code, ms), ...]}."""
        nsub = self.geom.noViews if nsub is None else nsub
        if op.startswith("proj"):
            self._check(d_in, d_out, view0, vstep, nsub)
        else:
            self._check(d_out, d_in, view0, vstep, nsub)
        best = C.c_int32()
        ms = (C.c_float * 8)()
        codes = (C.c_int32 * 8)()
        call("ctsim_plan_autotune", C.c_void_p(self.handle), op.encode(), view0, vstep, nsub, d_in.vp, d_in.dtype,
             d_out.vp, d_out.dtype, int(reps), C.byref(best), ms, codes, 8, C.c_void_p(stream.handle if stream else 0))
        n = 6 if op.startswith("proj") else 2
        return {"best": int(best.value), "candidates": [(int(codes[i]), round(float(ms[i]), 4)) for i in range(n)]}

    # host helpers (MATLAB-order numpy in, MATLAB-order numpy out) --------------------
    def upload_volume(self, vol: np.ndarray, dtype="f32", stream=None) -> DeviceArray:
        """MATLAB [nx ny nz] (or [nx ny]) -> device (ZXY for 3D)."""
        dt = dtype_code(dtype)
        if self.ndim == 2:
            return DeviceArray.from_numpy(np.asarray(vol).ravel(order="F"), dt, stream=stream)
        nx, ny, nz = self.geom.reconSize
        tmp = DeviceArray.from_numpy(np.asarray(vol).ravel(order="F"),
                                     F64 if dt == F64 else F32, stream=stream)
        out = DeviceArray(nx * ny * nz, dt)
        call("ctsim_xyz_to_zxy", tmp.vp, tmp.dtype, out.vp, out.dtype, nx, ny, nz,
             C.c_void_p(stream.handle if stream else 0))
        return out

    def download_volume(self, d: DeviceArray, stream=None) -> np.ndarray:
        if self.ndim == 2:
            return d.to_numpy(stream).reshape(self.geom.reconSize, order="F")
        nx, ny, nz = self.geom.reconSize
        tmp = DeviceArray(nx * ny * nz, F64 if d.dtype == F64 else F32)
        call("ctsim_zxy_to_xyz", d.vp, d.dtype, tmp.vp, tmp.dtype, nx, ny, nz,
             C.c_void_p(stream.handle if stream else 0))
        return tmp.to_numpy(stream).reshape((nx, ny, nz), order="F")

    def upload_sino(self, sino: np.ndarray, dtype="f32", stream=None) -> DeviceArray:
        return DeviceArray.from_numpy(np.asarray(sino).ravel(order="F"), dtype, stream=stream)

    def download_sino(self, d: DeviceArray, nviews=None, stream=None) -> np.ndarray:
        return d.to_numpy(stream).reshape(self.geom.sino_shape(nviews), order="F")

    def __del__(self):
        try:
            if self.handle:
                lib().ctsim_plan_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass
