"""Device-resident PWLS ordered-subset reconstruction.

PwlsOS restates statsrecon/pwls/reconPwlsSeNesterovSqs.m:37-137 (PwlsSQS the plain SQS
loop of statsrecon/pwls/reconPwlsSeSqs.m:36-110, PwlsLALM reconPwlsLALMOs.m, PwlsNUSQS
reconPwlsSeNusqs.m; each can replay its iterations as captured hipGraphs) with the separable-footprint
operators of geom/shell/loadPojectors.m and the Huber penalty of penalty/huberPenalty.m:

    precom = At(w .* A(ones))
    per subset:  d = Aos(z) - Os(y);  g = Atos(Os(w) .* d) + beta S(z);
                 du = g ./ (precom + beta T(x));  x = z - du;  s -= t du;
                 t' = (1 + sqrt(1 + 4 t^2)) / 2;  z = (1 - 1/t') x + s / t';  clamp
    per iteration: phi = sum_sub sum(w d^2)/2 + beta R(x);  rmsd = rms(x - xold)

Everything stays in HBM.  Sinogram-side data (y, w, the residual) is stored in
subset-major order, so Os(., isub) is a contiguous slice.

Three forms:
  * plain:  FP32 storage everywhere, one kernel launch per operator call;
  * mixed:  y, w and the residual in FP16 or BF16, FP32 footprint sums and volumes,
            FP64 step denominators and reductions, and each iteration replayed as one
            captured hipGraph.  By default the stored data are y - A x0 (x0 the start
            image, projected once in FP32), so that the half-precision rounding of the
            data is taken on the much smaller y - A x0: A is linear, so
            A z - y = (A z - A x0) - (y - A x0) exactly.  offset_mode "residual" (the
            default) projects z and subtracts the stored FP32 A x0 of the subset (10 MB)
            before the residual; "input" has the update kernel write z - x0 for the
            projector (two extra volume passes per subset).  offset=False stores y itself
            and projects an FP16 copy of z;
  * fp64:   every array, footprint and sum in double (precision="f64"), the accuracy
            reference the other two are measured against.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import F32, F64, call
from .device import DeviceArray, Graph, Stream, dtype_code
from .geometry import Geometry
from .projector import Projector
from .shell import subset_views


PENALTIES = {"huber": 0, "quad": 1, "hyper": 2, "aniso": 3, "isohuber": 4, "tv": 5, "isotv": 6}


def penalty_code(pfun) -> int:
    """penalty name of penalty/loadPenaltyOperator.m:22-50 ('huber', 'quad', 'hyper', 'aniso',
    'isohuber', 'tv', 'isotv') -> CTSIM_PEN_*"""
    if isinstance(pfun, int):
        return pfun
    try:
        return PENALTIES[pfun.lower()]
    except KeyError:
        raise ValueError("unknow penalty function") from None


def nesterov_t_sequence(n_steps: int, restart_every: int | None = None, steps_per_iter: int = 1) -> np.ndarray:
    """t_0 = 1, t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 (reconPwlsSeNesterovSqs.m:55,74); restarts to 1."""
    t = np.empty(n_steps + 1, np.float64)
    t[0] = 1.0
    for k in range(n_steps):
        t[k + 1] = (1.0 + math.sqrt(1.0 + 4.0 * t[k] * t[k])) / 2.0
    return t


class PwlsOS:
    def __init__(self, geom: Geometry, y: np.ndarray, w: np.ndarray, numos: int, beta: float, delta: float,
                 mixed: bool = True, sino_dtype: str = "f16", zcopy_dtype: str | None = None,
                 graph: bool | None = None, op: str = "tf", stream: Stream | None = None,
                 precision: str | None = None, pfun: str = "huber", offset: bool | None = None,
                 fused: bool = False, offset_mode: str = "residual"):
        if geom.ndim != 3:
            raise ValueError("PwlsOS drives the 3D cone-beam operators")
        self.f64 = precision == "f64"
        if precision not in (None, "f64", "f32", "mixed"):
            raise ValueError("precision is 'mixed', 'f32' or 'f64'")
        if precision is not None:
            mixed = precision == "mixed"
        self.geom = geom
        self.P = Projector(geom, compute="f64" if self.f64 else "f32")
        self.vdt = F64 if self.f64 else F32  # volume dtype
        self.M = int(numos)
        self.beta = float(beta)
        self.delta = float(delta)
        self.pfun = penalty_code(pfun)
        self.mixed = bool(mixed)
        self.offset = self.mixed and (True if offset is None else bool(offset))
        if offset_mode not in ("residual", "input"):
            raise ValueError("offset_mode is 'residual' or 'input'")
        # residual: A x0 stored and subtracted from A z; input: the projector reads z - x0
        self.offset_input = self.offset and offset_mode == "input"
        self.graph_mode = self.mixed if graph is None else bool(graph)
        if self.f64 and self.graph_mode:
            raise ValueError("the FP64 form runs without graph capture")
        self.kind = op
        self.stream = stream or Stream()
        nx, ny, nz = geom.reconSize
        self.nxyz = (nx, ny, nz)
        nv, nu, nviews = geom.sino_shape()
        self.cell = nu * nv
        self.sdt = dtype_code(sino_dtype) if self.mixed else self.vdt
        if zcopy_dtype is None:
            zcopy_dtype = "f32" if self.offset else "f16"
        self.zdt = dtype_code(zcopy_dtype) if self.mixed else None
        # subset-major permutation of the views
        self.subsets = []
        order = []
        off = 0
        for isub in range(1, self.M + 1):
            v0, vs, ns = Projector.subset(nviews, self.M, isub)
            self.subsets.append((v0, vs, ns, off))
            order.append(subset_views(nviews, self.M, isub))
            off += ns
        self.order = np.concatenate(order)
        hdt = np.float64 if self.f64 else np.float32
        y = np.asarray(y, hdt).reshape((nv, nu, nviews), order="F")[:, :, self.order]
        w = np.asarray(w, hdt).reshape((nv, nu, nviews), order="F")[:, :, self.order]
        # half-precision storage keeps w / max(w) (counts reach 2e4; w .* d would overflow
        # FP16); the scale returns in the backprojection factor and in the FP64 cost
        self.wscale = float(w.max()) if (self.mixed and w.size and w.max() > 0) else 1.0
        w = w / hdt(self.wscale) if self.wscale != 1.0 else w
        st = self.stream
        if self.offset:
            # y - A x0 is formed in setup(); y stays on the host in FP32 until then
            self._y_host = np.ascontiguousarray(y.ravel(order="F"), np.float32)
            self.y = DeviceArray(y.size, self.sdt)
        else:
            self.y = DeviceArray.from_numpy(y.ravel(order="F"), self.sdt, stream=st)
        self.w = DeviceArray.from_numpy(w.ravel(order="F"), self.sdt, stream=st)
        nmax = max(s[2] for s in self.subsets)
        nvox = nx * ny * nz
        vdt = self.vdt
        self.p = DeviceArray(nmax * self.cell, vdt)  # Aos(z), FP32 (or FP64) so that d = p - y keeps its bits
        self.r = DeviceArray(nmax * self.cell, self.sdt)  # Os(w) .* d
        self.x = [DeviceArray(nvox, vdt), DeviceArray(nvox, vdt)]
        self.z = [DeviceArray(nvox, vdt), DeviceArray(nvox, vdt)]
        self.s = DeviceArray(nvox, vdt)
        self.g = DeviceArray(nvox, vdt)
        self.precom = DeviceArray(nvox, vdt)
        self.xold = DeviceArray(nvox, vdt)
        # projector input copy: z - x0 (offset_mode "input") or a half-precision z (offset=False)
        self.zc = DeviceArray(nvox, self.zdt) if self.mixed and (self.offset_input or not self.offset) else None
        # A x0 of every view, subset-major, FP32 (offset_mode "residual")
        self.ax0 = DeviceArray(nviews * self.cell, F32) if self.offset and not self.offset_input else None
        # backprojection and update in one kernel (ctsim_plan_back_nesterov), on request only:
        # FP32, Huber, detectors of at most 64 rows, an FP32 projector input copy. At 512x512x128
        # with 41 views per subset it takes 1.53 ms against 1.02 + 0.30 ms for the two kernels
        # apart (the penalty's neighbour reads stall the backprojector's waves), so it is off
        can_fuse = (not self.f64 and self.pfun == 0 and op in ("dd", "tf") and nv <= 64 and nz <= 512
                    and (self.zc is None or self.zdt == F32))
        if fused and not can_fuse:
            raise ValueError("fused=True needs FP32 volumes, the Huber penalty, the dd or tf operator, a detector "
                             "of at most 64 rows, at most 512 slices and an FP32 projector input copy")
        self.fused = bool(fused)
        self.x0 = DeviceArray(nvox, F32) if self.offset_input else None  # reference image of the input form
        self.scal = DeviceArray(4, "f64")  # phi, R, rmsd numerator, spare
        self.cur = 0  # which of the x / z double buffers is current
        self.t = 1.0
        self.itn = 0
        self._graph = None
        self._tseq = None
        self._base = None

    # ------------------------------------------------------------------------------
    def _sv(self):
        return C.c_void_p(self.stream.handle)

    def _sub_ptr(self, arr: DeviceArray, off_views: int) -> C.c_void_p:
        esz = 2 if arr.dtype in (1, 2) else (8 if arr.dtype == 3 else 4)
        return C.c_void_p(arr.ptr + off_views * self.cell * esz)

    class _View:
        """a DeviceArray-like window for the plan API (pointer, count, dtype)"""

        def __init__(self, ptr, count, dtype):
            self.ptr, self.count, self.dtype = ptr, count, dtype

        @property
        def vp(self):
            return C.c_void_p(self.ptr)

    def _window(self, arr: DeviceArray, off_views: int, nviews: int):
        return PwlsOS._View(self._sub_ptr(arr, off_views).value, nviews * self.cell, arr.dtype)

    def setup(self, x0: np.ndarray) -> None:
        """x = z = s = x0 (MATLAB order), precom = At(w .* A(ones))."""
        st = self.stream
        vdt = self.vdt
        x0d = self.P.upload_volume(np.asarray(x0, np.float64 if self.f64 else np.float32), vdt, stream=st)
        for a in (self.x[0], self.z[0], self.s):
            a.copy_from(x0d, stream=st)
        ones = DeviceArray(self.P.n_img, vdt).fill(1.0, stream=st)
        a = DeviceArray(max(s[2] for s in self.subsets) * self.cell, vdt)
        wa = DeviceArray(a.count, vdt)
        if self.offset:
            # stored data y - A x0 (FP32 projection, one rounding to the storage dtype); the
            # input form's projector input z - x0 starts at zero, the residual form keeps A x0
            if self.offset_input:
                self.x0.copy_from(x0d, stream=st)
                self.zc.fill(0.0, stream=st)
            yf = DeviceArray.from_numpy(self._y_host, F32, stream=st)
            for v0, vs, ns, off in self.subsets:
                self.P.proj(x0d, a, f"proj,{self.kind}", v0, vs, ns, stream=st)
                call("ctsim_sub", self._sub_ptr(yf, off), a.vp, self._sub_ptr(self.y, off), self.y.dtype,
                     ns * self.cell, self._sv())
                if self.ax0 is not None:
                    DeviceArray.window(self.ax0, off * self.cell, ns * self.cell).copy_from(
                        DeviceArray.window(a, 0, ns * self.cell), stream=st)
            st.synchronize()
            del yf
        elif self.zc is not None:
            self.zc.copy_from(self.z[0], stream=st)
        # At(w .* A(ones)) summed subset by subset: every view is used once, scale 1
        for k, (v0, vs, ns, off) in enumerate(self.subsets):
            self.P.proj(ones, a, f"proj,{self.kind}", v0, vs, ns, stream=st)
            call("ctsim_mul", a.vp, vdt, self._sub_ptr(self.w, off), self.w.dtype, wa.vp, vdt, ns * self.cell,
                 self._sv())
            self.P.back(self._View(wa.ptr, ns * self.cell, vdt), self.precom, f"back,{self.kind}", v0, vs, ns,
                        scale=self.wscale, accumulate=k > 0, stream=st)
        self.scal.fill(0.0, stream=st)  # phi, R, rmsd accumulate from zero
        st.synchronize()
        self.cur = 0
        self.t = 1.0
        self.itn = 0

    # ------------------------------------------------------------------------------
    def _subset_step(self, k: int, t: float | None, t_next: float | None) -> None:
        """one ordered subset (index k into self.subsets) on the stream"""
        nx, ny, nz = self.nxyz
        v0, vs, ns, off = self.subsets[k]
        src = self.zc if self.zc is not None else self.z[self.cur]
        pw = self._window(self.p, 0, ns)
        self.P.proj(src, pw, f"proj,{self.kind}", v0, vs, ns, stream=self.stream)
        if self.ax0 is not None:
            # A z - A x0, in place (the stored data are y - A x0)
            call("ctsim_sub", pw.vp, self._sub_ptr(self.ax0, off), pw.vp, F32, ns * self.cell, self._sv())
        call("ctsim_pwls_residual", pw.vp, self.p.dtype, self._sub_ptr(self.y, off), self.y.dtype,
             self._sub_ptr(self.w, off), self.w.dtype, self.r.vp, self.r.dtype, ns * self.cell, self.wscale,
             C.c_void_p(self.scal.ptr), self._sv())
        a, b = self.cur, 1 - self.cur
        zc = self.zc.vp if self.zc is not None else None
        zcdt = self.zc.dtype if self.zc is not None else 0
        zref = self.x0.vp if self.x0 is not None else None
        if self.fused:
            rw = self._window(self.r, 0, ns)
            seq = t is None
            call("ctsim_plan_back_nesterov", C.c_void_p(self.P.handle), f"back,{self.kind}".encode(), v0, vs, ns,
                 rw.vp, rw.dtype, float(self.M) * self.wscale, self.x[a].vp, self.z[a].vp, self.s.vp,
                 self.precom.vp, self.x[b].vp, self.z[b].vp, zc, zref, self.beta, self.delta, self.pfun,
                 1.0 if seq else t, 1.0 if seq else t_next, self._tseq.vp if seq else None,
                 self._base.vp if seq else None, k if seq else 0, self._sv())
            self.cur = b
            return
        self.P.back(self._window(self.r, 0, ns), self.g, f"back,{self.kind}", v0, vs, ns,
                    scale=float(self.M) * self.wscale, stream=self.stream)
        if self.f64:
            call("ctsim_pwls_nesterov_update_f64", self.x[a].vp, self.z[a].vp, self.s.vp, self.g.vp,
                 self.precom.vp, self.x[b].vp, self.z[b].vp, nx, ny, nz, self.beta, self.delta, self.pfun, t, t_next,
                 self._sv())
        elif t is None:
            call("ctsim_pwls_nesterov_update_seq", self.x[a].vp, self.z[a].vp, self.s.vp, self.g.vp, self.precom.vp,
                 self.x[b].vp, self.z[b].vp, zc, zcdt, zref, nx, ny, nz, self.beta, self.delta, self.pfun, self._tseq.vp,
                 self._base.vp, k, self._sv())
        else:
            call("ctsim_pwls_nesterov_update", self.x[a].vp, self.z[a].vp, self.s.vp, self.g.vp, self.precom.vp,
                 self.x[b].vp, self.z[b].vp, zc, zcdt, zref, nx, ny, nz, self.beta, self.delta, self.pfun, t, t_next,
                 self._sv())
        self.cur = b

    def _iteration_body(self, device_t: bool) -> None:
        st = self.stream
        nx, ny, nz = self.nxyz
        self.xold.copy_from(self.x[self.cur], stream=st)
        call("ctsim_fill", C.c_void_p(self.scal.ptr), 3, 1, 0.0, self._sv())  # phi = 0 at isub == 1
        for k in range(self.M):
            if device_t:
                self._subset_step(k, None, None)
            else:
                t_next = (1.0 + math.sqrt(1.0 + 4.0 * self.t * self.t)) / 2.0
                self._subset_step(k, self.t, t_next)
                self.t = t_next
        xc = self.x[self.cur]
        sfx = "_f64" if self.f64 else ""
        call("ctsim_penalty" + sfx, xc.vp, nx, ny, nz, self.pfun, self.delta, 0, None, C.c_void_p(self.scal.ptr + 8),
             self._sv())
        call("ctsim_sqdiff_sum" + sfx, xc.vp, self.xold.vp, xc.count, C.c_void_p(self.scal.ptr + 16), self._sv())
        if device_t:
            call("ctsim_add_i32", self._base.vp, self.M, self._sv())

    def _cost_tail(self) -> None:
        """R(x) and the RMSD numerator of the iteration into the FP64 scalars (on the stream)"""
        nx, ny, nz = self.nxyz
        xc = self.x[self.cur]
        sfx = "_f64" if self.f64 else ""
        call("ctsim_penalty" + sfx, xc.vp, nx, ny, nz, self.pfun, self.delta, 0, None, C.c_void_p(self.scal.ptr + 8),
             self._sv())
        call("ctsim_sqdiff_sum" + sfx, xc.vp, self.xold.vp, xc.count, C.c_void_p(self.scal.ptr + 16), self._sv())

    def _scalars(self) -> dict:
        """phi, beta R(x) and rmsd of the finished iteration (after a stream synchronise)"""
        sc = self.scal.to_numpy()
        return dict(phi=float(sc[0] + self.beta * sc[1]), penalty=float(self.beta * sc[1]),
                    rmsd=float(math.sqrt(sc[2] / self.P.n_img)))

    def autotune(self, reps: int = 3) -> dict:
        """Pick the forward and backward launch schedules with the autotuner on the first
        subset's launches (cached per geometry, view selection and dtypes; see
        ctsim_amd.autotune).  Call after setup() and before prepare(): a captured iteration
        keeps the schedules it was captured with."""
        from .autotune import tuned
        v0, vs, ns, _ = self.subsets[0]
        src = self.zc if self.zc is not None else self.z[self.cur]
        pw = DeviceArray.window(self.p, 0, ns * self.cell)
        rw = DeviceArray.window(self.r, 0, ns * self.cell)
        f = tuned(self.P, src, pw, f"proj,{self.kind}", v0, vs, ns, reps=reps, stream=self.stream)
        b = tuned(self.P, rw, self.g, f"back,{self.kind}", v0, vs, ns, reps=reps, stream=self.stream)
        self.stream.synchronize()
        return {"proj": f, "back": b}

    def prepare(self, itnlim: int) -> None:
        """for the graph form: momentum sequence on the device, one iteration captured"""
        if not self.graph_mode:
            return
        seq = nesterov_t_sequence(itnlim * self.M + 1)
        self._tseq = DeviceArray.from_numpy(seq, "f64", stream=self.stream)
        self._base = DeviceArray(1, "f32")  # one int32 counter
        call("ctsim_fill", self._base.vp, F32, 1, 0.0, self._sv())  # 0.0f has the bits of int 0
        if self.M % 2:
            raise ValueError("graph replay needs an even number of subsets (x/z double buffers)")
        cur0 = self.cur
        with Graph(self.stream) as gr:
            self._iteration_body(device_t=True)
        self._graph = gr
        self.cur = cur0  # capture did not execute anything

    def iterate(self) -> dict:
        """one PWLS-OS iteration; returns phi, beta R(x), rmsd (read back after the iteration)"""
        if self.graph_mode:
            if self._graph is None:
                raise RuntimeError("call prepare(itnlim) before iterating the graph form")
            self._graph.launch()
        else:
            self._iteration_body(device_t=False)
        self.itn += 1
        self.stream.synchronize()
        sc = self.scal.to_numpy()
        nvox = self.P.n_img
        return dict(phi=float(sc[0] + self.beta * sc[1]), penalty=float(self.beta * sc[1]),
                    rmsd=float(math.sqrt(sc[2] / nvox)))

    def image(self) -> np.ndarray:
        """current x in MATLAB order"""
        self.stream.synchronize()
        return self.P.download_volume(self.x[self.cur], stream=self.stream)

    def run(self, x0: np.ndarray, itnlim: int, stop_crit: float = 1e-10):
        """reconPwlsSeNesterovSqs main loop (restart every 2*itnlim never triggers within itnlim)"""
        self.setup(x0)
        self.prepare(itnlim)
        phis, rmsds = [], []
        for itn in range(1, itnlim + 1):
            out = self.iterate()
            phis.append(out["phi"])
            rmsds.append(out["rmsd"])
            if itn > 5 and out["rmsd"] < stop_crit:
                break
        return self.image(), np.array(phis), np.array(rmsds)


class PwlsLALM:
    """LALM-OS on the device (statsrecon/pwls/reconPwlsLALMOs.m:37-110), FP32 storage:

        precom = At(w .* A(ones));  g = 0
        per subset:  xold = x;  r = Aos(x) - Os(y);  dL = Atos(Os(w) .* r)   (x M)
                     g = rho/(rho+1) dL + g/(rho+1);  s = rho dL + (1-rho) g
                     10 x:  x = max(x - (beta S(x) + rho precom (x - xold) + s) /
                                        (beta T(x) + rho precom), 0)
        per iteration: phi = sum_sub sum(w r^2)/2 + beta R(x);  rmsd = rms(x - xold)

    y and w sit in subset-major order as in PwlsOS; the inner steps ping-pong between two
    volume buffers."""

    def __init__(self, geom: Geometry, y: np.ndarray, w: np.ndarray, numos: int, beta: float, delta: float,
                 rho: float = 0.1, pfun: str = "huber", op: str = "tf", inner: int = 10, stream: Stream | None = None,
                 graph: bool = False):
        self.base = PwlsOS(geom, y, w, numos, beta, delta, precision="f32", graph=False, op=op, stream=stream,
                           pfun=pfun)
        self.rho = float(rho)
        self.inner = int(inner)
        nvox = self.base.P.n_img
        self.gacc = DeviceArray(nvox, F32)
        self.sdir = DeviceArray(nvox, F32)
        self.graph = _IterationGraph(self.base, self._body) if graph else None

    def setup(self, x0: np.ndarray) -> None:
        self.base.setup(x0)
        self.gacc.fill(0.0, stream=self.base.stream)
        self.itn = 0

    def iterate(self) -> dict:
        b = self.base
        if self.graph is not None:
            self.graph.run()
        else:
            self._body()
        b.stream.synchronize()
        self.itn += 1
        return b._scalars()

    def _body(self) -> None:
        """one iteration on the stream (every scalar is fixed, so it can be captured once)"""
        b = self.base
        st = b.stream
        nx, ny, nz = b.nxyz
        call("ctsim_fill", C.c_void_p(b.scal.ptr), 3, 1, 0.0, b._sv())
        for k, (v0, vs, ns, off) in enumerate(b.subsets):
            x = b.x[b.cur]
            b.xold.copy_from(x, stream=st)
            pw = b._window(b.p, 0, ns)
            b.P.proj(x, pw, f"proj,{b.kind}", v0, vs, ns, stream=st)
            call("ctsim_pwls_residual", pw.vp, b.p.dtype, b._sub_ptr(b.y, off), b.y.dtype, b._sub_ptr(b.w, off),
                 b.w.dtype, b.r.vp, b.r.dtype, ns * b.cell, b.wscale, C.c_void_p(b.scal.ptr), b._sv())
            b.P.back(b._window(b.r, 0, ns), b.g, f"back,{b.kind}", v0, vs, ns, scale=float(b.M) * b.wscale, stream=st)
            call("ctsim_pwls_lalm_sg", b.g.vp, self.gacc.vp, self.sdir.vp, b.g.count, self.rho, b._sv())
            for _ in range(self.inner):
                src, dst = b.x[b.cur], b.x[1 - b.cur]
                call("ctsim_pwls_lalm_inner", src.vp, b.xold.vp, self.sdir.vp, b.precom.vp, dst.vp, nx, ny, nz,
                     b.beta, b.delta, b.pfun, self.rho, b._sv())
                b.cur = 1 - b.cur
        b._cost_tail()

    def image(self) -> np.ndarray:
        return self.base.image()

    def run(self, x0: np.ndarray, itnlim: int, stop_crit: float = 1e-5):
        """reconPwlsLALMOs main loop (stops after iteration 5 once rmsd < stop_crit)"""
        self.setup(x0)
        phis, rmsds = [], []
        for itn in range(1, itnlim + 1):
            out = self.iterate()
            phis.append(out["phi"])
            rmsds.append(out["rmsd"])
            if itn > 5 and out["rmsd"] < stop_crit:
                break
        return self.image(), np.array(phis), np.array(rmsds)


class PwlsNUSQS:
    """Spatially non-uniform SQS on the device (statsrecon/pwls/reconPwlsSeNusqs.m:32-121), FP32:

        v = max(2 (|G1 x0| + |G2 x0|) + |x0|, 1e-3)            Sobel seed (:47-51)
        dL = At(w .* A(v)) ./ v                                (:54)
        itn = 1, 4, 7:  dLt = dL;  dRt = beta H(v)             (:66-68)
        itn = 2, 5, 8:  xref = x                               (:69-70)
        itn = 3, 6, 9:  v = DRA(max(|x - xref|, 1e-3), 40, 0.05);  dLt = At(w .* A(v)) ./ v   (:71-75)
        per subset:     x = max(x - (Atos(Os(w) .* (Aos(x) - Os(y))) + beta S(x)) ./ (dLt + dRt), 0)

    At(w .* A(v)) is summed subset by subset over the subset-major y/w of PwlsOS (every view
    once, scale 1).  H is the Huber curvature of huberPenalty.m mode 3 whatever pfun is, as
    loadPenaltyOperator.m:54 defines it."""

    EPS = 1e-3
    NLOOP, NFIX = 3, 10

    def __init__(self, geom: Geometry, y: np.ndarray, w: np.ndarray, numos: int, beta: float, delta: float,
                 pfun: str = "huber", op: str = "tf", stream: Stream | None = None, graph: bool = False):
        self.base = PwlsOS(geom, y, w, numos, beta, delta, precision="f32", graph=False, op=op, stream=stream,
                           pfun=pfun)
        if graph and self.base.M % 2:
            raise ValueError("graph replay needs an even number of subsets (x double buffers)")
        self.graph = _IterationGraph(self.base, self._body) if graph else None
        nvox = self.base.P.n_img
        self.v = DeviceArray(nvox, F32)
        self.dl = DeviceArray(nvox, F32)    # dL of the seed image
        self.dlt = DeviceArray(nvox, F32)   # dLtilde in use
        self.h = DeviceArray(nvox, F32)     # H(v); the update scales it by beta
        self.xref = DeviceArray(nvox, F32)
        self.tmp = DeviceArray(nvox, F32)

    def _data_curvature(self, out: DeviceArray) -> None:
        """out = At(w .* A(v)) ./ v"""
        b = self.base
        st = b.stream
        for k, (v0, vs, ns, off) in enumerate(b.subsets):
            pw = b._window(b.p, 0, ns)
            b.P.proj(self.v, pw, f"proj,{b.kind}", v0, vs, ns, stream=st)
            call("ctsim_mul", pw.vp, F32, b._sub_ptr(b.w, off), b.w.dtype, pw.vp, F32, ns * b.cell, b._sv())
            b.P.back(pw, self.tmp, f"back,{b.kind}", v0, vs, ns, scale=b.wscale, accumulate=k > 0, stream=st)
        call("ctsim_div", self.tmp.vp, self.v.vp, out.vp, out.count, b._sv())

    def setup(self, x0: np.ndarray) -> None:
        b = self.base
        b.setup(x0)
        nx, ny, nz = b.nxyz
        call("ctsim_nusqs_seed", b.x[b.cur].vp, self.v.vp, nx, ny, nz, self.EPS, b._sv())
        self._data_curvature(self.dl)
        b.stream.synchronize()
        self.itn = 0

    def iterate(self) -> dict:
        b = self.base
        st = b.stream
        nx, ny, nz = b.nxyz
        itn = self.itn + 1
        L, F = self.NLOOP, self.NFIX
        if itn % L == 1 and itn < F:
            self.dlt.copy_from(self.dl, stream=st)
            call("ctsim_penalty", self.v.vp, nx, ny, nz, 0, b.delta, 3, self.h.vp, None, b._sv())
        elif itn % L == L - 1 and itn <= F - 2:
            self.xref.copy_from(b.x[b.cur], stream=st)
        elif itn % L == 0 and itn <= F - 1:
            call("ctsim_absdiff_floor", b.x[b.cur].vp, self.xref.vp, self.tmp.vp, self.tmp.count, self.EPS, b._sv())
            call("ctsim_dynamic_range", self.tmp.vp, self.tmp.count, 40.0, 0.05, self.v.vp, b._sv())
            self._data_curvature(self.dlt)
        if self.graph is not None:
            self.graph.run()
        else:
            self._body()
        st.synchronize()
        self.itn = itn
        return b._scalars()

    def _body(self) -> None:
        """the subset loop of one iteration with the current dLtilde and H (capturable: the
        schedule above changes buffer contents only)"""
        b = self.base
        st = b.stream
        nx, ny, nz = b.nxyz
        b.xold.copy_from(b.x[b.cur], stream=st)
        call("ctsim_fill", C.c_void_p(b.scal.ptr), 3, 1, 0.0, b._sv())
        for k, (v0, vs, ns, off) in enumerate(b.subsets):
            x = b.x[b.cur]
            pw = b._window(b.p, 0, ns)
            b.P.proj(x, pw, f"proj,{b.kind}", v0, vs, ns, stream=st)
            call("ctsim_pwls_residual", pw.vp, b.p.dtype, b._sub_ptr(b.y, off), b.y.dtype, b._sub_ptr(b.w, off),
                 b.w.dtype, b.r.vp, b.r.dtype, ns * b.cell, b.wscale, C.c_void_p(b.scal.ptr), b._sv())
            b.P.back(b._window(b.r, 0, ns), b.g, f"back,{b.kind}", v0, vs, ns, scale=float(b.M) * b.wscale, stream=st)
            call("ctsim_pwls_nusqs_update", x.vp, b.x[1 - b.cur].vp, b.g.vp, self.dlt.vp, self.h.vp, nx, ny, nz,
                 b.beta, b.delta, b.pfun, b._sv())
            b.cur = 1 - b.cur
        b._cost_tail()

    def image(self) -> np.ndarray:
        return self.base.image()

    def run(self, x0: np.ndarray, itnlim: int, stop_crit: float = 1e-5):
        """reconPwlsSeNusqs main loop (stops after iteration 5 once rmsd < stop_crit)"""
        self.setup(x0)
        phis, rmsds = [], []
        for itn in range(1, itnlim + 1):
            out = self.iterate()
            phis.append(out["phi"])
            rmsds.append(out["rmsd"])
            if itn > 5 and out["rmsd"] < stop_crit:
                break
        return self.image(), np.array(phis), np.array(rmsds)


class _IterationGraph:
    """One captured hipGraph of an iteration body that ping-pongs the x buffers of a PwlsOS
    base.  The body is captured on the first run; its buffer choice is the one at capture
    time, so the body must swap the buffers an even number of times (then every replay
    starts and ends on the same buffer, as the eager loop does)."""

    def __init__(self, base: PwlsOS, body):
        self.base, self.body = base, body
        self._graph = None

    def run(self) -> None:
        b = self.base
        if self._graph is None:
            cur0 = b.cur
            with Graph(b.stream) as gr:
                self.body()
            if b.cur != cur0:
                raise ValueError("graph replay needs an even number of x buffer swaps per iteration "
                                 "(an even number of subsets)")
            self._graph = gr
        self._graph.launch()


class PwlsSQS:
    """Plain SQS with ordered subsets on the device (statsrecon/pwls/reconPwlsSeSqs.m:36-110),
    FP32 storage, FP64 denominators and reductions:

        precom = At(w .* A(ones))
        per subset:  d = Aos(x) - Os(y);  x = max(x - (Atos(Os(w) .* d) + beta S(x)) ./
                                                      (precom + beta T(x)), 0)
        per iteration: phi = sum_sub sum(w d^2)/2 + beta R(x);  rmsd = rms(x - xold)

    graph=True replays each iteration as one captured hipGraph (an even number of subsets)."""

    def __init__(self, geom: Geometry, y: np.ndarray, w: np.ndarray, numos: int, beta: float, delta: float,
                 pfun: str = "huber", op: str = "tf", stream: Stream | None = None, graph: bool = False):
        self.base = PwlsOS(geom, y, w, numos, beta, delta, precision="f32", graph=False, op=op, stream=stream,
                           pfun=pfun)
        self.graph = _IterationGraph(self.base, self._body) if graph else None
        if graph and self.base.M % 2:
            raise ValueError("graph replay needs an even number of subsets (x double buffers)")

    def setup(self, x0: np.ndarray) -> None:
        self.base.setup(x0)
        self.itn = 0

    def _body(self) -> None:
        b = self.base
        st = b.stream
        nx, ny, nz = b.nxyz
        b.xold.copy_from(b.x[b.cur], stream=st)
        call("ctsim_fill", C.c_void_p(b.scal.ptr), 3, 1, 0.0, b._sv())
        for k, (v0, vs, ns, off) in enumerate(b.subsets):
            x = b.x[b.cur]
            pw = b._window(b.p, 0, ns)
            b.P.proj(x, pw, f"proj,{b.kind}", v0, vs, ns, stream=st)
            call("ctsim_pwls_residual", pw.vp, b.p.dtype, b._sub_ptr(b.y, off), b.y.dtype, b._sub_ptr(b.w, off),
                 b.w.dtype, b.r.vp, b.r.dtype, ns * b.cell, b.wscale, C.c_void_p(b.scal.ptr), b._sv())
            b.P.back(b._window(b.r, 0, ns), b.g, f"back,{b.kind}", v0, vs, ns, scale=float(b.M) * b.wscale, stream=st)
            call("ctsim_pwls_sqs_update", x.vp, b.x[1 - b.cur].vp, b.g.vp, b.precom.vp, nx, ny, nz, b.beta, b.delta,
                 b.pfun, b._sv())
            b.cur = 1 - b.cur
        b._cost_tail()

    def iterate(self) -> dict:
        b = self.base
        if self.graph is not None:
            self.graph.run()
        else:
            self._body()
        b.stream.synchronize()
        self.itn += 1
        return b._scalars()

    def image(self) -> np.ndarray:
        return self.base.image()

    def run(self, x0: np.ndarray, itnlim: int, stop_crit: float = 1e-5):
        """reconPwlsSeSqs main loop (stops after iteration 5 once rmsd < stop_crit, :106)"""
        self.setup(x0)
        phis, rmsds = [], []
        for itn in range(1, itnlim + 1):
            out = self.iterate()
            phis.append(out["phi"])
            rmsds.append(out["rmsd"])
            if itn > 5 and out["rmsd"] < stop_crit:
                break
        return self.image(), np.array(phis), np.array(rmsds)
