"""Benchmark: PWLS-OS iterations per second at 512x512x128 (984 views, 1024x64 detector,
24 ordered subsets, Huber beta 1e4 and delta 0.001 /mm = 0.01 in the image's 1/cm units),
mixed-precision graph form vs plain FP32,
plus DD / SF forward+back projection GUPS at 512x512x256 (720 views, 1024x384).

A step is one PWLS-OS iteration: 24 x (forward projection of the subset, residual,
backprojection x 24, fused Nesterov-SQS update) plus the per-iteration penalty cost and
RMSD, all resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--sino-dtype bf16|f16]
                    [--accuracy-iters 20] [--no-pair] [--no-cpu] [--no-plain]

Accuracy: the all-FP64 form (FP64 plan, FP64 elementwise kernels), the plain FP32 form and
the mixed form with BF16 and with FP16 storage run --accuracy-iters iterations in lockstep from
the same FDK image, untimed; the JSON line reports, after every iteration, the relative RMSE
||x - x64|| / ||x64|| over the volume of each form against the FP64 iterate.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_counts():
    with open(os.path.join(ROOT, "tests", "golden", "work_counts.json")) as f:
        return json.load(f)


def flops(c):
    """algorithmic FLOPs of one 3D DD/SF operator: 17 U + 3 C + 60 K"""
    return 17.0 * c["U"] + 3.0 * c["C"] + 60.0 * c["K"]


def measured_traffic():
    """HBM bytes per launch of the forward projector on one subset, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes, each scaled by what the counter reads for a
    copy kernel of known size (profiles/cbct_proj_tf_pwls_subset_traffic.json); None when
    no such measurement is committed"""
    path = os.path.join(ROOT, "profiles", "cbct_proj_tf_pwls_subset_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    fetch = t["fetch_kib"] * t["copy_bytes"] / t["copy_fetch_kib"]
    write = t["write_kib"] * t["copy_bytes"] / t["copy_write_kib"]
    return round(fetch + write)


def simulate_pwls_data(geom, seed=4):
    """body phantom -> SF projection on the GPU -> Poisson counts at I0 = 2e4 -> y, w = counts"""
    from ctsim_amd import workloads as W
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    vol = W.body_phantom(geom.reconSize, geom.reconSpacing, 10, seed)
    P = Projector(geom)
    d = DeviceArray(P.n_sino(), "f32")
    P.proj(P.upload_volume(vol), d, "proj,tf")
    p = P.download_sino(d).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    I0 = 2.0e4
    counts = rng.poisson(I0 * np.exp(-p)).astype(np.float64)
    y = -np.log(np.maximum(counts, 10.0) / I0)  # dummyLowPhotonCount = 10
    w = counts  # computeWeightsPwls with r = 0, v = 0
    return vol, y.astype(np.float32), w.astype(np.float32)


class Timer:
    def __init__(self, stream):
        from ctsim_amd.device import Event
        self.stream = stream
        self.a, self.b = Event(), Event()

    def __enter__(self):
        self.a.record(self.stream)
        return self

    def __exit__(self, *exc):
        self.b.record(self.stream)
        self.ms = self.a.elapsed_ms(self.b)
        return False


def kernel_times(solver, reps=1):
    """average launch duration (ms) of each PWLS kernel over the 24 subsets, by HIP events
    on the solver's stream"""
    import ctypes as C
    from ctsim_amd._lib import call
    st = solver.stream
    nx, ny, nz = solver.nxyz
    t = {"proj": 0.0, "back": 0.0, "residual": 0.0, "update": 0.0}
    src = solver.zc if solver.zc is not None else solver.z[solver.cur]
    for _ in range(reps):
        for k, (v0, vs, ns, off) in enumerate(solver.subsets):
            pw = solver._window(solver.p, 0, ns)
            with Timer(st) as tm:
                solver.P.proj(src, pw, f"proj,{solver.kind}", v0, vs, ns, stream=st)
            t["proj"] += tm.ms
            with Timer(st) as tm:
                if solver.ax0 is not None:
                    call("ctsim_sub", pw.vp, solver._sub_ptr(solver.ax0, off), pw.vp, solver.p.dtype, ns * solver.cell,
                         solver._sv())
                call("ctsim_pwls_residual", pw.vp, solver.p.dtype, solver._sub_ptr(solver.y, off), solver.y.dtype,
                     solver._sub_ptr(solver.w, off), solver.w.dtype, solver.r.vp, solver.r.dtype, ns * solver.cell,
                     solver.wscale, C.c_void_p(solver.scal.ptr), solver._sv())
            t["residual"] += tm.ms
            with Timer(st) as tm:
                solver.P.back(solver._window(solver.r, 0, ns), solver.g, f"back,{solver.kind}", v0, vs, ns,
                              scale=float(solver.M) * solver.wscale, stream=st)
            t["back"] += tm.ms
            a, b = solver.cur, 1 - solver.cur
            with Timer(st) as tm:
                call("ctsim_pwls_nesterov_update", solver.x[a].vp, solver.z[a].vp, solver.s.vp, solver.g.vp,
                     solver.precom.vp, solver.x[b].vp, solver.z[b].vp,
                     solver.zc.vp if solver.zc is not None else None, solver.zc.dtype if solver.zc is not None else 0,
                     solver.x0.vp if solver.x0 is not None else None, nx, ny, nz, solver.beta, solver.delta,
                     solver.pfun, 1.0, 1.0, solver._sv())
            t["update"] += tm.ms
            if solver.fused:
                with Timer(st) as tm:
                    rw = solver._window(solver.r, 0, ns)
                    call("ctsim_plan_back_nesterov", C.c_void_p(solver.P.handle), f"back,{solver.kind}".encode(), v0,
                         vs, ns, rw.vp, rw.dtype, float(solver.M) * solver.wscale, solver.x[a].vp, solver.z[a].vp,
                         solver.s.vp, solver.precom.vp, solver.x[b].vp, solver.z[b].vp,
                         solver.zc.vp if solver.zc is not None else None,
                         solver.x0.vp if solver.x0 is not None else None, solver.beta, solver.delta, solver.pfun,
                         1.0, 1.0, None, None, 0, solver._sv())
                t["back+update fused"] = t.get("back+update fused", 0.0) + tm.ms
    n = reps * len(solver.subsets)
    return {k: v / n for k, v in t.items()}


def run_pwls(form, geom, y, w, x0, steps, warmup, sino_dtype="bf16", numos=24, beta=1e4, delta=1e-2,
             total_iters=0, fused=False):
    """warmup + steps iterations (steps timed) of the form 'mixed', 'plain' or 'f64', then
    untimed iterations up to total_iters"""
    from ctsim_amd.device import synchronize
    from ctsim_amd.pwls import PwlsOS
    precision = {"mixed": "mixed", "plain": "f32", "f64": "f64"}[form]
    n_all = max(warmup + steps, total_iters)
    solver = PwlsOS(geom, y, w, numos, beta, delta, precision=precision, sino_dtype=sino_dtype,
                    graph=form == "mixed", fused=fused)
    solver.setup(x0)
    # the batched (mixed) form takes the autotuner's schedules; the plain form keeps the
    # plan's defaults
    solver.tuning = solver.autotune(reps=3) if form == "mixed" else None
    solver.prepare(n_all)
    for _ in range(warmup):
        solver.iterate()
    synchronize()
    t0 = time.perf_counter()
    outs = [solver.iterate() for _ in range(steps)]
    synchronize()
    dt = time.perf_counter() - t0
    for _ in range(n_all - warmup - steps):
        outs.append(solver.iterate())
    return solver, dt, outs


def rel_rmse(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def pair_gups(kinds=("dd", "tf"), reps=2):
    """forward + back projection of the 512x512x256 / 720-view / 1024x384 pair"""
    from ctsim_amd import workloads as W
    from ctsim_amd.device import DeviceArray, Stream
    from ctsim_amd.projector import Projector
    geom = W.sf_pair_512()
    P = Projector(geom)
    st = Stream()
    rng = np.random.default_rng(1)
    vol = DeviceArray.from_numpy(rng.uniform(0, 1, P.n_img).astype(np.float32), "f32", stream=st)
    sino = DeviceArray(P.n_sino(), "f32")
    img = DeviceArray(P.n_img, "f32")
    counts = load_counts()
    out = {}
    for kind in kinds:
        # the autotuner's schedules for this geometry (times of every candidate reported)
        tp_sched = P.autotune(vol, sino, f"proj,{kind}", reps=1, stream=st)
        tb_sched = P.autotune(sino, img, f"back,{kind}", reps=1, stream=st)
        P.proj(vol, sino, f"proj,{kind}", stream=st)
        P.back(sino, img, f"back,{kind}", stream=st)
        st.synchronize()
        tp = tb = 0.0
        for _ in range(reps):
            with Timer(st) as a:
                P.proj(vol, sino, f"proj,{kind}", stream=st)
            with Timer(st) as b:
                P.back(sino, img, f"back,{kind}", stream=st)
            tp += a.ms
            tb += b.ms
        tp /= reps
        tb /= reps
        c = counts[f"sf_pair_512/{kind}"]
        pt, bt = flops(c) / (tp * 1e-3) / 1e12, flops(c) / (tb * 1e-3) / 1e12
        out[kind] = dict(proj_ms=round(tp, 3), back_ms=round(tb, 3),
                         gups_pair=round(2 * c["U"] / ((tp + tb) * 1e-3) / 1e9, 2),
                         proj_tflops=round(pt, 2), back_tflops=round(bt, 2),
                         proj_frac=round(pt / PEAK_FP32_TFLOPS, 4), back_frac=round(bt / PEAK_FP32_TFLOPS, 4),
                         schedule={"proj": tp_sched, "back": tb_sched,
                                   "codes": "proj form|split<<2 (form 0 auto, 1 ring, 2 register, 3 ring with two row chunks; split 0 auto, "
                                            "1 whole, 2 shared); back 0 linear, 1 super-tile"})
        log(f"pair {kind}: proj {tp:.1f} ms  back {tb:.1f} ms  {out[kind]['gups_pair']} GUPS")
    return out


def fan_pair_bench(reps=5):
    """2D DD forward + matched back projection, FP32: 256^2 at 0.8 mm, 10 random ellipses,
    360 views over 360 degrees, SAD 1000 / SDD 1500 mm, 512 flat bins at 1.0 mm"""
    from ctsim_amd import workloads as W
    from ctsim_amd.device import DeviceArray, Stream
    from ctsim_amd.projector import Projector
    geom = W.fan_256()
    P = Projector(geom)
    st = Stream()
    img = DeviceArray.from_numpy(W.ellipse_image(256, 0.8, 10, 1).ravel(order="F"), "f32", stream=st)
    sino = DeviceArray(P.n_sino(), "f32")
    back = DeviceArray(P.n_img, "f32")
    P.proj(img, sino, "proj,dd", stream=st)
    P.back(sino, back, "back,dd", stream=st)
    tp = tb = 0.0
    for _ in range(reps):
        with Timer(st) as a:
            P.proj(img, sino, "proj,dd", stream=st)
        with Timer(st) as b:
            P.back(sino, back, "back,dd", stream=st)
        tp += a.ms
        tb += b.ms
    tp /= reps
    tb /= reps
    U = 256 * 256 * 360
    return dict(workload="fan_dd_256x256_360views_512bins", proj_ms=round(tp, 4), back_ms=round(tb, 4),
                gups_pair=round(2 * U / ((tp + tb) * 1e-3) / 1e9, 2))


def batch2_bench(reps=3):
    """64 uneven 2D slices (128^2..512^2, 180..984 views, 256..1024 bins) projected and
    backprojected (DD) in one launch each, FP16 images and sinograms, FP32 accumulation,
    launch schedule chosen by the autotuner"""
    from ctsim_amd import workloads as W
    from ctsim_amd.batch2 import Batch2
    from ctsim_amd.device import DeviceArray, Stream
    geoms, imgs = W.batch2_uneven()
    B = Batch2(geoms)
    st = Stream()
    d_img = DeviceArray.from_numpy(B.pack_images(imgs), "f16", stream=st)
    d_sino = DeviceArray(B.sino_elems, "f16")
    d_back = DeviceArray(B.img_elems, "f16")
    default = B.schedule
    tuned = B.autotune(d_img, d_sino, d_back, "proj,dd", "back,dd", reps=2, stream=st)
    tp = tb = 0.0
    for _ in range(reps):
        with Timer(st) as a:
            B.proj(d_img, d_sino, "proj,dd", stream=st)
        with Timer(st) as b:
            B.back(d_sino, d_back, "back,dd", stream=st)
        tp += a.ms
        tb += b.ms
    tp /= reps
    tb /= reps
    U = float(sum(g.reconSize[0] * g.reconSize[1] * g.noViews for g in geoms))
    log(f"batch2: proj {tp:.2f} ms back {tb:.2f} ms, schedule {B.schedule} (default {default})")
    return dict(workload="batch2_64_uneven_slices_dd_fp16", slices=len(geoms), pixel_views=U, proj_ms=round(tp, 3),
                back_ms=round(tb, 3), gups_pair=round(2 * U / ((tp + tb) * 1e-3) / 1e9, 2),
                schedule={"proj_split": tuned["proj_split"], "back_split": tuned["back_split"],
                          "default": list(default)})


def fdk_bench(reps=2):
    """FDK at 512^3 / 0.5 mm from 720 views on a flat 1024x768 detector at 0.388 mm:
    ellipsoid phantom projected on the GPU, Poisson noise at I0 = 1e5, cosine weight and
    Hann ramp filter on the GPU, filtered views stored in FP32 / FP16 / BF16, FP32
    back,hpd accumulation.  Each storage form is compared with the all-FP64 FDK (FP64
    filter and backprojector on the GPU)."""
    from ctsim_amd import workloads as W
    from ctsim_amd.device import DeviceArray, Stream
    from ctsim_amd.fdk import FdkFilter, fdk_scale
    from ctsim_amd.projector import Projector
    geom = W.fdk_512()
    P = Projector(geom)
    st = Stream()
    vol = W.ellipsoid_phantom(geom.reconSize, geom.reconSpacing, 20, 3)
    d_p = DeviceArray(P.n_sino(), "f32")
    P.proj(P.upload_volume(vol, stream=st), d_p, "proj,tf", stream=st)
    p = d_p.to_numpy(st)
    del d_p
    rng = np.random.default_rng(12)
    I0 = 1.0e5
    y = np.empty(p.size, np.float32)
    step = 1 << 24
    for a in range(0, p.size, step):
        c = rng.poisson(I0 * np.exp(-p[a:a + step].astype(np.float64)))
        y[a:a + step] = -np.log(np.maximum(c, 10) / I0)
    del p
    d_y = DeviceArray.from_numpy(y, "f32", stream=st)
    scale = fdk_scale(geom)
    U = float(np.prod(geom.reconSize)) * geom.noViews
    out = {}
    imgs = {}
    filt = FdkFilter(geom, "hann")
    for storage in ("f16", "bf16", "f32"):
        d_f = DeviceArray(d_y.count, storage)
        d_v = DeviceArray(P.n_img, "f32")
        filt(d_y, d_f, stream=st)
        P.back(d_f, d_v, "back,hpd", scale=scale, stream=st)
        tf = tb = 0.0
        for _ in range(reps):
            with Timer(st) as a:
                filt(d_y, d_f, stream=st)
            with Timer(st) as b:
                P.back(d_f, d_v, "back,hpd", scale=scale, stream=st)
            tf += a.ms
            tb += b.ms
        tf /= reps
        tb /= reps
        imgs[storage] = P.download_volume(d_v, stream=st)
        bt = 19.0 * U / (tb * 1e-3) / 1e12
        out[storage] = dict(filter_ms=round(tf, 2), back_ms=round(tb, 2), total_ms=round(tf + tb, 2),
                            back_gups=round(U / (tb * 1e-3) / 1e9, 1), back_tflops=round(bt, 2),
                            back_frac=round(bt / PEAK_FP32_TFLOPS, 4))
        log(f"fdk {storage}: filter {tf:.1f} ms  back {tb:.1f} ms")
        del d_f, d_v
    del filt
    P64 = Projector(geom, compute="f64")
    d_y64 = DeviceArray(d_y.count, "f64")
    from ctsim_amd._lib import call
    call("ctsim_convert", d_y.vp, d_y.dtype, d_y64.vp, d_y64.dtype, d_y.count, C.c_void_p(st.handle))
    del d_y
    f64 = FdkFilter(geom, "hann", compute="f64")
    d_f64 = f64(d_y64, DeviceArray(d_y64.count, "f64"), stream=st)
    del d_y64
    d_v64 = DeviceArray(P.n_img, "f64")
    P64.back(d_f64, d_v64, "back,hpd", scale=scale, stream=st)
    ref = P64.download_volume(d_v64, stream=st)
    for storage, img in imgs.items():
        out[storage]["rel_l2_vs_fp64"] = rel_rmse(img, ref)
    log("fdk accuracy:", {k: v["rel_l2_vs_fp64"] for k, v in out.items()})
    return dict(workload="fdk_512x512x512_720views_1024x768det_hann", voxel_views=U, **out)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return os.uname().machine


def cpu_baseline(geom, nsample=8):
    """reference CPU projector (oracle/_ref, else the oracle port), one core, on nsample
    evenly spaced views of the PWLS geometry; extrapolated to one iteration (every view
    forward and back once)."""
    from oracle import oracle as O
    kind = "reference" if O.ref_available() else "port"
    idx = np.linspace(0, geom.noViews - 1, nsample).astype(int)
    g = O.G3(geom.reconSize, geom.reconSpacing, geom.reconOffset, geom.detSize, geom.detSpacing, geom.detOffset,
             geom.SAD, geom.ADD, geom.SDD, geom.betas[idx], couchz=geom.couchZ[idx], dettype=geom.flatPanel)
    rng = np.random.default_rng(2)
    vol = rng.uniform(0, 1, geom.reconSize).astype(np.float32)
    sino = rng.uniform(0, 1, g.sino_shape()).astype(np.float32)
    mp = np.asarray(geom.map, bool)
    run = (lambda op, x: O.ref_cbct(op, g, x, mp)) if kind == "reference" else (
        lambda op, x: O.cbct(op, g, x, mp, nthreads=1))
    t0 = time.perf_counter()
    run("proj,tf", vol)
    t1 = time.perf_counter()
    run("back,tf", sino)
    t2 = time.perf_counter()
    per_iter = (t2 - t0) * geom.noViews / nsample
    return dict(value=round(1.0 / per_iter, 6), unit="iter/s", cores=1, kind=kind, cpu=cpu_model(),
                host_cores=os.cpu_count(),
                sample=f"proj,tf + back,tf of {nsample} of 984 views at 512x512x128 with the FOV map, "
                       f"{t1 - t0:.1f} s + {t2 - t1:.1f} s, scaled x{geom.noViews / nsample:.0f} to one iteration; "
                       f"the elementwise PWLS work (residual, update, penalty) is excluded, which favours the CPU; "
                       f"the reference C compiled with g++ -O2 from its own sources, 1 thread (it is single-threaded)")


def accuracy_lockstep(geom, y, w, x0, n_iter, numos=24, beta=1e4, delta=1e-2):
    """FP64, plain FP32 and mixed (BF16 and FP16 storage) PWLS-OS from the same start, in lockstep;
    relative RMSE against the FP64 iterate after every iteration.  A second FP64 run starts
    from x0 with every non-zero voxel moved by one FP32 ulp (random direction): how far two
    all-FP64 runs drift apart from a start difference at FP32 rounding level is the floor any
    FP32-storing form can reach against the FP64 iterate."""
    from ctsim_amd.pwls import PwlsOS
    rng = np.random.default_rng(77)
    x0 = np.asarray(x0, np.float32)
    away = np.where(rng.uniform(size=x0.shape) < 0.5, np.float32(np.inf), np.float32(-np.inf))
    x0p = np.where(x0 != 0, np.nextafter(x0, away), x0).astype(np.float32)
    forms = {"f64": dict(precision="f64", graph=False), "f64_ulp": dict(precision="f64", graph=False),
             "plain": dict(precision="f32", graph=False),
             "mixed_bf16": dict(precision="mixed", sino_dtype="bf16"),
             "mixed_f16": dict(precision="mixed", sino_dtype="f16")}
    solvers = {}
    for name, kw in forms.items():
        s = PwlsOS(geom, y, w, numos, beta, delta, **kw)
        s.setup(x0p if name == "f64_ulp" else x0)
        s.prepare(n_iter)
        solvers[name] = s
    curves = {k: [] for k in forms if k != "f64"}
    rmsd64 = []
    phi = {}
    for it in range(n_iter):
        imgs = {}
        for name, s in solvers.items():
            out = s.iterate()
            imgs[name] = s.image()
            phi[name] = out["phi"]
            if name == "f64":
                rmsd64.append(float(f"{out['rmsd']:.3e}"))
        for k in curves:
            curves[k].append(float(f"{rel_rmse(imgs[k], imgs['f64']):.3e}"))
        log(f"accuracy iteration {it + 1}:", {k: v[-1] for k, v in curves.items()})
    del solvers
    # iterations where two FP64 runs one FP32 ulp apart still agree to 1e-4: there the FP64
    # iterate is a well-posed target
    stable = [i + 1 for i, e in enumerate(curves["f64_ulp"]) if e <= 1e-4]
    worst_stable = {k: max((curves[k][i - 1] for i in stable), default=None) for k in curves if k != "f64_ulp"}
    return {"iterations": n_iter,
            "rel_rmse_vs_fp64_final": {k: v[-1] for k, v in curves.items()},
            "rel_rmse_vs_fp64_per_iter": curves,
            "fp64_rmsd_per_iter": rmsd64,
            "fp64_stable_iterations": stable,
            "worst_rel_rmse_while_fp64_stable": worst_stable,
            "phi_final": phi,
            "note": "f64_ulp is the all-FP64 loop started one FP32 ulp away from x0: its distance from the "
                    "FP64 iterate is what FP32 rounding of the start alone produces in the restart-free "
                    "Nesterov loop of reconPwlsSeNesterovSqs.m (nrestart = 2 itnlim); the forms are compared "
                    "with 1e-3 over the iterations where f64_ulp stays within 1e-4"}


def accuracy_m1(geom, y, w, x0, n_iter, beta=1e4, delta=1e-2):
    """One subset (M = 1): the same loop without ordered subsets, where the iterates do not
    oscillate.  Relative RMSE of each form against the FP64 iterate after every iteration
    (the mixed forms without graph replay: one subset is an odd count for the x/z double
    buffers; the storage precisions are those of the timed form)."""
    from ctsim_amd.pwls import PwlsOS
    forms = {"f64": dict(precision="f64", graph=False), "plain": dict(precision="f32", graph=False),
             "mixed_bf16": dict(precision="mixed", sino_dtype="bf16", graph=False),
             "mixed_f16": dict(precision="mixed", sino_dtype="f16", graph=False)}
    solvers = {}
    for name, kw in forms.items():
        sv = PwlsOS(geom, y, w, 1, beta, delta, **kw)
        sv.setup(x0)
        sv.prepare(n_iter)
        solvers[name] = sv
    curves = {k: [] for k in forms if k != "f64"}
    for it in range(n_iter):
        imgs = {}
        for name, sv in solvers.items():
            sv.iterate()
            imgs[name] = sv.image()
        for k in curves:
            curves[k].append(float(f"{rel_rmse(imgs[k], imgs['f64']):.3e}"))
    log("accuracy M=1 final:", {k: v[-1] for k, v in curves.items()})
    del solvers
    return {"subsets": 1, "iterations": n_iter, "rel_rmse_vs_fp64_per_iter": curves,
            "rel_rmse_vs_fp64_final": {k: v[-1] for k, v in curves.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sino-dtype", default="bf16", choices=["bf16", "f16"])
    ap.add_argument("--accuracy-iters", type=int, default=20)
    ap.add_argument("--no-pair", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fdk", action="store_true")
    ap.add_argument("--no-2d", action="store_true")
    ap.add_argument("--fuse", action="store_true",
                    help="run the backprojection and the Nesterov-SQS update as one kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from ctsim_amd import workloads as W
    from ctsim_amd.device import device_count
    if device_count() < 1:
        raise SystemExit("no GPU visible")
    geom = W.pwls_512x128()
    log("simulating PWLS data ...")
    vol, y, w = simulate_pwls_data(geom)
    # FDK initial image (reconPwlsSeNesterovSqs initialises from FBP)
    from ctsim_amd.fdk import recon_fdk
    x0 = recon_fdk(y, geom, "hamming")
    x0 = np.maximum(x0, 0).astype(np.float32)
    log("FDK init done")

    counts = load_counts()
    c_it = counts["pwls_512x128/tf"]
    f_iter = 2 * flops(c_it)  # forward + back over all views per iteration

    n_acc = args.accuracy_iters
    solver, dt, outs = run_pwls("mixed", geom, y, w, x0, args.steps, args.warmup, args.sino_dtype,
                                   fused=args.fuse)
    ms_mixed = dt / args.steps * 1e3
    log(f"mixed: {ms_mixed:.1f} ms/iter, phi {outs[-1]['phi']:.6e}, rmsd {outs[-1]['rmsd']:.3e}")
    log("autotuned schedules:", solver.tuning)
    tuning = solver.tuning
    sched = solver.P.schedule()
    kt = kernel_times(solver)
    log("kernel ms (avg per subset):", {k: round(v, 3) for k, v in kt.items()})
    del solver

    plain = None
    if not args.no_plain:
        sp, dtp, outs_p = run_pwls("plain", geom, y, w, x0, args.steps, args.warmup)
        plain = dtp / args.steps * 1e3
        log(f"plain: {plain:.1f} ms/iter, phi {outs_p[-1]['phi']:.6e}")
        kt_plain = kernel_times(sp)
        del sp
    accuracy = accuracy_lockstep(geom, y, w, x0, n_acc) if n_acc else None
    if accuracy is not None:
        accuracy["m1"] = accuracy_m1(geom, y, w, x0, n_acc)
    M = 24
    # dominant kernel: the forward projector (per-subset launch)
    f_launch = flops(c_it) / M
    achieved = f_launch / (kt["proj"] * 1e-3) / 1e12
    ring = (sched[0] & 3) != 2
    split = (sched[0] >> 2) & 3
    kname = ("cbct_proj_ring_kernel<float,SF,16,8,2,8>" if ring and split != 1 else
             "cbct_proj_ring_kernel<float,SF,16,1,2,4>" if ring else "cbct_proj_kernel<float,float,float,SF,16,1,*>")
    roof = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                frac=round(achieved / PEAK_FP32_TFLOPS, 4), traffic=measured_traffic(),
                kernel=f"{kname} (one 41-view subset, FP32 input z - x0, autotuned schedule {sched[0]})",
                note="FP32 vector ALU bound, no MFMA (not a dense contraction); peak = MI355X's published FP32 "
                     "figure; achieved = (17U+3C+60K)/24 per launch / HIP-event time on the launch stream")
    res = {
        "metric": "PWLS-OS iter/s at 512²×128 and DD proj+backproj GUPS, mixed vs plain FP32",
        "value": round(1e3 / ms_mixed, 4),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (water cylinder + 10 random ellipsoid inserts, Poisson I0=2e4, FDK init)",
        "config": {"workload": "pwls_os_sf_512x512x128_984views_1024x64det_24subsets", "volume": [512, 512, 128],
                   "views": 984, "detector": [1024, 64], "subsets": 24, "beta": 1e4, "delta_per_mm": 1e-3,
                   "storage": f"{args.sino_dtype} data (y - A x0), weights and residual; fp32 volumes, projector "
                              "input z - x0 and footprint sums; fp64 denominators and reductions",
                   "graph": True},
        "plain_fp32_iter_per_s": round(1e3 / plain, 4) if plain else None,
        "mixed_over_plain": round(plain / ms_mixed, 3) if plain else None,
        "gups_iteration": round(2 * c_it["U"] / (ms_mixed * 1e-3) / 1e9, 2),
        "iteration_tflops": round(f_iter / (ms_mixed * 1e-3) / 1e12, 3),
        "kernel_ms_per_subset": {k: round(v, 4) for k, v in kt.items()},
        "autotune": {"schedule": {"proj": sched[0], "back": sched[1]}, "proj": tuning["proj"],
                     "back": tuning["back"],
                     "codes": "proj form|split<<2 (form 0 auto, 1 ring, 2 register, 3 ring with two row chunks; split 0 auto, 1 whole, "
                              "2 shared); back 0 linear, 1 super-tile; times in ms per 41-view launch"},
        "roofline": roof,
        "accuracy": accuracy,
    }
    if plain:
        res["kernel_ms_per_subset_plain"] = {k: round(v, 4) for k, v in kt_plain.items()}
    if not args.no_pair:
        res["pair_512x512x256"] = pair_gups()
    if not args.no_fdk:
        res["fdk_512"] = fdk_bench()
    if not args.no_2d:
        res["fan_256"] = fan_pair_bench()
        res["batch2_uneven"] = batch2_bench()
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(geom)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
