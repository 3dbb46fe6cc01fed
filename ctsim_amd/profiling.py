"""Short, fixed GPU workloads for profiler runs (rocprofv3 kernel traces and counter
passes), so that a profile names only the kernels it is about.

    python -m ctsim_amd.profiling copy [reps]          # 256 MiB FP32 copy kernel (ctsim_convert)
    python -m ctsim_amd.profiling pwls-subset [reps]   # proj,tf + back,tf of one 41-view subset,
                                                       # 512x512x128 volume, 1024x64 detector: FP32
                                                       # projector input and output, BF16 sinogram
                                                       # into the backprojector (as in PwlsOS mixed)
    python -m ctsim_amd.profiling pair-slab [reps]     # proj,tf + back,tf of 24 of 720 views,
                                                       # 512x512x256 volume, 1024x384 detector, FP32
    python -m ctsim_amd.profiling pair-run [reps]      # the same pair over 48 consecutive views (the
                                                       # view-to-view locality of the full 720-view pair)
    python -m ctsim_amd.profiling fdk-back [reps]      # back,hpd of 720 filtered views (FP16) into a
                                                       # 512^3 FP32 volume, 1024x768 detector
    python -m ctsim_amd.profiling pwls-update [reps]   # fused Nesterov-SQS update (Huber, FP16 z copy)
                                                       # on 512x512x128 FP32 volumes
    python -m ctsim_amd.profiling fan-pair [reps]      # proj,dd + back,dd of the 2D fan workload:
                                                       # 256x256 image, 360 views, 512 bins, FP32
    python -m ctsim_amd.profiling batch2 [reps]        # proj,dd + back,dd of the 64 uneven 2D slices
                                                       # (FP16 images and sinograms), schedule (1, 1)

The copy workload calibrates the FETCH_SIZE / WRITE_SIZE counters: it reads and writes
exactly 256 MiB per launch."""
from __future__ import annotations

import sys

from . import workloads
from ._lib import F32, call
from .device import DeviceArray, synchronize
from .projector import Projector

COPY_ELEMS = 64 * 1024 * 1024  # 256 MiB of float32


def copy(reps: int) -> None:
    a = DeviceArray(COPY_ELEMS, "f32").fill(1.0)
    b = DeviceArray(COPY_ELEMS, "f32")
    for _ in range(reps):
        call("ctsim_convert", a.vp, F32, b.vp, F32, COPY_ELEMS, None)
    synchronize()


def _pair(g, vol, nsub, vstep, dtype, reps, back_dtype=None):
    """proj,tf of the volume stored in dtype into a dtype sinogram, then back,tf of a
    back_dtype copy of it (default: the same buffer) into an FP32 volume"""
    P = Projector(g)
    x = P.upload_volume(vol, dtype=dtype)
    s = DeviceArray(P.n_sino(nsub), dtype)
    sb = s if back_dtype is None else DeviceArray(P.n_sino(nsub), back_dtype)
    b = DeviceArray(P.n_img, "f32")
    for _ in range(reps):
        P.proj(x, s, op="proj,tf", view0=0, vstep=vstep, nsub=nsub)
        if sb is not s:
            sb.copy_from(s)
        P.back(sb, b, op="back,tf", view0=0, vstep=vstep, nsub=nsub)
    synchronize()


def pwls_subset(reps: int) -> None:
    g = workloads.pwls_512x128(with_map=True)
    vol = workloads.body_phantom(g.reconSize, g.reconSpacing, 10, 7)
    _pair(g, vol, 41, 24, "f32", reps, back_dtype="bf16")


def pair_slab(reps: int) -> None:
    g = workloads.sf_pair_512()
    vol = workloads.ellipsoid_phantom(g.reconSize, g.reconSpacing, 20, 2)
    _pair(g, vol, 24, 30, "f32", reps)


def pair_run(reps: int) -> None:
    g = workloads.sf_pair_512()
    vol = workloads.ellipsoid_phantom(g.reconSize, g.reconSpacing, 20, 2)
    _pair(g, vol, 48, 1, "f32", reps)


def fdk_back(reps: int) -> None:
    g = workloads.fdk_512()
    P = Projector(g)
    s = DeviceArray(P.n_sino(g.noViews), "f16").fill(1.0)
    b = DeviceArray(P.n_img, "f32")
    for _ in range(reps):
        P.back(s, b, op="back,hpd", scale=1.0)
    synchronize()


def pwls_update(reps: int) -> None:
    nx, ny, nz = 512, 512, 128
    n = nx * ny * nz
    x = [DeviceArray(n, "f32").fill(0.02), DeviceArray(n, "f32")]
    z = [DeviceArray(n, "f32").fill(0.02), DeviceArray(n, "f32")]
    s = DeviceArray(n, "f32").fill(0.02)
    g = DeviceArray(n, "f32").fill(1e-3)
    pre = DeviceArray(n, "f32").fill(10.0)
    zc = DeviceArray(n, "f16")
    for k in range(reps):
        a, b = k % 2, 1 - k % 2
        call("ctsim_pwls_nesterov_update", x[a].vp, z[a].vp, s.vp, g.vp, pre.vp, x[b].vp, z[b].vp, zc.vp, zc.dtype,
             None, nx, ny, nz, 1e4, 1e-2, 0, 1.0 + k, 2.0 + k, None)
    synchronize()


def fan_pair(reps: int) -> None:
    g = workloads.fan_256()
    P = Projector(g)
    img = DeviceArray.from_numpy(workloads.ellipse_image(256, 0.8, 10, 1).ravel(order="F"), "f32")
    s = DeviceArray(P.n_sino(), "f32")
    b = DeviceArray(P.n_img, "f32")
    for _ in range(reps):
        P.proj(img, s, "proj,dd")
        P.back(s, b, "back,dd")
    synchronize()


def batch2(reps: int) -> None:
    from .batch2 import Batch2
    geoms, imgs = workloads.batch2_uneven()
    B = Batch2(geoms)
    B.set_schedule(1, 1)
    img = DeviceArray.from_numpy(B.pack_images(imgs), "f16")
    s = DeviceArray(B.sino_elems, "f16")
    b = DeviceArray(B.img_elems, "f16")
    for _ in range(reps):
        B.proj(img, s, "proj,dd")
        B.back(s, b, "back,dd")
    synchronize()


WORKLOADS = {"copy": copy, "pwls-subset": pwls_subset, "pair-slab": pair_slab, "pair-run": pair_run, "fdk-back": fdk_back,
             "pwls-update": pwls_update, "fan-pair": fan_pair, "batch2": batch2}


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in WORKLOADS:
        print("usage: python -m ctsim_amd.profiling {%s} [reps]" % "|".join(WORKLOADS), file=sys.stderr)
        return 2
    reps = int(argv[1]) if len(argv) > 1 else 3
    print("profiling workload", argv[0], "reps", reps, flush=True)
    WORKLOADS[argv[0]](reps)
    print("done", argv[0], flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
