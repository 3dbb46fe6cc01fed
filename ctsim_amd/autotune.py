"""Schedule cache of the 3D autotuner (ctsim_plan_autotune).

A tuned schedule depends on the operator, the geometry, the view selection, the storage
dtypes and the GPU, not on the data, so it is measured once and kept in a JSON file
(``$CTSIM_SCHEDULE_CACHE``, default ``~/.cache/ctsim_amd/schedules.json``):

    from ctsim_amd.autotune import tuned
    tuned(P, d_vol, d_sino, "proj,tf", view0, vstep, nsub)   # sets P's schedule (cached or measured)

The reference has no counterpart: its CUDA path launches fixed 16 x 16 blocks with one
block per detector column (geom/dd_cuda/cbct_dd_proj.c:441-442).
"""
from __future__ import annotations

import hashlib
import json
import os

from .projector import Projector
from .device import DeviceArray


def cache_path() -> str:
    return os.environ.get("CTSIM_SCHEDULE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "ctsim_amd",
                                                                  "schedules.json")


def _device_name() -> str:
    try:
        with open("/sys/class/drm/card0/device/product_name") as f:
            return f.read().strip()
    except OSError:
        return "gpu"


def schedule_key(P: Projector, op: str, view0: int, vstep: int, nsub: int, in_dt: int, out_dt: int) -> str:
    g = P.geom
    fields = [op, g.reconSize, g.reconSpacing, g.reconOffset, g.detSize, g.detSpacing, g.detOffset, g.SAD, g.SDD,
              g.flatPanel, view0, vstep, nsub, in_dt, out_dt, P.compute, _device_name(),
              hashlib.sha1(g.betas.tobytes()).hexdigest()[:12]]
    return hashlib.sha1(repr(fields).encode()).hexdigest()[:20]


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tuned(P: Projector, d_in: DeviceArray, d_out: DeviceArray, op: str, view0=0, vstep=1, nsub=None, reps=3,
          stream=None, refresh=False) -> dict:
    """Set P's schedule for op from the cache, or measure it with the autotuner and store it.
    Returns the cache entry ({"best": code, "candidates": [...]})."""
    nsub = P.geom.noViews if nsub is None else nsub
    key = schedule_key(P, op, view0, vstep, nsub, d_in.dtype, d_out.dtype)
    path = cache_path()
    cache = _load(path)
    entry = None if refresh else cache.get(key)
    if entry is None:
        entry = P.autotune(d_in, d_out, op, view0, vstep, nsub, reps=reps, stream=stream)
        entry["candidates"] = [list(c) for c in entry["candidates"]]  # as JSON stores them
        entry["op"] = op
        cache[key] = entry
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    if op.startswith("proj"):
        P.set_schedule(proj_sched=entry["best"])
    else:
        P.set_schedule(back_sched=entry["best"])
    return entry
