"""The 3D launch schedules and the autotuner (ctsim_plan_set_schedule / ctsim_plan_autotune).

Every forward schedule (automatic, LDS ring with whole or shared work items, register
pipeline with whole or shared items, the ring with two row chunks per item) and both
backprojector column orders against the oracle, on geometries on either side of the ring
forms' slice limits (128, 192 and 256 slices, FP32 and FP16 volumes) and detectors of one
to three 128-row items, with the same tolerance as the other parity tests; the
autotuner on a subset launch returns one of its candidates, keeps it in the plan, and the
schedule cache stores and reuses it."""
import os

import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_parity import _cone_case, _tol, geometry_from_g3

pytestmark = pytest.mark.gpu

PROJ_SCHEDS = [0, 1 | 1 << 2, 1 | 2 << 2, 2 | 1 << 2, 2 | 2 << 2, 3 | 1 << 2]
CASES = [
    dict(seed=21, n=(40, 36, 100), nuv=(80, 40), nviews=12),               # ring, two slots
    dict(seed=22, n=(36, 40, 180), nuv=(70, 70), nviews=9, dz=0.3, dv=0.9),  # ring, three slots
    dict(seed=23, n=(30, 34, 240), nuv=(60, 90), nviews=8, dz=0.3, dv=0.9),  # ring, four slots (on request)
    dict(seed=24, n=(30, 28, 120), nuv=(70, 300), nviews=6, dv=0.25),        # 300 rows: three two-chunk items
    dict(seed=25, n=(26, 30, 250), nuv=(50, 200), nviews=5, dz=0.25, dv=0.3, dettype=0, couch=True),
]


@pytest.mark.parametrize("case", CASES, ids=[str(c["seed"]) for c in CASES])
@pytest.mark.parametrize("vdt", ["f32", "f16"])
def test_every_schedule_matches_the_oracle(case, vdt):
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(**case)
    v = vol.astype(np.float16).astype(np.float32) if vdt == "f16" else vol
    P = Projector(geometry_from_g3(g, mp))
    d_vol = P.upload_volume(v, dtype=vdt)
    out = DeviceArray(P.n_sino(), "f32")
    print(f"schedules case {case['seed']} n={g.n.tolist()} {vdt}", flush=True)
    for kind in ("dd", "tf"):
        want = O.cbct(f"proj,{kind}", g, v, mp)
        tol = _tol(f"proj,{kind}", g, v, mp)
        for sc in PROJ_SCHEDS:
            P.set_schedule(proj_sched=sc)
            P.proj(d_vol, out, f"proj,{kind}")
            e = rel_l2(P.download_sino(out), want)
            assert e <= tol, (kind, sc, e)
    if vdt == "f32":
        d_sino = P.upload_sino(sino)
        img = DeviceArray(P.n_img, "f32")
        for kind in ("dd", "tf"):
            want = O.cbct(f"back,{kind}", g, sino, mp)
            for sb in (0, 1):
                P.set_schedule(back_sched=sb)
                P.back(d_sino, img, f"back,{kind}")
                e = rel_l2(P.download_volume(img), want)
                assert e <= _tol(f"back,{kind}", g, sino, mp), (kind, sb, e)


def test_automatic_forward_takes_the_two_chunk_form_at_scale():
    """a detector over 64 rows with enough unsplit items (here 342 views x 16 tiles x 3 row
    chunks = 16416) makes the automatic schedule take the two-chunk ring form: its result is
    form 3's, bit for bit, and the oracle's to the parity tolerance"""
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(seed=26, n=(16, 18, 40), nuv=(256, 130), nviews=342, dz=0.6, dv=0.35)
    P = Projector(geometry_from_g3(g, mp))
    d_vol = P.upload_volume(vol)
    out = DeviceArray(P.n_sino(), "f32")
    print("automatic two-chunk form", g.n.tolist(), flush=True)
    for kind in ("dd", "tf"):
        P.set_schedule(proj_sched=0)
        P.proj(d_vol, out, f"proj,{kind}")
        auto = P.download_sino(out)
        P.set_schedule(proj_sched=3 | 1 << 2)
        P.proj(d_vol, out, f"proj,{kind}")
        np.testing.assert_array_equal(auto, P.download_sino(out))
        want = O.cbct(f"proj,{kind}", g, vol, mp)
        assert rel_l2(auto, want) <= _tol(f"proj,{kind}", g, vol, mp), kind


FDK_CASES = [
    dict(seed=31),
    dict(seed=32, dettype=0, couch=True, offsets=True),
    dict(seed=33, n=(37, 53, 70), nuv=(70, 130), nviews=9),                  # nz not a multiple of 64
    dict(seed=34, n=(24, 24, 300), nuv=(50, 40), nviews=5, dz=0.2, dv=1.5),  # voxels above and below the rows
    dict(seed=35, n=(20, 22, 16), nuv=(40, 1100), nviews=4, dz=2.0, dv=0.05),  # taller than the staged rows
]


@pytest.mark.parametrize("case", FDK_CASES, ids=[str(c["seed"]) for c in FDK_CASES])
def test_fdk_staged_and_direct_forms_match_the_oracle(case):
    """Both FDK backprojection forms (back_sched bit 1) against the oracle, and against each
    other on an FP16 sinogram: the staged form interpolates in u first, so the two agree to
    FP32 rounding."""
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(**case)
    P = Projector(geometry_from_g3(g, mp))
    d_sino = P.upload_sino(sino)
    d16 = P.upload_sino(sino, dtype="f16")
    img = DeviceArray(P.n_img, "f32")
    print(f"fdk forms case {case['seed']} n={g.n.tolist()} nuv={list(case.get('nuv', (96, 40)))}", flush=True)
    for op in ("back,pd", "back,hpd"):
        want = O.cbct(op, g, sino, mp)
        tol = _tol(op, g, sino, mp)
        got16 = {}
        for sb in (1, 3):
            P.set_schedule(back_sched=sb)
            P.back(d_sino, img, op)
            e = rel_l2(P.download_volume(img), want)
            assert e <= tol, (op, sb, e)
            P.back(d16, img, op)
            got16[sb] = P.download_volume(img)
        assert rel_l2(got16[1], got16[3]) < 1e-6, op
    P.set_schedule(back_sched=1)
    r = P.autotune(d_sino, img, "back,hpd", reps=1)
    assert sorted(c for c, _ in r["candidates"]) == [1, 3] and r["best"] in (1, 3)
    assert P.schedule()[1] == r["best"]


def test_autotuner_picks_and_caches(tmp_path, monkeypatch):
    from ctsim_amd import autotune
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    monkeypatch.setenv("CTSIM_SCHEDULE_CACHE", str(tmp_path / "schedules.json"))
    g, vol, sino, mp = _cone_case(**CASES[0])
    P = Projector(geometry_from_g3(g, mp))
    d_vol = P.upload_volume(vol)
    d_sino = DeviceArray(P.n_sino(6), "f32")
    img = DeviceArray(P.n_img, "f32")
    e = autotune.tuned(P, d_vol, d_sino, "proj,tf", 1, 2, 6, reps=2)
    print("forward candidates", e["candidates"], "best", e["best"], flush=True)
    assert e["best"] in PROJ_SCHEDS and P.schedule()[0] == e["best"]
    assert [c for c, _ in e["candidates"]] == PROJ_SCHEDS and all(ms > 0 for _, ms in e["candidates"])
    b = autotune.tuned(P, d_sino, img, "back,tf", 1, 2, 6, reps=2)
    assert b["best"] in (0, 1) and P.schedule() == (e["best"], b["best"])
    assert os.path.exists(tmp_path / "schedules.json")
    # a second plan of the same geometry takes the cached schedules without timing
    P2 = Projector(geometry_from_g3(g, mp))
    P2.set_schedule(proj_sched=0, back_sched=0)
    e2 = autotune.tuned(P2, d_vol, d_sino, "proj,tf", 1, 2, 6)
    assert e2 == e and P2.schedule()[0] == e["best"]
