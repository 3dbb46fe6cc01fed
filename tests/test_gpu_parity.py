"""GPU parity of the HIP projector family against the oracle and the reference's golden
vectors, through the C ABI.

Tolerances (stated by the target for FP32 footprints): relative L2 <= 1e-5 against the
reference CPU path, and matched-pair adjoint mismatch |<Ax,y> - <x,A^T y>| / (|Ax| |y|)
< 1e-6.  The FP64 kernels are held to 1e-11 against the FP64 oracle.
"""
import os
import sys

import numpy as np
import pytest

from conftest import golden_files, load_golden, rel_l2
from helpers import ellipse_phantom2, g2_from, g3_from, geometry_from_g2, geometry_from_g3

pytestmark = pytest.mark.gpu

OPS2 = ("proj,dd", "proj,tf", "back,dd", "back,tf", "back,pd")
OPS3 = ("proj,dd", "proj,tf", "back,dd", "back,tf", "back,pd", "back,hpd")
TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    from ctsim_amd import device
    if device.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


# ---------------------------------------------------------------------------------
# golden vectors of the reference's compiled kernels, through the MEX gateways
# ---------------------------------------------------------------------------------
@pytest.mark.parametrize("path", golden_files("fan_"), ids=os.path.basename)
def test_fbct_gateway_vs_reference_golden(path):
    from ctsim_amd import mex
    rec = load_golden(path)
    say("golden 2D", os.path.basename(path))
    for op in OPS2:
        x = rec["img"] if op.startswith("proj") else rec["sino"]
        got = mex.fbct_geom_mex(op, rec["nxy"], rec["dxy"], rec["oxy"], rec["nu"], rec["du"], rec["ou"], rec["sad"],
                                rec["add"], rec["sdd"], len(rec["betas"]), rec["betas"], rec["dettype"],
                                np.asfortranarray(x))
        want = rec[op.replace(",", "_")]
        assert got.shape == want.shape
        assert rel_l2(got, want) <= TOL, (op, rel_l2(got, want))


@pytest.mark.parametrize("path", golden_files("cone_"), ids=os.path.basename)
def test_cbct_gateway_vs_reference_golden(path):
    from ctsim_amd import mex
    rec = load_golden(path)
    say("golden 3D", os.path.basename(path))
    mp = rec["map"] if bool(rec["has_map"]) else None
    for op in OPS3:
        x = rec["img"] if op.startswith("proj") else rec["sino"]
        got = mex.cbct_geom_mex(op, rec["n"], rec["d"], rec["o"], rec["nuv"], rec["duv"], rec["ouv"], rec["sad"],
                                rec["add"], rec["sdd"], len(rec["betas"]), rec["betas"], rec["couchz"],
                                rec["dettype"], np.asfortranarray(x), mp)
        want = rec[op.replace(",", "_")]
        assert got.shape == want.shape
        assert rel_l2(got, want) <= TOL, (op, rel_l2(got, want))


def test_cuda_gateway_names():
    from ctsim_amd import mex
    rec = load_golden(golden_files("cone_flat_map")[0])
    args = (rec["n"], rec["d"], rec["o"], rec["nuv"], rec["duv"], rec["ouv"], rec["sad"], rec["add"], rec["sdd"],
            len(rec["betas"]), rec["betas"], rec["couchz"], rec["dettype"])
    mp = rec["map"]
    assert rel_l2(mex.cbct_dd_cuda("proj,dd", *args, rec["img"], mp), rec["proj_dd"]) <= TOL
    assert rel_l2(mex.cbct_sf_cuda("back,tf", *args, rec["sino"], mp), rec["back_tf"]) <= TOL
    assert rel_l2(mex.cbct_back_pd_cuda("back,hpd", *args, rec["sino"], mp), rec["back_hpd"]) <= TOL


def _per_op(name, g, betas, couchz, mp, x, out_shape):
    import ctypes as C
    from ctsim_amd._lib import lib
    x = np.asfortranarray(x, dtype=np.float32)
    out = np.zeros(out_shape, dtype=np.float32, order="F")
    betas = np.ascontiguousarray(betas, dtype=np.float64)
    err = C.create_string_buffer(256)
    f32p = C.POINTER(C.c_float)
    cz = None if couchz is None else np.ascontiguousarray(couchz, dtype=np.float32)
    mpp = None if mp is None else np.ascontiguousarray(np.asarray(mp, dtype=np.uint8).ravel(order="F"))
    rc = getattr(lib(), name)(C.byref(g), len(betas), betas.ctypes.data_as(C.POINTER(C.c_double)),
                              None if cz is None else cz.ctypes.data_as(f32p),
                              None if mpp is None else mpp.ctypes.data_as(C.POINTER(C.c_uint8)),
                              x.ctypes.data_as(f32p), out.ctypes.data_as(f32p), err, 256)
    assert rc == 0, err.value
    return out


def test_per_operator_entry_points_match_golden():
    from ctsim_amd._lib import CtsimGeom
    rec = load_golden(golden_files("cone_flat_map")[0])
    n, nuv = [int(v) for v in rec["n"]], [int(v) for v in rec["nuv"]]
    g = CtsimGeom(*n, *map(float, rec["d"]), *map(float, rec["o"]), *nuv, *map(float, rec["duv"]),
                  *map(float, rec["ouv"]), float(rec["sad"]), float(rec["add"]), float(rec["sdd"]), int(rec["dettype"]))
    sino_shape = (nuv[1], nuv[0], len(rec["betas"]))
    for op in OPS3:
        proj = op.startswith("proj")
        got = _per_op("ctsim_cbct_" + op.replace(",", "_"), g, rec["betas"], rec["couchz"], rec["map"],
                      rec["img"] if proj else rec["sino"], sino_shape if proj else tuple(n))
        assert rel_l2(got, rec[op.replace(",", "_")]) <= TOL, op
    rec = load_golden(golden_files("fan_flat_small")[0])
    n = [int(v) for v in rec["nxy"]]
    g = CtsimGeom(n[0], n[1], 1, *map(float, rec["dxy"]), 1.0, *map(float, rec["oxy"]), 0.0, int(rec["nu"]), 1,
                  float(rec["du"]), 1.0, float(rec["ou"]), 0.0, float(rec["sad"]), float(rec["add"]),
                  float(rec["sdd"]), int(rec["dettype"]))
    for op in OPS2:
        proj = op.startswith("proj")
        got = _per_op("ctsim_fbct_" + op.replace(",", "_"), g, rec["betas"], None, None,
                      rec["img"] if proj else rec["sino"], (int(rec["nu"]), len(rec["betas"])) if proj else tuple(n))
        assert rel_l2(got, rec[op.replace(",", "_")]) <= TOL, op


# ---------------------------------------------------------------------------------
# random seeded cases against the oracle, larger than the fixtures
# ---------------------------------------------------------------------------------
def _cone_case(seed, n=(48, 44, 24), nuv=(96, 40), nviews=20, dettype=1, offsets=False, couch=False, map_=True,
               dz=0.6, dv=0.4):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    betas = np.sort(rng.uniform(0, 2 * np.pi, nviews))
    o = (0.7, -1.2, 0.5) if offsets else (0, 0, 0)
    ouv = (1.5, -2.5) if offsets else (0, 0)
    g = O.G3(n, (0.5, 0.5, dz), o, nuv, (0.4 if dettype == 1 else 0.45, dv), ouv, 1000.0, 0.0, 1500.0, betas,
             couchz=rng.uniform(-3, 3, nviews) if couch else None, dettype=dettype)
    vol = rng.uniform(0, 1, n).astype(np.float32)
    vol[rng.uniform(size=n) < 0.3] = 0
    sino = rng.uniform(0, 1, (nuv[1], nuv[0], nviews)).astype(np.float32)
    mp = (rng.uniform(size=n[:2]) < 0.9) if map_ else None
    return g, vol, sino, mp


CONE_CASES = [
    dict(seed=1),
    dict(seed=2, offsets=True, map_=False),
    dict(seed=3, dettype=0, couch=True),
    dict(seed=4, n=(37, 53, 70), nuv=(70, 130), nviews=9),  # nv > 64: several row chunks, odd sizes
    dict(seed=5, n=(64, 64, 8), nuv=(200, 12), nviews=7),  # thin detector, wide fan
    dict(seed=6, n=(40, 40, 150), nuv=(90, 60), nviews=8, dz=0.3, dv=1.2),  # slices finer than rows: multi-pass
    dict(seed=7, n=(30, 30, 6), nuv=(70, 140), nviews=6, dz=4.0, dv=0.3),  # 20 rows per slice: 16-cell cap
    dict(seed=9, n=(30, 30, 100), nuv=(60, 64), nviews=6, dz=0.5, dv=0.45),  # 61-64 rows per column, nv = 64
    dict(seed=10, n=(30, 30, 40), nuv=(60, 62), nviews=6, dz=1.0, dv=0.45),  # nv = 62: rows reach the pads
    dict(seed=11, n=(24, 24, 300), nuv=(50, 40), nviews=5, dz=0.2, dv=1.5),  # 300 voxels per column: two windows
]


def _tol(op, g, x, mp):
    """1e-5, widened to 3x the oracle's own FP32-vs-FP64 gap where that gap is larger.
    Footprints wider than 16 cells hit the reference's cap, which anchors the cell window
    at floor(clo): there the operator is discontinuous at the last ulp and even the
    reference's FP32 and FP64 results differ by ~1e-3."""
    from oracle import oracle as O
    cond = rel_l2(O.cbct(op, g, x, mp), O.cbct(op, g, x, mp, prec=64))
    return max(TOL, 3.0 * cond)


@pytest.mark.parametrize("case", CONE_CASES, ids=[str(c["seed"]) for c in CONE_CASES])
def test_cbct_plan_vs_oracle(case):
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(**case)
    say("cone case", case)
    P = Projector(geometry_from_g3(g, mp))
    d_vol = P.upload_volume(vol)
    d_sino = P.upload_sino(sino)
    out_s = DeviceArray(P.n_sino(), "f32")
    out_v = DeviceArray(P.n_img, "f32")
    for kind in ("dd", "tf"):
        P.proj(d_vol, out_s, f"proj,{kind}")
        got = P.download_sino(out_s)
        want = O.cbct(f"proj,{kind}", g, vol, mp)
        assert rel_l2(got, want) <= _tol(f"proj,{kind}", g, vol, mp), (kind, rel_l2(got, want))
        # strip pruning must not drop any column: exhaustive candidate lists give the same bits
        ex = DeviceArray(P.n_sino(), "f32")
        P.proj(d_vol, ex, f"proj,{kind}", exhaustive=True)
        assert np.array_equal(P.download_sino(ex), got), kind
    for op in ("back,dd", "back,tf", "back,pd", "back,hpd"):
        P.back(d_sino, out_v, op)
        got = P.download_volume(out_v)
        want = O.cbct(op, g, sino, mp)
        assert rel_l2(got, want) <= _tol(op, g, sino, mp), (op, rel_l2(got, want))


@pytest.mark.parametrize("seed", [1, 4, 9])
def test_cbct_register_pipeline_vs_oracle(seed):
    """FP16 volumes take the register-pipeline forward kernel (FP32 volumes of at most 192
    slices take the LDS ring): against the oracle on the FP16-rounded volume"""
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    case = next(c for c in CONE_CASES if c["seed"] == seed)
    g, vol, sino, mp = _cone_case(**case)
    say("register pipeline case", case)
    v16 = vol.astype(np.float16).astype(np.float32)
    P = Projector(geometry_from_g3(g, mp))
    d_vol = P.upload_volume(v16, dtype="f16")
    out = DeviceArray(P.n_sino(), "f32")
    for kind in ("dd", "tf"):
        P.proj(d_vol, out, f"proj,{kind}")
        got = P.download_sino(out)
        want = O.cbct(f"proj,{kind}", g, v16, mp)
        assert rel_l2(got, want) <= _tol(f"proj,{kind}", g, v16, mp), (kind, rel_l2(got, want))


@pytest.mark.parametrize("seed", range(8))
def test_cbct_random_geometries_vs_oracle(seed):
    """DD and SF, both directions, on random geometries: voxel and detector pitches, offsets,
    couch positions, maps and both detector types; slice counts on either side of the
    forward kernels' 192-slice switch and row counts on either side of the backprojector's
    64-row form"""
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    rng = np.random.default_rng(300 + seed)
    n = (int(rng.integers(12, 28)), int(rng.integers(12, 28)), int(rng.choice([int(rng.integers(4, 190)), 250])))
    nuv = (int(rng.integers(30, 90)), int(rng.choice([int(rng.integers(8, 64)), int(rng.integers(65, 140))])))
    nviews = 5
    dettype = int(seed % 2)
    betas = np.sort(rng.uniform(0, 2 * np.pi, nviews))
    d = tuple(float(v) for v in rng.uniform(0.4, 1.2, 3))
    o = tuple(float(v) for v in rng.uniform(-1.5, 1.5, 3))
    duv = tuple(float(v) for v in rng.uniform(0.3, 1.6, 2))
    ouv = tuple(float(v) for v in rng.uniform(-2, 2, 2))
    g = O.G3(n, d, o, nuv, duv, ouv, 1000.0, 0.0, 1500.0, betas,
             couchz=rng.uniform(-3, 3, nviews) if seed % 3 == 0 else None, dettype=dettype)
    vol = rng.uniform(0, 1, n).astype(np.float32)
    sino = rng.uniform(0, 1, (nuv[1], nuv[0], nviews)).astype(np.float32)
    mp = (rng.uniform(size=n[:2]) < 0.9) if seed % 2 == 0 else None
    say("random cone case", seed, n, nuv, d, duv, dettype)
    P = Projector(geometry_from_g3(g, mp))
    d_vol = P.upload_volume(vol)
    d_sino = P.upload_sino(sino)
    out_s = DeviceArray(P.n_sino(), "f32")
    out_v = DeviceArray(P.n_img, "f32")
    for kind in ("dd", "tf"):
        P.proj(d_vol, out_s, f"proj,{kind}")
        got = P.download_sino(out_s)
        want = O.cbct(f"proj,{kind}", g, vol, mp)
        assert rel_l2(got, want) <= _tol(f"proj,{kind}", g, vol, mp), (kind, rel_l2(got, want))
        P.back(d_sino, out_v, f"back,{kind}")
        got = P.download_volume(out_v)
        want = O.cbct(f"back,{kind}", g, sino, mp)
        assert rel_l2(got, want) <= _tol(f"back,{kind}", g, sino, mp), (kind, rel_l2(got, want))


def test_cbct_unsplit_launch_vs_oracle():
    """>= 16384 work items: each wave walks its whole strip (no LDS combine)"""
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(8, n=(24, 24, 12), nuv=(1024, 200), nviews=64, dz=0.9, dv=0.4)
    P = Projector(geometry_from_g3(g, mp))
    out = DeviceArray(P.n_sino(), "f32")
    d_vol = P.upload_volume(vol)
    for kind in ("dd", "tf"):
        P.proj(d_vol, out, f"proj,{kind}")
        assert rel_l2(P.download_sino(out), O.cbct(f"proj,{kind}", g, vol, mp)) <= TOL, kind


FAN_CASES = [
    dict(seed=11, n=(64, 64), nu=128, nviews=36, dettype=1),
    dict(seed=12, n=(45, 77), nu=100, nviews=25, dettype=1, offsets=True),
    dict(seed=13, n=(50, 50), nu=90, nviews=30, dettype=0),
    dict(seed=14, n=(24, 20), nu=600, nviews=1000, dettype=1),  # 10000 (view, tile) items: one wave per item
]


def _fan_case(seed, n, nu, nviews, dettype, offsets=False):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    betas = np.sort(rng.uniform(0, 2 * np.pi, nviews))
    g = O.G2(n[0], n[1], 0.8, 0.7 if offsets else 0.8, 1.2 if offsets else 0.0, -0.7 if offsets else 0.0, nu, 1.0,
             2.0 if offsets else 0.0, 1000.0, 0.0, 1500.0, betas, dettype)
    img = rng.uniform(0, 1, n).astype(np.float32)
    sino = rng.uniform(0, 1, (nu, nviews)).astype(np.float32)
    return g, img, sino


@pytest.mark.parametrize("case", FAN_CASES, ids=[str(c["seed"]) for c in FAN_CASES])
def test_fbct_plan_vs_oracle(case):
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, img, sino = _fan_case(**case)
    say("fan case", case)
    P = Projector(geometry_from_g2(g))
    d_img = P.upload_volume(img)
    d_sino = P.upload_sino(sino)
    out_s = DeviceArray(P.n_sino(), "f32")
    out_i = DeviceArray(P.n_img, "f32")
    for kind in ("dd", "tf"):
        P.proj(d_img, out_s, f"proj,{kind}")
        got = P.download_sino(out_s)
        assert rel_l2(got, O.fbct(f"proj,{kind}", g, img)) <= TOL, kind
        ex = DeviceArray(P.n_sino(), "f32")
        P.proj(d_img, ex, f"proj,{kind}", exhaustive=True)
        assert np.array_equal(P.download_sino(ex), got), kind
    for op in ("back,dd", "back,tf", "back,pd"):
        P.back(d_sino, out_i, op)
        got = P.download_volume(out_i)
        assert rel_l2(got, O.fbct(op, g, sino)) <= TOL, op


# ---------------------------------------------------------------------------------
# adjointness of the GPU pairs, ordered subsets, fp64 kernels, storage dtypes
# ---------------------------------------------------------------------------------
def _adjoint_mismatch(ax, y, x, aty):
    ax, y, x, aty = (np.asarray(a, np.float64).ravel() for a in (ax, y, x, aty))
    return abs(ax.dot(y) - x.dot(aty)) / (np.linalg.norm(ax) * np.linalg.norm(y))


@pytest.mark.parametrize("kind", ["dd", "tf"])
def test_gpu_pairs_are_adjoint(kind):
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(21, n=(40, 40, 32), nuv=(90, 48), nviews=24)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(g.img_shape()).astype(np.float32)
    y = rng.standard_normal(g.sino_shape()).astype(np.float32)
    P = Projector(geometry_from_g3(g, mp))
    ax = DeviceArray(P.n_sino(), "f32")
    aty = DeviceArray(P.n_img, "f32")
    P.proj(P.upload_volume(x), ax, f"proj,{kind}")
    P.back(P.upload_sino(y), aty, f"back,{kind}")
    m = _adjoint_mismatch(P.download_sino(ax), y, x, P.download_volume(aty))
    say("adjoint 3D", kind, m)
    assert m < 1e-6
    g2, img, s2 = _fan_case(31, (64, 64), 128, 40, 1)
    x2 = rng.standard_normal(g2.img_shape()).astype(np.float32)
    y2 = rng.standard_normal(g2.sino_shape()).astype(np.float32)
    P2 = Projector(geometry_from_g2(g2))
    a2 = DeviceArray(P2.n_sino(), "f32")
    t2 = DeviceArray(P2.n_img, "f32")
    P2.proj(P2.upload_volume(x2), a2, f"proj,{kind}")
    P2.back(P2.upload_sino(y2), t2, f"back,{kind}")
    m2 = _adjoint_mismatch(P2.download_sino(a2), y2, x2, P2.download_volume(t2))
    say("adjoint 2D", kind, m2)
    assert m2 < 1e-6


def test_ordered_subsets_and_scale():
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    from ctsim_amd.shell import subset_views
    g, vol, sino, mp = _cone_case(41, nviews=24)
    P = Projector(geometry_from_g3(g, mp))
    d_vol = P.upload_volume(vol)
    M = 5
    for isub in (1, 3, 5):
        v0, vs, ns = Projector.subset(g.nviews, M, isub)
        idx = subset_views(g.nviews, M, isub)
        gs = O.G3(g.n, g.d, g.o, g.nuv, g.duv, g.ouv, g.sad, g.add, g.sdd, g.betas[idx], couchz=g.couchz[idx],
                  dettype=g.dettype)
        out = DeviceArray(P.n_sino(ns), "f32")
        P.proj(d_vol, out, "proj,tf", v0, vs, ns)
        assert rel_l2(P.download_sino(out, ns), O.cbct("proj,tf", gs, vol, mp)) <= TOL
        sub = np.asfortranarray(sino[:, :, idx])
        img = DeviceArray(P.n_img, "f32")
        P.back(P.upload_sino(sub), img, "back,tf", v0, vs, ns, scale=M)
        want = np.float32(M) * O.cbct("back,tf", gs, sub, mp)
        assert rel_l2(P.download_volume(img), want) <= TOL


def test_fp64_kernels_vs_fp64_oracle():
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(51, n=(32, 30, 20), nuv=(70, 36), nviews=12)
    P = Projector(geometry_from_g3(g, mp), compute="f64")
    d_vol = P.upload_volume(vol.astype(np.float64), dtype="f64")
    d_sino = P.upload_sino(sino.astype(np.float64), dtype="f64")
    for kind in ("dd", "tf"):
        out = DeviceArray(P.n_sino(), "f64")
        P.proj(d_vol, out, f"proj,{kind}")
        assert rel_l2(P.download_sino(out), O.cbct(f"proj,{kind}", g, vol, mp, prec=64)) < 1e-11
    for op in ("back,dd", "back,tf", "back,hpd"):
        out = DeviceArray(P.n_img, "f64")
        P.back(d_sino, out, op)
        assert rel_l2(P.download_volume(out), O.cbct(op, g, sino, mp, prec=64)) < 1e-11, op
    g2, img, s2 = _fan_case(52, (40, 40), 80, 16, 1)
    P2 = Projector(geometry_from_g2(g2), compute="f64")
    out = DeviceArray(P2.n_sino(), "f64")
    P2.proj(P2.upload_volume(img.astype(np.float64), "f64"), out, "proj,tf")
    assert rel_l2(P2.download_sino(out), O.fbct("proj,tf", g2, img, prec=64)) < 1e-11


@pytest.mark.parametrize("dt,tol", [("f16", 2e-3), ("bf16", 1.2e-2)])
def test_half_storage_close_to_fp32(dt, tol):
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(61)
    P = Projector(geometry_from_g3(g, mp))
    ref = DeviceArray(P.n_sino(), "f32")
    P.proj(P.upload_volume(vol), ref, "proj,tf")
    half = DeviceArray(P.n_sino(), dt)
    P.proj(P.upload_volume(vol, dtype=dt), half, "proj,tf")
    assert rel_l2(half.to_numpy(), ref.to_numpy()) < tol
    v32 = DeviceArray(P.n_img, "f32")
    P.back(P.upload_sino(sino), v32, "back,tf")
    v16 = DeviceArray(P.n_img, "f32")
    P.back(P.upload_sino(sino, dtype=dt), v16, "back,tf")
    assert rel_l2(v16.to_numpy(), v32.to_numpy()) < tol


def test_batched_uneven_2d_matches_single_slices():
    import ctypes as C
    from ctsim_amd import _lib
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    rng = np.random.default_rng(71)
    geoms, imgs, betas_list = [], [], []
    for k in range(6):
        n = int(rng.choice([32, 48, 64]))
        nv = int(rng.choice([10, 20, 30]))
        nu = int(rng.choice([64, 96]))
        b = np.sort(rng.uniform(0, 2 * np.pi, nv))
        from oracle import oracle as O
        g = O.G2(n, n, 0.8, 0.8, 0, 0, nu, 1.0, 0, 1000.0, 0, 1500.0, b, 1)
        geoms.append(g)
        imgs.append(ellipse_phantom2(n, n, 0.8, 4, rng))
        betas_list.append(b)
    cg = (_lib.CtsimGeom * len(geoms))(*[geometry_from_g2(g).to_c() for g in geoms])
    nvs = np.array([g.nviews for g in geoms], np.int32)
    bptrs = (C.POINTER(C.c_double) * len(geoms))(*[b.ctypes.data_as(C.POINTER(C.c_double)) for b in betas_list])
    h = C.c_void_p()
    _lib.call("ctsim_batch2_create", C.byref(h), len(geoms), cg, nvs.ctypes.data_as(C.POINTER(C.c_int32)), bptrs, 0)
    ni, ns = C.c_int64(), C.c_int64()
    _lib.call("ctsim_batch2_sizes", h, C.byref(ni), C.byref(ns))
    packed = np.concatenate([im.ravel(order="F") for im in imgs])
    assert packed.size == ni.value
    d_img = DeviceArray.from_numpy(packed, "f32")
    d_sino = DeviceArray(ns.value, "f32")
    # the batch's default schedule splits the forward items over 4 waves below 8192 items,
    # as the single-slice launches do: the same sums in the same order
    for kind in ("dd", "tf"):
        _lib.call("ctsim_batch2_proj", h, f"proj,{kind}".encode(), d_img.vp, 0, d_sino.vp, 0, None)
        got = d_sino.to_numpy()
        off = 0
        for g, im in zip(geoms, imgs):
            P = Projector(geometry_from_g2(g))
            one = DeviceArray(P.n_sino(), "f32")
            P.proj(P.upload_volume(im), one, f"proj,{kind}")
            np.testing.assert_array_equal(got[off:off + one.count], one.to_numpy())
            off += one.count
        d_back = DeviceArray(ni.value, "f32")
        _lib.call("ctsim_batch2_back", h, f"back,{kind}".encode(), d_sino.vp, 0, d_back.vp, 0, None)
        gb = d_back.to_numpy()
        off_i, off_s = 0, 0
        for g in geoms:
            P = Projector(geometry_from_g2(g))
            n_s = P.n_sino()
            s = DeviceArray.from_numpy(got[off_s:off_s + n_s], "f32")
            one = DeviceArray(P.n_img, "f32")
            P.back(s, one, f"back,{kind}")
            np.testing.assert_array_equal(gb[off_i:off_i + P.n_img], one.to_numpy())
            off_i += P.n_img
            off_s += n_s
    _lib.call("ctsim_batch2_destroy", h)


def test_batched_schedules_and_autotune_vs_oracle():
    """every launch schedule of the batched 2D operators matches the FP32 oracle per slice
    (the view or strip shares change only the summation order), FP16 storage stays within
    its precision, and the autotuner leaves a valid schedule"""
    from oracle import oracle as O
    from ctsim_amd.batch2 import Batch2
    from ctsim_amd.device import DeviceArray
    rng = np.random.default_rng(72)
    gs, imgs = [], []
    for k in range(5):
        n = int(rng.choice([24, 40, 72]))
        nv = int(rng.choice([7, 16, 33]))
        nu = int(rng.choice([50, 96, 130]))
        b = np.sort(rng.uniform(0, 2 * np.pi, nv))
        gs.append(O.G2(n, n, 0.8, 0.8, 0.3, -0.2, nu, 1.0, 0.4, 1000.0, 0, 1500.0, b, 1))
        imgs.append(ellipse_phantom2(n, n, 0.8, 4, rng))
    B = Batch2([geometry_from_g2(g) for g in gs])
    d_img = DeviceArray.from_numpy(B.pack_images(imgs), "f32")
    d_sino = DeviceArray(B.sino_elems, "f32")
    d_back = DeviceArray(B.img_elems, "f32")
    for kind in ("dd", "tf"):
        want_p = [O.fbct(f"proj,{kind}", g, im) for g, im in zip(gs, imgs)]
        for ps in (1, 2, 4):
            for bs in (1, 2, 4):
                B.set_schedule(ps, bs)
                assert B.schedule == (ps, bs)
                B.proj(d_img, d_sino, f"proj,{kind}")
                sinos = B.unpack_sinos(d_sino.to_numpy())
                for s, w in zip(sinos, want_p):
                    assert rel_l2(s, w) <= TOL, (kind, ps, bs)
                B.back(d_sino, d_back, f"back,{kind}")
                backs = B.unpack_images(d_back.to_numpy())
                for g, s, bk in zip(gs, sinos, backs):
                    assert rel_l2(bk, O.fbct(f"back,{kind}", g, s)) <= TOL, (kind, ps, bs)
    tuned = B.autotune(d_img, d_sino, d_back, "proj,dd", "back,dd", reps=1)
    assert tuned["proj_split"] in (1, 2, 4) and tuned["back_split"] in (1, 2, 4)
    assert tuned["proj_ms"] > 0 and tuned["back_ms"] > 0
    # FP16 storage of images and sinograms, FP32 accumulation
    d_img16 = DeviceArray.from_numpy(B.pack_images(imgs), "f16")
    d_sino16 = DeviceArray(B.sino_elems, "f16")
    B.proj(d_img16, d_sino16, "proj,dd")
    for s, g, im in zip(B.unpack_sinos(d_sino16.to_numpy()), gs, imgs):
        assert rel_l2(s, O.fbct("proj,dd", g, im)) <= 2e-3


# ---------------------------------------------------------------------------------
# ray-driven projection ("proj,rd"): the reference's float march, one for one
# ---------------------------------------------------------------------------------
RD_TOL = 1e-6  # the march reproduces the reference's float operations; only the device's
               # libm-free arithmetic order could differ


@pytest.mark.parametrize("path", golden_files("fan_") + golden_files("cone_"), ids=os.path.basename)
def test_proj_rd_gateways_vs_reference_golden(path):
    from ctsim_amd import mex
    rec = load_golden(path)
    say("golden rd", os.path.basename(path))
    if "nxy" in rec:
        got = mex.fbct_geom_mex("proj,rd", rec["nxy"], rec["dxy"], rec["oxy"], rec["nu"], rec["du"], rec["ou"],
                                rec["sad"], rec["add"], rec["sdd"], len(rec["betas"]), rec["betas"], rec["dettype"],
                                np.asfortranarray(rec["img"]))
    else:
        got = mex.cbct_geom_mex("proj,rd", rec["n"], rec["d"], rec["o"], rec["nuv"], rec["duv"], rec["ouv"],
                                rec["sad"], rec["add"], rec["sdd"], len(rec["betas"]), rec["betas"], rec["couchz"],
                                rec["dettype"], np.asfortranarray(rec["img"]))
    want = rec["proj_rd"]
    assert got.shape == want.shape
    print("rd rel L2", rel_l2(got, want), "bit-equal", np.array_equal(got, want))
    assert rel_l2(got, want) <= RD_TOL, rel_l2(got, want)


def test_proj_rd_with_map_fails_like_the_reference():
    from ctsim_amd import mex
    from ctsim_amd._lib import CtsimError
    rec = load_golden(golden_files("cone_flat_map")[0])
    with pytest.raises((CtsimError, ValueError), match="cbct_geom_mex"):
        mex.cbct_geom_mex("proj,rd", rec["n"], rec["d"], rec["o"], rec["nuv"], rec["duv"], rec["ouv"], rec["sad"],
                          rec["add"], rec["sdd"], len(rec["betas"]), rec["betas"], rec["couchz"], rec["dettype"],
                          np.asfortranarray(rec["img"]), rec["map"])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_proj_rd_random_geometries_vs_oracle(seed):
    """per-operator entry points (2D and 3D, both detector types, offsets, couch z) against
    the FP32 oracle, which equals the compiled reference bit for bit"""
    from oracle import oracle as O
    from ctsim_amd._lib import CtsimGeom
    rng = np.random.default_rng(100 + seed)
    nx, ny, nz = (int(v) for v in rng.integers(10, 30, 3))
    nu, nv, nviews = int(rng.integers(16, 48)), int(rng.integers(40, 80)), 6
    dettype = seed % 2
    betas = rng.uniform(0, 2 * np.pi, nviews)
    cz = rng.uniform(-2, 2, nviews).astype(np.float32)
    d = [float(v) for v in rng.uniform(0.6, 1.3, 3)]
    o = [float(v) for v in rng.uniform(-1, 1, 3)]
    duv = [float(v) for v in rng.uniform(0.6, 1.4, 2)]
    ouv = [float(v) for v in rng.uniform(-1, 1, 2)]
    g3 = O.G3([nx, ny, nz], d, o, [nu, nv], duv, ouv, 600.0, 0, 1000.0, betas, couchz=cz, dettype=dettype)
    vol = rng.uniform(0, 1, (nx, ny, nz)).astype(np.float32)
    want = O.cbct("proj,rd", g3, vol)
    g = CtsimGeom(nx, ny, nz, *d, *o, nu, nv, *duv, *ouv, 600.0, 0.0, 1000.0, dettype)
    got = _per_op("ctsim_cbct_proj_rd", g, betas, cz, None, vol, (nv, nu, nviews))
    say("rd 3D seed", seed, rel_l2(got, want), np.array_equal(got, want))
    assert rel_l2(got, want) <= RD_TOL
    g2 = O.G2(nx, ny, d[0], d[1], o[0], o[1], nu, duv[0], ouv[0], 600.0, 0, 1000.0, betas, dettype)
    img = vol[:, :, 0].copy()
    want2 = O.fbct("proj,rd", g2, img)
    got2 = _per_op("ctsim_fbct_proj_rd", g, betas, None, None, img, (nu, nviews))
    assert rel_l2(got2, want2) <= RD_TOL


def test_cbct_back_taller_than_one_launch_vs_oracle():
    """more than 512 slices: the backprojectors run in z-slabs of 512 (a lane holds at most
    eight voxels of a column); every slab against the oracle's whole-column loop"""
    from oracle import oracle as O
    from ctsim_amd.device import DeviceArray
    from ctsim_amd.projector import Projector
    g, vol, sino, mp = _cone_case(12, n=(10, 12, 1100), nuv=(40, 180), nviews=6, dz=0.2, dv=1.4)
    say("slab case", g.n, g.nuv)
    P = Projector(geometry_from_g3(g, mp))
    d_sino = P.upload_sino(sino)
    out = DeviceArray(P.n_img, "f32")
    for op in ("back,dd", "back,tf", "back,pd", "back,hpd"):
        P.back(d_sino, out, op)
        got = P.download_volume(out)
        want = O.cbct(op, g, sino, mp)
        assert rel_l2(got, want) <= _tol(op, g, sino, mp), (op, rel_l2(got, want))
    # accumulate adds every slab once (onto a back,tf result)
    P.back(d_sino, out, "back,tf")
    P.back(d_sino, out, "back,tf", accumulate=True)
    assert rel_l2(P.download_volume(out), 2 * O.cbct("back,tf", g, sino, mp)) <= TOL
