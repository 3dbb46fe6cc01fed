"""The MATLAB boundary: the mexFunction shims of ctsim_amd/csrc/mex/ compiled against a
stand-in mex.h (tests/mexstub/) and called with hand-made mxArrays in the argument
order of geom/shell/forwardProjectMex.m:41-85 and backProjectMex.m:140-184.

CPU: every shim compiles with -Wall -Werror against MATLAB's mx* signatures and the
header, and the argument checks answer with the reference gateways' messages
(geom/fbct_os/fbct_geom_mex.c:122-135,266, geom/cbct_os/cbct_geom_mex.c:135-155).
GPU: a valid call through each shim equals the oracle (relative L2 <= 1e-5), with the
scalars and size vectors passed in the classes MATLAB callers use (int32, single,
double: the reference reads them through unchecked casts, geom/dd_cuda/ct_def.h:19-22).
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_l2

sys.path.insert(0, os.path.join(ROOT, "tests", "mexstub"))


@pytest.fixture(scope="module")
def shims():
    import build_mex
    from mexcall import MexShims
    return MexShims(build_mex.build())


def fan_args(op, data, nviews=12, betas=None):
    betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False) if betas is None else betas
    return [op, np.array([24, 20], np.int32), np.array([0.8, 0.8], np.float32), np.array([0.5, -0.25], np.float32),
            np.int32(48), np.float32(1.0), np.float32(0.25), np.float32(1000.0), np.float32(500.0),
            np.float32(1500.0), np.int32(nviews), betas, np.int32(1), data]


def cone_args(op, data, map_=True, nviews=10, dettype=1, classes="matlab"):
    betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    i32 = np.int32 if classes == "matlab" else np.float64
    f32 = np.float32 if classes == "matlab" else np.float64
    a = [op, np.array([20, 18, 12], i32), np.array([0.6, 0.6, 0.8], f32), np.array([0.25, -0.5, 0.5], f32),
         np.array([40, 24], i32), np.array([0.5, 0.6], f32), np.array([-0.5, 0.25], f32), f32(1000.0),
         f32(500.0), f32(1500.0), i32(nviews), betas, np.linspace(-1, 1, nviews).astype(np.float32),
         i32(dettype), data]
    if map_ is not None:
        m = np.ones((20, 18), bool)
        m[:3, :] = False
        a.append(m)
    return a


def test_shims_compile_against_the_mex_signatures(shims):
    assert set(shims.fn) == {"fbct_geom_mex", "cbct_geom_mex", "cbct_dd_cuda", "cbct_sf_cuda", "cbct_back_pd_cuda",
                             "cbct_geom_pmat_mex"}


def test_argument_count_and_op_checks(shims):
    img2 = np.zeros((24, 20), np.float32)
    out, err = shims.call("fbct_geom_mex", *fan_args("proj,dd", img2)[:13])
    assert out is None and err == "Error: incorrect number of arguments."
    out, err = shims.call("fbct_geom_mex", *fan_args("proj,dd", img2), nlhs=2)
    assert err == "Error: incorrect number of arguments."
    a = fan_args("proj,dd", img2)
    a[0] = np.float64(3.0)
    assert shims.call("fbct_geom_mex", *a)[1] == "The 1st argument is invalid!"
    vol = np.zeros((20, 18, 12), np.float32)
    # the CUDA-named gateways take the 16-argument form only; cbct_geom_mex also 15
    assert shims.call("cbct_dd_cuda", *cone_args("proj,dd", vol, map_=None))[1] == \
        "Error: incorrect number of arguments."
    assert shims.call("cbct_geom_mex", *cone_args("proj,xx", vol, map_=None))[1] == "Error: unknown geom type."
    assert shims.call("cbct_sf_cuda", *cone_args("proj,dd", vol))[1] == "Error: unknown geom type."


def test_pmat_argument_checks(shims):
    n, nuv, nviews = (8, 8, 4), (16, 8), 3
    a = ["proj,dd", np.array(n, np.int32), np.ones(3, np.float32), np.zeros(3, np.float32), np.array(nuv, np.int32),
         np.ones(2, np.float32), np.int32(nviews), np.zeros((3, 4, nviews), np.float32),
         np.zeros((3, nviews), np.float32), np.ones(nviews, np.float32), np.zeros((2, nviews), np.float32),
         np.zeros(n, np.float32), np.ones(n[:2], bool)]
    assert shims.call("cbct_geom_pmat_mex", *a[:12])[1] == "Error: incorrect number of arguments."
    b = list(a)
    b[7] = b[7].astype(np.float64)
    assert shims.call("cbct_geom_pmat_mex", *b)[1] == "Error: mx_pmat must have [3 4 noViews ] float array."
    b = list(a)
    b[8] = b[8][:2]
    assert shims.call("cbct_geom_pmat_mex", *b)[1] == "Error: mx_cameras must have noviews X 2 float array."
    b = list(a)
    b[9] = b[9][:2]
    assert shims.call("cbct_geom_pmat_mex", *b)[1] == "Error: mx_sdds must have noviews X 1 float array."
    b = list(a)
    b[10] = b[10].astype(np.float64)
    assert shims.call("cbct_geom_pmat_mex", *b)[1] == "Error: mx_cuvs must have noviews X 2 float array."
    b = list(a)
    b[12] = b[12].astype(np.uint8)
    assert shims.call("cbct_geom_pmat_mex", *b)[1] == "Error: mx_map must have [nx ny] bool array."
    b = list(a)
    b[0] = "back,xx"
    assert shims.call("cbct_geom_pmat_mex", *b)[1] == "Error: unknown geom type."


def test_argument_class_checks(shims):
    vol = np.zeros((20, 18, 12), np.float32)
    a = cone_args("proj,tf", vol)
    a[11] = a[11].astype(np.float32)
    assert shims.call("cbct_geom_mex", *a)[1] == "Error: mx_betas must have noviews X 1 double array."
    a = cone_args("proj,tf", vol)
    a[11] = a[11][:-1]
    assert shims.call("cbct_geom_mex", *a)[1] == "Error: mx_betas must have noviews X 1 double array."
    a = cone_args("proj,tf", vol)
    a[12] = a[12].astype(np.float64)
    assert shims.call("cbct_geom_mex", *a)[1] == "Error: mx_couchz must have noviews X 1 double array."
    a = cone_args("proj,tf", vol)
    a[14] = vol.astype(np.float64)
    assert shims.call("cbct_geom_mex", *a)[1] == "Error: mx_image must have [nx ny nz] float array."
    a = cone_args("back,tf", vol)  # a volume where the sinogram belongs
    assert shims.call("cbct_geom_mex", *a)[1] == "Error: mx_proj must have [nv nu noviews] float array."
    a = cone_args("proj,tf", vol)
    a[15] = a[15][:, :17]
    assert shims.call("cbct_geom_mex", *a)[1] == "Error: mx_map must have [nx ny] bool array."
    a = cone_args("proj,tf", vol)
    a[1] = np.array([20, 18], np.int32)
    assert "nxyz must be a real numeric array of 3" in shims.call("cbct_geom_mex", *a)[1]


@pytest.mark.gpu
def test_shim_calls_match_the_oracle(shims, oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(7)
    # 2D fan beam, 14 arguments
    img2 = rng.uniform(0, 1, (24, 20)).astype(np.float32)
    g2 = O.G2(24, 20, 0.8, 0.8, 0.5, -0.25, 48, 1.0, 0.25, 1000.0, 500.0, 1500.0,
              np.linspace(0, 2 * np.pi, 12, endpoint=False), 1)
    for op in ("proj,dd", "proj,tf"):
        out, err = shims.call("fbct_geom_mex", *fan_args(op, img2))
        assert not err, err
        assert out.shape == (48, 12)
        assert rel_l2(out, O.fbct(op, g2, img2)) <= 1e-5, op
    sino2 = rng.uniform(0, 1, (48, 12)).astype(np.float32)
    for op in ("back,dd", "back,tf", "back,pd"):
        out, err = shims.call("fbct_geom_mex", *fan_args(op, sino2))
        assert not err, err
        assert out.shape == (24, 20)
        assert rel_l2(out, O.fbct(op, g2, sino2)) <= 1e-5, op
    # 3D cone beam, 16 arguments with the map and 15 without; scalars as MATLAB classes or doubles
    nviews = 10
    betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    g3 = O.G3((20, 18, 12), (0.6, 0.6, 0.8), (0.25, -0.5, 0.5), (40, 24), (0.5, 0.6), (-0.5, 0.25), 1000.0, 500.0,
              1500.0, betas, couchz=np.linspace(-1, 1, nviews).astype(np.float32), dettype=1)
    m = np.ones((20, 18), bool)
    m[:3, :] = False
    # Fortran order: the shim reads these buffers themselves, so a write would show
    vol = np.asfortranarray(rng.uniform(0, 1, (20, 18, 12)).astype(np.float32))
    sino = np.asfortranarray(rng.uniform(0, 1, (24, 40, nviews)).astype(np.float32))
    vol0, sino0 = vol.copy(), sino.copy()
    cases = [("cbct_geom_mex", "proj,tf", vol, True, "matlab"), ("cbct_geom_mex", "proj,dd", vol, False, "double"),
             ("cbct_geom_mex", "back,tf", sino, True, "matlab"), ("cbct_geom_mex", "back,hpd", sino, False, "matlab"),
             ("cbct_dd_cuda", "proj,dd", vol, True, "double"), ("cbct_dd_cuda", "back,dd", sino, True, "matlab"),
             ("cbct_sf_cuda", "proj,tf", vol, True, "matlab"), ("cbct_sf_cuda", "back,tf", sino, True, "double"),
             ("cbct_back_pd_cuda", "back,pd", sino, True, "matlab"),
             ("cbct_back_pd_cuda", "back,hpd", sino, True, "matlab")]
    for name, op, data, with_map, classes in cases:
        out, err = shims.call(name, *cone_args(op, data, map_=True if with_map else None, classes=classes))
        assert not err, (name, op, err)
        want = O.cbct(op, g3, data, map_=m if with_map else None)
        assert out.shape == want.shape, (name, op)
        e = rel_l2(out, want)
        print(f"  {name} {op} ({'map' if with_map else 'no map'}, {classes} classes): rel-L2 {e:.2e}", flush=True)
        assert e <= 1e-5, (name, op, e)
    assert np.array_equal(vol, vol0) and np.array_equal(sino, sino0)  # inputs untouched


@pytest.mark.gpu
def test_pmat_shim_matches_the_oracle(shims, oracle_mod):
    """cbct_geom_pmat_mex through its shim: a three-element duv read as the reference reads it
    (duv[1], duv[2]), a two-element one as (du, dv), against the PMat oracle."""
    O = oracle_mod
    n, nuv, nviews = (20, 18, 12), (40, 24), 9
    P, cams, sdds, ouv = O.circular_pmat(np.linspace(0, 2 * np.pi, nviews, endpoint=False), 1000.0, 1500.0, nuv,
                                         (0.5, 0.6))
    m = np.ones(n[:2], bool)
    m[:3, :] = False
    g = O.GP(n, (0.6, 0.6, 0.8), (0.25, -0.5, 0.5), nuv, 0.5, 0.6, P, cams, sdds, ouv, map_=m)
    rng = np.random.default_rng(5)
    vol = np.asfortranarray(rng.uniform(0, 1, n).astype(np.float32))
    sino = np.asfortranarray(rng.uniform(0, 1, g.sino_shape()).astype(np.float32))
    for duv in (np.array([0.5, 0.6], np.float32), np.array([9.0, 0.5, 0.6], np.float32)):
        for op in ("proj,dd", "back,dd", "back,hpd"):
            data = vol if op.startswith("proj") else sino
            out, err = shims.call("cbct_geom_pmat_mex", op, np.array(n, np.int32), g.d, g.o, np.array(nuv, np.int32),
                                  duv, np.int32(nviews), g.pmat, g.cameras, g.sdds, g.ouv, data, m)
            assert not err, (op, err)
            e = rel_l2(out, O.pmat(op, g, data))
            print(f"  cbct_geom_pmat_mex {op} (duv of {duv.size}): rel-L2 {e:.2e}", flush=True)
            assert e <= 1e-5, (op, e)
    out, err = shims.call("cbct_geom_pmat_mex", "proj,rd", np.array(n, np.int32), g.d, g.o, np.array(nuv, np.int32),
                          np.array([0.5, 0.6], np.float32), np.int32(nviews), g.pmat, g.cameras, g.sdds, g.ouv, vol, m)
    assert out is None and "ray driven method is not supported" in err
