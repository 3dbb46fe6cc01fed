"""CPU checks of the C ABI: the library loads, exports every function the header
declares, and rejects bad arguments with the reference's messages before touching
any device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ctsim_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ctsim_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_gateways():
    names = declared_functions()
    for n in ("ctsim_fbct_geom_mex", "ctsim_cbct_geom_mex", "ctsim_cbct_dd_cuda", "ctsim_cbct_sf_cuda",
              "ctsim_cbct_back_pd_cuda", "ctsim_plan_create", "ctsim_plan_proj", "ctsim_plan_back"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from ctsim_amd import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert not [n for n in declared_functions() if n not in _lib.SIGNATURES]


def _geom3_args():
    return (np.array([4, 4, 4], np.int32).ctypes.data_as(C.POINTER(C.c_int32)),
            np.ones(3, np.float32).ctypes.data_as(C.POINTER(C.c_float)),
            np.zeros(3, np.float32).ctypes.data_as(C.POINTER(C.c_float)),
            np.array([8, 8], np.int32).ctypes.data_as(C.POINTER(C.c_int32)),
            np.ones(2, np.float32).ctypes.data_as(C.POINTER(C.c_float)),
            np.zeros(2, np.float32).ctypes.data_as(C.POINTER(C.c_float)))


def test_gateway_rejects_unknown_op_without_device():
    from ctsim_amd import _lib
    L = _lib.lib()
    betas = np.zeros(2)
    cz = np.zeros(2, np.float32)
    buf = np.zeros(512, np.float32)
    err = C.create_string_buffer(256)
    keep = _geom3_args()
    rc = L.ctsim_cbct_geom_mex(b"proj,xx", *keep, 500.0, 0.0, 900.0, 2, betas.ctypes.data_as(C.POINTER(C.c_double)),
                               cz.ctypes.data_as(C.POINTER(C.c_float)), 1, buf.ctypes.data_as(C.POINTER(C.c_float)),
                               None, buf.ctypes.data_as(C.POINTER(C.c_float)), err, 256)
    assert rc == 1
    assert err.value.decode() == "Error: unknown geom type."
    rc = L.ctsim_cbct_sf_cuda(b"proj,dd", *keep, 500.0, 0.0, 900.0, 2, betas.ctypes.data_as(C.POINTER(C.c_double)),
                              cz.ctypes.data_as(C.POINTER(C.c_float)), 1, buf.ctypes.data_as(C.POINTER(C.c_float)),
                              None, buf.ctypes.data_as(C.POINTER(C.c_float)), err, 256)
    assert rc == 1 and "unknown geom type" in err.value.decode()


def test_per_operator_entry_points_validate_without_device():
    from ctsim_amd import _lib
    L = _lib.lib()
    err = C.create_string_buffer(256)
    buf = np.zeros(64, np.float32)
    fp = buf.ctypes.data_as(C.POINTER(C.c_float))
    betas = np.zeros(2)
    bp = betas.ctypes.data_as(C.POINTER(C.c_double))
    assert L.ctsim_cbct_proj_tf(None, 2, bp, None, None, fp, fp, err, 256) == 1
    assert err.value.decode() == "Error: null geometry."
    g = _lib.CtsimGeom(4, 4, 4, 1.0, 1.0, 1.0, 0, 0, 0, 8, 8, 1.0, 1.0, 0, 0, 500.0, 0.0, 900.0, 1)
    assert L.ctsim_fbct_back_pd(C.byref(g), 0, bp, None, None, fp, fp, err, 256) == 1
    assert "noviews" in err.value.decode()


def test_plan_rejects_bad_geometry_without_device():
    from ctsim_amd import _lib
    L = _lib.lib()
    g = _lib.CtsimGeom()
    g.nx, g.ny, g.nz, g.nu, g.nv = 0, 4, 4, 8, 8
    g.dx = g.dy = g.dz = g.du = g.dv = 1.0
    p = C.c_void_p()
    betas = np.zeros(3)
    rc = L.ctsim_plan_create(C.byref(p), 3, C.byref(g), 3, betas.ctypes.data_as(C.POINTER(C.c_double)), None, None, 0)
    assert rc == 1 and b"sizes must be positive" in L.ctsim_last_error()
    g.nx = 4
    rc = L.ctsim_plan_create(C.byref(p), 3, C.byref(g), 0, betas.ctypes.data_as(C.POINTER(C.c_double)), None, None, 0)
    assert rc == 1 and b"mx_betas" in L.ctsim_last_error()
    rc = L.ctsim_plan_create(C.byref(p), 2, C.byref(g), 3, betas.ctypes.data_as(C.POINTER(C.c_double)), None, None, 0)
    assert rc == 1 and b"nz = nv = 1" in L.ctsim_last_error()
    rc = L.ctsim_plan_create(C.byref(p), 3, C.byref(g), 3, betas.ctypes.data_as(C.POINTER(C.c_double)), None, None, 7)
    assert rc == 1


def test_python_mex_validation_messages():
    from ctsim_amd import mex
    with pytest.raises(ValueError, match="mx_betas"):
        mex.cbct_geom_mex("proj,tf", [4, 4, 4], [1, 1, 1], [0, 0, 0], [8, 8], [1, 1], [0, 0], 500, 0, 900, 3,
                          np.zeros(2), np.zeros(3, np.float32), 1, np.zeros((4, 4, 4), np.float32))
    with pytest.raises(ValueError, match="mx_couchz"):
        mex.cbct_geom_mex("proj,tf", [4, 4, 4], [1, 1, 1], [0, 0, 0], [8, 8], [1, 1], [0, 0], 500, 0, 900, 2,
                          np.zeros(2), np.zeros(2), 1, np.zeros((4, 4, 4), np.float32))
    with pytest.raises(ValueError, match=r"\[nx ny nz\] float"):
        mex.cbct_geom_mex("proj,tf", [4, 4, 4], [1, 1, 1], [0, 0, 0], [8, 8], [1, 1], [0, 0], 500, 0, 900, 2,
                          np.zeros(2), np.zeros(2, np.float32), 1, np.zeros((4, 4, 4), np.float64))
    with pytest.raises(ValueError, match=r"\[nv nu noviews\]"):
        mex.cbct_geom_mex("back,tf", [4, 4, 4], [1, 1, 1], [0, 0, 0], [8, 8], [1, 1], [0, 0], 500, 0, 900, 2,
                          np.zeros(2), np.zeros(2, np.float32), 1, np.zeros((8, 8, 3), np.float32))
    with pytest.raises(ValueError, match="mx_map"):
        mex.cbct_geom_mex("back,tf", [4, 4, 4], [1, 1, 1], [0, 0, 0], [8, 8], [1, 1], [0, 0], 500, 0, 900, 2,
                          np.zeros(2), np.zeros(2, np.float32), 1, np.zeros((8, 8, 2), np.float32),
                          np.ones((4, 4), np.uint8))
    with pytest.raises(ValueError, match="unknown geom type"):
        mex.fbct_geom_mex("back,hpd", [4, 4], [1, 1], [0, 0], 8, 1, 0, 500, 0, 900, 2, np.zeros(2), 1,
                          np.zeros((8, 2), np.float32))
    with pytest.raises(ValueError, match="unknown geom type"):
        mex.cbct_back_pd_cuda("proj,dd", [4, 4, 4], [1, 1, 1], [0, 0, 0], [8, 8], [1, 1], [0, 0], 500, 0, 900, 2,
                              np.zeros(2), np.zeros(2, np.float32), 1, np.zeros((4, 4, 4), np.float32),
                              np.ones((4, 4), bool))
