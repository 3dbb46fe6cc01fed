"""Python mirrors of CTSim's projector MEX functions, running on the GPU.

Same argument lists, classes and validation messages as the MEX gateways:
``fbct_geom_mex`` (geom/fbct_os/fbct_geom_mex.c:75-326, 14 arguments) and
``cbct_geom_mex`` (geom/cbct_os/cbct_geom_mex.c:83-600, 15 or 16 arguments), plus the
names of the reference's CUDA gateways (``cbct_dd_cuda``, ``cbct_sf_cuda``,
``cbct_back_pd_cuda``).  Arrays are numpy in MATLAB column-major order; outputs are new
float32 Fortran-ordered arrays, as ``mxCreateNumericArray`` would return.  Every call
goes through the C ABI of libctsim_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import P_f32, P_f64, P_i32, P_u8, call

_OPS2 = ("proj,rd", "proj,dd", "proj,tf", "back,pd", "back,dd", "back,tf")
_OPS3 = ("proj,rd", "proj,dd", "proj,tf", "back,pd", "back,dd", "back,tf", "back,hpd")


def _arr(x, dt, n=None):
    a = np.ascontiguousarray(np.atleast_1d(x), dt)
    if n is not None and a.size < n:
        raise ValueError("Error: incorrect number of arguments.")
    return a


def _ptr(a, ct):
    return a.ctypes.data_as(ct)


def _check_op(op, ops):
    if not isinstance(op, str):
        raise ValueError("The 1st argument is invalid!")
    if op not in ops:
        raise ValueError("Error: unknown geom type.")


def _f32_data(x, shape, msg):
    a = np.asarray(x)
    if a.dtype != np.float32 or tuple(a.shape[:len(shape)]) != tuple(shape) or a.size != int(np.prod(shape)):
        raise ValueError(msg)
    return np.asfortranarray(a)


def fbct_geom_mex(op, nxy, dxy, oxy, nu, du, ou, sad, add, sdd, noviews, betas, detector_type, data):
    """2D fan-beam projector gateway (fbct_geom_mex)."""
    _check_op(op, _OPS2)
    nxy = _arr(nxy, np.int32, 2)
    dxy = _arr(dxy, np.float32, 2)
    oxy = _arr(oxy, np.float32, 2)
    nu = int(np.atleast_1d(nu)[0])
    noviews = int(np.atleast_1d(noviews)[0])
    b = np.asarray(betas)
    if b.size != noviews or b.dtype != np.float64:
        raise ValueError("Error: mx_betas must have noviews X 1 double array.")
    b = np.ascontiguousarray(b.ravel())
    if op.startswith("proj"):
        src = _f32_data(data, (nxy[0], nxy[1]), "Error: mx_image must have [nx ny] float array.")
        out = np.empty((nu, noviews), np.float32, order="F")
    else:
        src = _f32_data(data, (nu, noviews), "Error: mx_proj must have [nu noviews] float array.")
        out = np.empty((int(nxy[0]), int(nxy[1])), np.float32, order="F")
    err = C.create_string_buffer(512)
    call("ctsim_fbct_geom_mex", op.encode(), _ptr(nxy, P_i32), _ptr(dxy, P_f32), _ptr(oxy, P_f32), nu,
         float(np.float32(np.atleast_1d(du)[0])), float(np.float32(np.atleast_1d(ou)[0])),
         float(np.float32(sad)), float(np.float32(add)), float(np.float32(sdd)), noviews, _ptr(b, P_f64),
         int(np.atleast_1d(detector_type)[0]), src.ctypes.data_as(P_f32), out.ctypes.data_as(P_f32), err, 512)
    return out


def _cbct(entry, ops, op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews, betas, couchz, detector_type,
          data, map_=None, require_map=False):
    _check_op(op, ops)
    nxyz = _arr(nxyz, np.int32, 3)
    dxyz = _arr(dxyz, np.float32, 3)
    cxyz = _arr(cxyz, np.float32, 3)
    nuv = _arr(nuv, np.int32, 2)
    duv = _arr(duv, np.float32, 2)
    cuv = _arr(cuv, np.float32, 2)
    noviews = int(np.atleast_1d(noviews)[0])
    b = np.asarray(betas)
    if b.size != noviews or b.dtype != np.float64:
        raise ValueError("Error: mx_betas must have noviews X 1 double array.")
    cz = np.asarray(couchz)
    if cz.size != noviews or cz.dtype != np.float32:
        raise ValueError("Error: mx_couchz must have noviews X 1 double array.")
    b = np.ascontiguousarray(b.ravel())
    cz = np.ascontiguousarray(cz.ravel())
    nx, ny, nz = (int(v) for v in nxyz)
    nu, nv = int(nuv[0]), int(nuv[1])
    if op.startswith("proj"):
        src = _f32_data(data, (nx, ny, nz), "Error: mx_image must have [nx ny nz] float array.")
        out = np.empty((nv, nu, noviews), np.float32, order="F")
    else:
        src = _f32_data(data, (nv, nu, noviews), "Error: mx_proj must have [nv nu noviews] float array.")
        out = np.empty((nx, ny, nz), np.float32, order="F")
    mp = None
    if map_ is not None:
        m = np.asarray(map_)
        if m.dtype != np.bool_ or m.shape[:2] != (nx, ny) or m.size != nx * ny:
            raise ValueError("Error: mx_map must have [nx ny] bool array.")
        mp = np.asfortranarray(m).view(np.uint8)
        mp = np.ascontiguousarray(mp.ravel(order="F"))
    elif require_map:
        raise ValueError("Error: incorrect number of arguments.")
    err = C.create_string_buffer(512)
    call(entry, op.encode(), _ptr(nxyz, P_i32), _ptr(dxyz, P_f32), _ptr(cxyz, P_f32), _ptr(nuv, P_i32),
         _ptr(duv, P_f32), _ptr(cuv, P_f32), float(np.float32(sad)), float(np.float32(add)), float(np.float32(sdd)),
         noviews, _ptr(b, P_f64), _ptr(cz, P_f32), int(np.atleast_1d(detector_type)[0]), src.ctypes.data_as(P_f32),
         None if mp is None else mp.ctypes.data_as(P_u8), out.ctypes.data_as(P_f32), err, 512)
    return out


def cbct_geom_mex(op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews, betas, couchz, detector_type, data,
                  map=None):
    """3D cone-beam projector gateway (cbct_geom_mex); map=None is the 15-argument form."""
    return _cbct("ctsim_cbct_geom_mex", _OPS3, op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews, betas,
                 couchz, detector_type, data, map)


def cbct_dd_cuda(op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews, betas, couchz, detector_type, data,
                 map):
    """Name and 16-argument signature of geom/dd_cuda/cbct_dd_cuda.cu ('proj,dd', 'back,dd')."""
    return _cbct("ctsim_cbct_dd_cuda", ("proj,dd", "back,dd"), op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd,
                 noviews, betas, couchz, detector_type, data, map, require_map=True)


def cbct_sf_cuda(op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews, betas, couchz, detector_type, data,
                 map):
    """Name and 16-argument signature of geom/sf_cuda/cbct_sf_cuda.cu ('proj,tf', 'back,tf')."""
    return _cbct("ctsim_cbct_sf_cuda", ("proj,tf", "back,tf"), op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd,
                 noviews, betas, couchz, detector_type, data, map, require_map=True)


def cbct_back_pd_cuda(op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews, betas, couchz, detector_type,
                      data, map):
    """Name and 16-argument signature of geom/fdk_cuda/cbct_back_pd_cuda.cu ('back,pd', 'back,hpd')."""
    if op.startswith("proj"):
        raise ValueError("Error: unknown geom type.")  # the CUDA FDK gateway rejects 'proj' (:155-157)
    return _cbct("ctsim_cbct_back_pd_cuda", ("back,pd", "back,hpd"), op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add,
                 sdd, noviews, betas, couchz, detector_type, data, map, require_map=True)


_OPS_PMAT = ("proj,rd", "proj,dd", "proj,tf", "back,pd", "back,hpd", "back,dd", "back,tf")


def cbct_geom_pmat_mex(op, nxyz, dxyz, cxyz, nuv, duv, noviews, pmat, cameras, sdds, cuvs, data, map):
    """Projection-matrix cone-beam gateway (geom/pcbct/cbct_geom_pmat_mex.c, 13 arguments).

    pmat [3 4 noviews], cameras [3 noviews], sdds [noviews], cuvs [2 noviews], all single;
    map logical [nx ny] (required).  duv of three elements is read as the reference reads it
    (duv[1], duv[2]); two elements as (du, dv)."""
    _check_op(op, _OPS_PMAT)
    nxyz = _arr(nxyz, np.int32, 3)
    dxyz = _arr(dxyz, np.float32, 3)
    cxyz = _arr(cxyz, np.float32, 3)
    nuv = _arr(nuv, np.int32, 2)
    duv = _arr(duv, np.float32, 2)
    du, dv = (float(duv[1]), float(duv[2])) if duv.size >= 3 else (float(duv[0]), float(duv[1]))
    noviews = int(np.atleast_1d(noviews)[0])

    def single(x, n, msg):
        a = np.asarray(x)
        if a.dtype != np.float32 or a.size != n:
            raise ValueError(msg)
        return np.asfortranarray(a)

    pm = single(pmat, 12 * noviews, "Error: mx_pmat must have [3 4 noViews ] float array.")
    cams = single(cameras, 3 * noviews, "Error: mx_cameras must have noviews X 2 float array.")
    sd = single(sdds, noviews, "Error: mx_sdds must have noviews X 1 float array.")
    cu = single(cuvs, 2 * noviews, "Error: mx_cuvs must have noviews X 2 float array.")
    nx, ny, nz = (int(v) for v in nxyz)
    nu, nv = int(nuv[0]), int(nuv[1])
    if op.startswith("proj"):
        src = _f32_data(data, (nx, ny, nz), "Error: mx_image must have [nx ny nz] float array.")
        out = np.empty((nv, nu, noviews), np.float32, order="F")
    else:
        src = _f32_data(data, (nv, nu, noviews), "Error: mx_proj must have [nv nu noviews] float array.")
        out = np.empty((nx, ny, nz), np.float32, order="F")
    m = np.asarray(map)
    if m.dtype != np.bool_ or m.size != nx * ny:
        raise ValueError("Error: mx_map must have [nx ny] bool array.")
    mp = np.ascontiguousarray(np.asfortranarray(m.reshape((nx, ny), order="F")).view(np.uint8).ravel(order="F"))
    err = C.create_string_buffer(512)
    call("ctsim_cbct_geom_pmat_mex", op.encode(), _ptr(nxyz, P_i32), _ptr(dxyz, P_f32), _ptr(cxyz, P_f32),
         _ptr(nuv, P_i32), du, dv, noviews, pm.ctypes.data_as(P_f32), cams.ctypes.data_as(P_f32),
         sd.ctypes.data_as(P_f32), cu.ctypes.data_as(P_f32), src.ctypes.data_as(P_f32), mp.ctypes.data_as(P_u8),
         out.ctypes.data_as(P_f32), err, 512)
    return out
