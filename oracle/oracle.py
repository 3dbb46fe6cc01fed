"""ctypes front end of the CPU oracle (liboracle.so) and of the reference's own
compiled CPU kernels (_ref/libctsim_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg.  The product package ``ctsim_amd`` never imports it.

Both libraries take the MEX argument lists of ``fbct_geom_mex`` (14 arguments,
geom/fbct_os/fbct_geom_mex.c:250-315) and ``cbct_geom_mex`` (15 / 16 arguments,
geom/cbct_os/cbct_geom_mex.c:516-586), in MATLAB column-major layout, as numpy arrays:
image ``[nx, ny(, nz)]`` Fortran order, sinogram ``[nu, nviews]`` or ``[nv, nu, nviews]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CTSIM_ORACLE_LIB lets a test session load an alternative build of the oracle
_LIB = os.environ.get("CTSIM_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libctsim_ref.so")

_lib = None
_ref = None


def build(ref: bool = True) -> None:
    """Compile liboracle.so and, when the reference tree is present, _ref/libctsim_ref.so."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/geom"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build(ref=False)
        L = C.CDLL(_LIB)
        f = L.oracle_fbct
        f.restype = C.c_int
        f.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                      C.POINTER(C.c_double), C.c_int32, C.c_void_p, C.c_void_p, C.c_int]
        f = L.oracle_cbct
        f.restype = C.c_int
        f.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float,
                      C.c_float, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_int32, C.c_void_p,
                      C.POINTER(C.c_uint8), C.c_void_p, C.c_int]
        f = L.oracle_pmat
        f.restype = C.c_int
        f.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.POINTER(C.c_int32), C.c_float, C.c_float, C.c_int32, C.POINTER(C.c_float),
                      C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                      C.c_void_p, C.c_void_p, C.c_int]
        f = L.oracle_cbct_count
        f.restype = C.c_int
        f.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float,
                      C.c_float, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_int32,
                      C.POINTER(C.c_uint8), C.c_int, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(_REF)
        f = L.ref_fbct
        f.restype = C.c_int
        f.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                      C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.POINTER(C.c_double),
                      C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        f = L.ref_cbct
        f.restype = C.c_int
        f.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float,
                      C.c_float, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_int,
                      C.POINTER(C.c_float), C.POINTER(C.c_uint8), C.POINTER(C.c_float)]
        _ref = L
    return _ref


class G2:
    """2D fan-beam geometry in MEX argument form (geom/fbct_os/fbct_geom_mex.c:99-114)."""

    def __init__(self, nx, ny, dx, dy, ox, oy, nu, du, ou, sad, add, sdd, betas, dettype=1):
        self.nxy = np.array([nx, ny], np.int32)
        self.dxy = np.array([dx, dy], np.float32)
        self.oxy = np.array([ox, oy], np.float32)
        self.nu, self.du, self.ou = int(nu), float(du), float(ou)
        self.sad, self.add, self.sdd = float(sad), float(add), float(sdd)
        self.betas = np.ascontiguousarray(betas, np.float64)
        self.nviews = len(self.betas)
        self.dettype = int(dettype)

    def img_shape(self):
        return (int(self.nxy[0]), int(self.nxy[1]))

    def sino_shape(self):
        return (self.nu, self.nviews)


class G3:
    """3D cone-beam geometry in MEX argument form (geom/cbct_os/cbct_geom_mex.c:103-125)."""

    def __init__(self, n, d, o, nuv, duv, ouv, sad, add, sdd, betas, couchz=None, dettype=1):
        self.n = np.array(n, np.int32)
        self.d = np.array(d, np.float32)
        self.o = np.array(o, np.float32)
        self.nuv = np.array(nuv, np.int32)
        self.duv = np.array(duv, np.float32)
        self.ouv = np.array(ouv, np.float32)
        self.sad, self.add, self.sdd = float(sad), float(add), float(sdd)
        self.betas = np.ascontiguousarray(betas, np.float64)
        self.nviews = len(self.betas)
        self.couchz = (np.zeros(self.nviews, np.float32) if couchz is None
                       else np.ascontiguousarray(couchz, np.float32))
        self.dettype = int(dettype)

    def img_shape(self):
        return tuple(int(x) for x in self.n)

    def sino_shape(self):
        return (int(self.nuv[1]), int(self.nuv[0]), self.nviews)


def _out_shape(g, op):
    return g.sino_shape() if op.startswith("proj") else g.img_shape()


def fbct(op: str, g: G2, data: np.ndarray, prec: int = 32, nthreads: int = 1) -> np.ndarray:
    dt = np.float32 if prec == 32 else np.float64
    src = np.asfortranarray(data, dt)
    out = np.zeros(_out_shape(g, op), dt, order="F")
    rc = lib().oracle_fbct(op.encode(), prec, _p(g.nxy, C.c_int32), _p(g.dxy, C.c_float), _p(g.oxy, C.c_float),
                           g.nu, g.du, g.ou, g.sad, g.add, g.sdd, g.nviews, _p(g.betas, C.c_double), g.dettype,
                           src.ctypes.data, out.ctypes.data, nthreads)
    if rc:
        raise ValueError(f"oracle_fbct({op!r}) failed with {rc}")
    return out


def cbct(op: str, g: G3, data: np.ndarray, map_=None, prec: int = 32, nthreads: int = 1) -> np.ndarray:
    dt = np.float32 if prec == 32 else np.float64
    src = np.asfortranarray(data, dt)
    m = None if map_ is None else np.asfortranarray(map_, np.uint8)
    out = np.zeros(_out_shape(g, op), dt, order="F")
    rc = lib().oracle_cbct(op.encode(), prec, _p(g.n, C.c_int32), _p(g.d, C.c_float), _p(g.o, C.c_float),
                           _p(g.nuv, C.c_int32), _p(g.duv, C.c_float), _p(g.ouv, C.c_float), g.sad, g.add, g.sdd,
                           g.nviews, _p(g.betas, C.c_double), _p(g.couchz, C.c_float), g.dettype, src.ctypes.data,
                           _p(m, C.c_uint8), out.ctypes.data, nthreads)
    if rc:
        raise ValueError(f"oracle_cbct({op!r}) failed with {rc}")
    return out


def cbct_count(kind: str, g: G3, map_=None, nthreads: int = 8):
    """(U, C, K): voxel-view updates, voxel-cell weight pairs and column-view pairs of a 3D DD/SF op."""
    m = None if map_ is None else np.asfortranarray(map_, np.uint8)
    out = np.zeros(3, np.float64)
    rc = lib().oracle_cbct_count(kind.encode(), _p(g.n, C.c_int32), _p(g.d, C.c_float), _p(g.o, C.c_float),
                                 _p(g.nuv, C.c_int32), _p(g.duv, C.c_float), _p(g.ouv, C.c_float), g.sad, g.add,
                                 g.sdd, g.nviews, _p(g.betas, C.c_double), _p(g.couchz, C.c_float), g.dettype,
                                 _p(m, C.c_uint8), nthreads, _p(out, C.c_double))
    if rc:
        raise ValueError("oracle_cbct_count failed")
    return tuple(float(x) for x in out)


def ref_fbct(op: str, g: G2, data: np.ndarray) -> np.ndarray:
    """The reference's own compiled 2D kernel (oracle/_ref)."""
    src = np.asfortranarray(data, np.float32).copy(order="F")
    out = np.zeros(_out_shape(g, op), np.float32, order="F")
    nxy = g.nxy.astype(np.intc)
    betas = g.betas.copy()
    rc = ref().ref_fbct(op.encode(), _p(nxy, C.c_int), _p(g.dxy, C.c_float), _p(g.oxy, C.c_float), g.nu, g.du,
                        g.ou, g.sad, g.add, g.sdd, g.nviews, _p(betas, C.c_double), g.dettype,
                        _p(src, C.c_float), _p(out, C.c_float))
    if rc:
        raise ValueError(f"ref_fbct({op!r}) failed")
    return out


def ref_cbct(op: str, g: G3, data: np.ndarray, map_=None) -> np.ndarray:
    """The reference's own compiled 3D kernel (oracle/_ref); map None = 15-argument family."""
    src = np.asfortranarray(data, np.float32).copy(order="F")
    m = None if map_ is None else np.asfortranarray(map_, np.uint8).copy(order="F")
    out = np.zeros(_out_shape(g, op), np.float32, order="F")
    n = g.n.astype(np.intc)
    nuv = g.nuv.astype(np.intc)
    betas = g.betas.copy()
    couchz = g.couchz.copy()
    rc = ref().ref_cbct(op.encode(), _p(n, C.c_int), _p(g.d, C.c_float), _p(g.o, C.c_float), _p(nuv, C.c_int),
                        _p(g.duv, C.c_float), _p(g.ouv, C.c_float), g.sad, g.add, g.sdd, g.nviews,
                        _p(betas, C.c_double), _p(couchz, C.c_float), g.dettype, _p(src, C.c_float),
                        _p(m, C.c_uint8), _p(out, C.c_float))
    if rc:
        raise ValueError(f"ref_cbct({op!r}) failed")
    return out


class GP:
    """Projection-matrix cone-beam geometry in the argument form of the PMat gateway
    (geom/pcbct/cbct_geom_pmat_mex.c:19-128): pmat [3 4 nviews], cameras [3 nviews],
    sdds [nviews], offset_uv [2 nviews] (float32, Fortran order), map [nx ny]."""

    def __init__(self, n, d, o, nuv, du, dv, pmat, cameras, sdds, ouv, map_=None):
        self.n = np.array(n, np.int32)
        self.d = np.array(d, np.float32)
        self.o = np.array(o, np.float32)
        self.nuv = np.array(nuv, np.int32)
        self.du, self.dv = float(du), float(dv)
        self.pmat = np.asfortranarray(pmat, np.float32)
        self.nviews = int(self.pmat.shape[2])
        self.cameras = np.asfortranarray(cameras, np.float32)
        self.sdds = np.ascontiguousarray(sdds, np.float32)
        self.ouv = np.asfortranarray(ouv, np.float32)
        self.map = (np.ones((int(n[0]), int(n[1])), np.uint8) if map_ is None
                    else np.asfortranarray(map_, np.uint8))

    def img_shape(self):
        return tuple(int(x) for x in self.n)

    def sino_shape(self):
        return (int(self.nuv[1]), int(self.nuv[0]), self.nviews)


def circular_pmat(betas, sad, sdd, nuv, duv, ouv=(0.0, 0.0)):
    """Projection matrices, camera positions and distances of a circular flat-detector orbit
    in the conventions of geom/cbct_os (source at (-SAD sin b, -SAD cos b), detector
    coordinate t = x cos b - y sin b, depth D = SAD + x sin b + y cos b): u, v are detector
    cell indices with cell centres at integers, u = (SDD / du) t / D - wu,
    v = (SDD / dv) z / D - wv, w = -(n - 1) / 2 + offset.  Float64 arrays."""
    betas = np.asarray(betas, np.float64)
    nu, nv = nuv
    wu = -(nu - 1) / 2.0 + ouv[0]
    wv = -(nv - 1) / 2.0 + ouv[1]
    P = np.zeros((3, 4, len(betas)))
    cams = np.zeros((3, len(betas)))
    for k, b in enumerate(betas):
        s, c = np.sin(b), np.cos(b)
        depth = np.array([s, c, 0.0, sad])
        P[0, :, k] = (sdd / duv[0]) * np.array([c, -s, 0.0, 0.0]) - wu * depth
        P[1, :, k] = (sdd / duv[1]) * np.array([0.0, 0.0, 1.0, 0.0]) - wv * depth
        P[2, :, k] = depth
        cams[:, k] = (-sad * s, -sad * c, 0.0)
    ouvs = np.tile(np.asarray(ouv, np.float64)[:, None], (1, len(betas)))
    return P, cams, np.full(len(betas), float(sdd)), ouvs


def pmat(op: str, g: GP, data: np.ndarray, prec: int = 32, nthreads: int = 1) -> np.ndarray:
    dt = np.float32 if prec == 32 else np.float64
    src = np.asfortranarray(data, dt)
    out = np.zeros(g.sino_shape() if op.startswith("proj") else g.img_shape(), dt, order="F")
    rc = lib().oracle_pmat(op.encode(), prec, _p(g.n, C.c_int32), _p(g.d, C.c_float), _p(g.o, C.c_float),
                           _p(g.nuv, C.c_int32), g.du, g.dv, g.nviews, _p(g.pmat, C.c_float),
                           _p(g.cameras, C.c_float), _p(g.sdds, C.c_float), _p(g.ouv, C.c_float),
                           _p(g.map, C.c_uint8), src.ctypes.data, out.ctypes.data, nthreads)
    if rc:
        raise ValueError(f"oracle_pmat({op!r}) failed with {rc}")
    return out
