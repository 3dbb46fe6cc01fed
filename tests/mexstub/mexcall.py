"""Drive the compiled mexFunction shims from Python through the stand-in array model.

    m = MexShims(build_mex.build())
    out, err = m.call("cbct_geom_mex", "proj,tf", nxyz, dxyz, ...)   # MATLAB argument order

Arguments are numpy arrays (their dtype picks the mxClassID: float64 double, float32
single, int32 int32, bool logical, uint8 uint8) or str (a char array).  Returns the
first output as a numpy array in MATLAB (Fortran) order, or None with the message
that mexErrMsgTxt raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LOGICAL, CHAR, DOUBLE, SINGLE, INT8, UINT8, INT16, UINT16, INT32, UINT32, INT64, UINT64 = (
    3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
_CLASS = {np.dtype(np.float64): DOUBLE, np.dtype(np.float32): SINGLE, np.dtype(np.int32): INT32,
          np.dtype(np.uint32): UINT32, np.dtype(np.int16): INT16, np.dtype(np.uint16): UINT16,
          np.dtype(np.int8): INT8, np.dtype(np.uint8): UINT8, np.dtype(np.int64): INT64,
          np.dtype(np.uint64): UINT64, np.dtype(np.bool_): LOGICAL}
_NP = {SINGLE: np.float32, DOUBLE: np.float64}

MEXFN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p))


class MexShims:
    def __init__(self, paths: dict):
        lib_dir = os.path.dirname(next(iter(paths.values())))
        self.stub = C.CDLL(os.path.join(lib_dir, "libmexstub.so"), mode=C.RTLD_GLOBAL)
        s = self.stub
        s.stub_wrap.restype = C.c_void_p
        s.stub_wrap.argtypes = [C.c_int, C.c_int, C.c_size_t, C.c_void_p]
        s.stub_free.argtypes = [C.c_void_p]
        s.stub_ndim.restype = C.c_size_t
        s.stub_ndim.argtypes = [C.c_void_p]
        s.stub_dim.restype = C.c_size_t
        s.stub_dim.argtypes = [C.c_void_p, C.c_int]
        s.mxGetData.restype = C.c_void_p
        s.mxGetData.argtypes = [C.c_void_p]
        s.mxGetClassID.restype = C.c_int
        s.mxGetClassID.argtypes = [C.c_void_p]
        s.stub_call.restype = C.c_int
        s.stub_call.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p),
                                C.c_char_p, C.c_size_t]
        self.fn = {}
        self._libs = {}
        for name, path in paths.items():
            L = C.CDLL(path)
            self._libs[name] = L
            self.fn[name] = C.cast(L.mexFunction, C.c_void_p).value

    def _arr(self, a, keep):
        if isinstance(a, str):
            d = np.frombuffer(a.encode("ascii"), np.uint8).astype(np.uint16)
            keep.append(d)
            return self.stub.stub_wrap(CHAR, 0, d.size, d.ctypes.data)
        d = np.asfortranarray(a)
        if d.dtype not in _CLASS:
            raise TypeError(f"no mxClassID for {d.dtype}")
        keep.append(d)
        return self.stub.stub_wrap(_CLASS[d.dtype], 0, d.size, d.ctypes.data)

    def call(self, name: str, *args, nlhs: int = 1):
        keep = []
        prhs = (C.c_void_p * max(1, len(args)))(*[self._arr(a, keep) for a in args])
        plhs = (C.c_void_p * max(1, nlhs))()
        msg = C.create_string_buffer(1024)
        try:
            rc = self.stub.stub_call(self.fn[name], nlhs, plhs, len(args), prhs, msg, 1024)
            out = None
            if plhs[0]:
                a = plhs[0]
                dims = tuple(int(self.stub.stub_dim(a, k)) for k in range(int(self.stub.stub_ndim(a))))
                cls = self.stub.mxGetClassID(a)
                n = int(np.prod(dims))
                buf = (C.c_char * (n * np.dtype(_NP[cls]).itemsize)).from_address(self.stub.mxGetData(a))
                out = np.frombuffer(buf, _NP[cls]).reshape(dims, order="F").copy(order="F")
                self.stub.stub_free(a)
            return (None, msg.value.decode()) if rc else (out, "")
        finally:
            for p in prhs:
                self.stub.stub_free(p)
