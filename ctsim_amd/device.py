"""Device memory, streams, events and graphs through the C ABI (no torch dependency)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BF16, F16, F32, F64, call, lib

_NP = {F32: np.float32, F16: np.float16, F64: np.float64, BF16: np.uint16}
_SIZE = {F32: 4, F16: 2, BF16: 2, F64: 8}
DTYPES = {"f32": F32, "f16": F16, "bf16": BF16, "f64": F64,
          "float32": F32, "float16": F16, "bfloat16": BF16, "float64": F64}


def dtype_code(dt) -> int:
    if isinstance(dt, int):
        return dt
    if isinstance(dt, str):
        return DTYPES[dt]
    dt = np.dtype(dt)
    return {np.dtype(np.float32): F32, np.dtype(np.float16): F16, np.dtype(np.float64): F64}[dt]


def bf16_from_f32(a: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even float32 -> bfloat16 bit patterns (uint16)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + 0x7FFF
    out = ((u + r) >> 16).astype(np.uint16)
    nan = np.isnan(a.reshape(-1)).reshape(a.shape)
    out[nan] = 0x7FC0
    return out


def f32_from_bf16(a: np.ndarray) -> np.ndarray:
    return (np.asarray(a, np.uint16).astype(np.uint32) << 16).view(np.float32)


class DeviceArray:
    """A device buffer of `count` elements of a CTSIM dtype, freed on garbage collection."""

    def __init__(self, count: int, dtype=F32, shape=None):
        self.dtype = dtype_code(dtype)
        self.count = int(count)
        self.shape = tuple(shape) if shape is not None else (self.count,)
        self.nbytes = self.count * _SIZE[self.dtype]
        p = C.c_void_p()
        call("ctsim_malloc", C.byref(p), self.nbytes)
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, a: np.ndarray, dtype=None, stream=None) -> "DeviceArray":
        """Copy a host array (C order of its memory as given; pass Fortran arrays as
        ravel(order='F') when MATLAB order is meant)."""
        dt = dtype_code(dtype if dtype is not None else a.dtype)
        if dt == BF16:
            host = bf16_from_f32(np.asarray(a, np.float32))
        else:
            host = np.ascontiguousarray(a, _NP[dt])
        d = cls(host.size, dt, shape=a.shape)
        call("ctsim_memcpy_h2d", C.c_void_p(d.ptr), host.ctypes.data_as(C.c_void_p), d.nbytes,
             C.c_void_p(stream.handle if stream else 0))
        return d

    def to_numpy(self, stream=None) -> np.ndarray:
        host = np.empty(self.count, _NP[self.dtype])
        call("ctsim_memcpy_d2h", host.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes,
             C.c_void_p(stream.handle if stream else 0))
        if self.dtype == BF16:
            host = f32_from_bf16(host)
        return host.reshape(self.shape)

    def fill(self, value: float, stream=None) -> "DeviceArray":
        call("ctsim_fill", C.c_void_p(self.ptr), self.dtype, self.count, float(value),
             C.c_void_p(stream.handle if stream else 0))
        return self

    def copy_from(self, other: "DeviceArray", stream=None) -> None:
        if other.count != self.count:
            raise ValueError("size mismatch")
        if other.dtype == self.dtype:
            call("ctsim_memcpy_d2d", C.c_void_p(self.ptr), C.c_void_p(other.ptr), self.nbytes,
                 C.c_void_p(stream.handle if stream else 0))
        else:
            call("ctsim_convert", C.c_void_p(other.ptr), other.dtype, C.c_void_p(self.ptr), self.dtype, self.count,
                 C.c_void_p(stream.handle if stream else 0))

    @property
    def vp(self) -> C.c_void_p:
        return C.c_void_p(self.ptr)

    @classmethod
    def window(cls, base: "DeviceArray", offset: int, count: int) -> "DeviceArray":
        """A view of count elements of base from element offset on; it does not own the
        memory (freeing it is a no-op) and must not outlive base."""
        if offset < 0 or count < 0 or offset + count > base.count:
            raise ValueError("window outside the array")
        v = cls.__new__(cls)
        v.dtype, v.count, v.shape = base.dtype, int(count), (int(count),)
        v.nbytes = v.count * _SIZE[v.dtype]
        v.ptr = base.ptr + offset * _SIZE[base.dtype]
        v._view = True
        return v

    def free(self) -> None:
        if getattr(self, "_view", False):
            self.ptr = 0
            return
        if getattr(self, "ptr", 0):
            try:
                lib().ctsim_free(C.c_void_p(self.ptr))
            finally:
                self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        p = C.c_void_p()
        call("ctsim_stream_create", C.byref(p))
        self.handle = p.value

    @property
    def vp(self):
        return C.c_void_p(self.handle)

    def synchronize(self) -> None:
        call("ctsim_stream_synchronize", C.c_void_p(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().ctsim_stream_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        call("ctsim_event_create", C.byref(p))
        self.handle = p.value

    def record(self, stream: Stream | None = None) -> None:
        call("ctsim_event_record", C.c_void_p(self.handle), C.c_void_p(stream.handle if stream else 0))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        call("ctsim_event_elapsed_ms", C.c_void_p(self.handle), C.c_void_p(end.handle), C.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib().ctsim_event_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


class Graph:
    """Stream capture into an instantiated hipGraph."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.handle = None

    def __enter__(self):
        call("ctsim_graph_begin", C.c_void_p(self.stream.handle))
        return self

    def __exit__(self, et, ev, tb):
        h = C.c_void_p()
        status = lib().ctsim_graph_end(C.c_void_p(self.stream.handle), C.byref(h))
        if et is None:
            from ._lib import check
            check(status)
            self.handle = h.value
        return False

    def launch(self) -> None:
        call("ctsim_graph_launch", C.c_void_p(self.handle), C.c_void_p(self.stream.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().ctsim_graph_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


def synchronize() -> None:
    call("ctsim_device_synchronize")


def device_count() -> int:
    n = C.c_int32()
    status = lib().ctsim_device_count(C.byref(n))
    return int(n.value) if status == 0 else 0
