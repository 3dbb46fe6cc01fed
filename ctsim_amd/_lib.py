"""ctypes binding of libctsim_hip.so (the C ABI in include/ctsim_amd.h).

The library is loaded from this package directory.  There is no fallback: when the
shared object is missing or fails to load, every call raises, so a GPU run can never
silently execute anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctsim_hip.so")

F32, F16, BF16, F64 = 0, 1, 2, 3

vp = C.c_void_p
i32 = C.c_int32
f32 = C.c_float
f64 = C.c_double
sz = C.c_size_t
cp = C.c_char_p
P_i32 = C.POINTER(C.c_int32)
P_f32 = C.POINTER(C.c_float)
P_f64 = C.POINTER(C.c_double)
P_u8 = C.POINTER(C.c_uint8)
P_i64 = C.POINTER(C.c_int64)


class CtsimGeom(C.Structure):
    """struct ctsim_geom"""

    _fields_ = [("nx", i32), ("ny", i32), ("nz", i32),
                ("dx", f32), ("dy", f32), ("dz", f32),
                ("ox", f32), ("oy", f32), ("oz", f32),
                ("nu", i32), ("nv", i32),
                ("du", f32), ("dv", f32),
                ("ou", f32), ("ov", f32),
                ("sad", f32), ("add", f32), ("sdd", f32),
                ("detector_type", i32)]


_GATEWAY3 = [cp, P_i32, P_f32, P_f32, P_i32, P_f32, P_f32, f32, f32, f32, i32, P_f64, P_f32, i32, P_f32, P_u8, P_f32,
             C.c_char_p, sz]

_PER_OP = [C.POINTER(CtsimGeom), i32, P_f64, P_f32, P_u8, P_f32, P_f32, C.c_char_p, sz]

SIGNATURES = {
    "ctsim_last_error": (C.c_char_p, []),
    "ctsim_version": (C.c_int, []),
    "ctsim_fbct_geom_mex": (C.c_int, [cp, P_i32, P_f32, P_f32, i32, f32, f32, f32, f32, f32, i32, P_f64, i32, P_f32,
                                      P_f32, C.c_char_p, sz]),
    "ctsim_cbct_geom_mex": (C.c_int, _GATEWAY3),
    "ctsim_cbct_dd_cuda": (C.c_int, _GATEWAY3),
    "ctsim_cbct_sf_cuda": (C.c_int, _GATEWAY3),
    "ctsim_cbct_back_pd_cuda": (C.c_int, _GATEWAY3),
    "ctsim_cbct_geom_pmat_mex": (C.c_int, [C.c_char_p, P_i32, P_f32, P_f32, P_i32, f32, f32, i32, P_f32, P_f32,
                                           P_f32, P_f32, P_f32, P_u8, P_f32, C.c_char_p, sz]),
    "ctsim_fbct_proj_rd": (C.c_int, _PER_OP),
    "ctsim_cbct_proj_rd": (C.c_int, _PER_OP),
    "ctsim_fbct_proj_dd": (C.c_int, _PER_OP),
    "ctsim_fbct_proj_tf": (C.c_int, _PER_OP),
    "ctsim_fbct_back_dd": (C.c_int, _PER_OP),
    "ctsim_fbct_back_tf": (C.c_int, _PER_OP),
    "ctsim_fbct_back_pd": (C.c_int, _PER_OP),
    "ctsim_cbct_proj_dd": (C.c_int, _PER_OP),
    "ctsim_cbct_proj_tf": (C.c_int, _PER_OP),
    "ctsim_cbct_back_dd": (C.c_int, _PER_OP),
    "ctsim_cbct_back_tf": (C.c_int, _PER_OP),
    "ctsim_cbct_back_pd": (C.c_int, _PER_OP),
    "ctsim_cbct_back_hpd": (C.c_int, _PER_OP),
    "ctsim_plan_create": (C.c_int, [C.POINTER(vp), i32, C.POINTER(CtsimGeom), i32, P_f64, P_f32, P_u8, i32]),
    "ctsim_plan_destroy": (C.c_int, [vp]),
    "ctsim_plan_set_schedule": (C.c_int, [vp, i32, i32]),
    "ctsim_plan_get_schedule": (C.c_int, [vp, P_i32, P_i32]),
    "ctsim_plan_autotune": (C.c_int, [vp, C.c_char_p, i32, i32, i32, vp, i32, vp, i32, i32, P_i32, P_f32, P_i32, i32,
                                      vp]),
    "ctsim_plan_proj": (C.c_int, [vp, cp, i32, i32, i32, vp, i32, vp, i32, vp]),
    "ctsim_plan_proj_exhaustive": (C.c_int, [vp, cp, i32, i32, i32, vp, i32, vp, i32, vp]),
    "ctsim_plan_back": (C.c_int, [vp, cp, i32, i32, i32, vp, i32, vp, i32, f64, i32, vp]),
    "ctsim_plan_back_nesterov": (C.c_int, [vp, cp, i32, i32, i32, vp, i32, f64, vp, vp, vp, vp, vp, vp, vp, vp, f64,
                                           f64, i32, f64, f64, vp, vp, i32, vp]),
    "ctsim_xyz_to_zxy": (C.c_int, [vp, i32, vp, i32, i32, i32, i32, vp]),
    "ctsim_zxy_to_xyz": (C.c_int, [vp, i32, vp, i32, i32, i32, i32, vp]),
    "ctsim_batch2_create": (C.c_int, [C.POINTER(vp), i32, C.POINTER(CtsimGeom), P_i32, C.POINTER(P_f64), i32]),
    "ctsim_batch2_destroy": (C.c_int, [vp]),
    "ctsim_batch2_sizes": (C.c_int, [vp, P_i64, P_i64]),
    "ctsim_batch2_proj": (C.c_int, [vp, cp, vp, i32, vp, i32, vp]),
    "ctsim_batch2_back": (C.c_int, [vp, cp, vp, i32, vp, i32, vp]),
    "ctsim_batch2_set_schedule": (C.c_int, [vp, i32, i32]),
    "ctsim_batch2_get_schedule": (C.c_int, [vp, P_i32, P_i32]),
    "ctsim_batch2_autotune": (C.c_int, [vp, cp, cp, vp, i32, vp, i32, vp, i32, P_f32, vp]),
    "ctsim_malloc": (C.c_int, [C.POINTER(vp), sz]),
    "ctsim_free": (C.c_int, [vp]),
    "ctsim_memcpy_h2d": (C.c_int, [vp, vp, sz, vp]),
    "ctsim_memcpy_d2h": (C.c_int, [vp, vp, sz, vp]),
    "ctsim_memcpy_d2d": (C.c_int, [vp, vp, sz, vp]),
    "ctsim_fill": (C.c_int, [vp, i32, sz, f64, vp]),
    "ctsim_convert": (C.c_int, [vp, i32, vp, i32, sz, vp]),
    "ctsim_stream_create": (C.c_int, [C.POINTER(vp)]),
    "ctsim_stream_destroy": (C.c_int, [vp]),
    "ctsim_stream_synchronize": (C.c_int, [vp]),
    "ctsim_device_synchronize": (C.c_int, []),
    "ctsim_event_create": (C.c_int, [C.POINTER(vp)]),
    "ctsim_event_destroy": (C.c_int, [vp]),
    "ctsim_event_record": (C.c_int, [vp, vp]),
    "ctsim_event_elapsed_ms": (C.c_int, [vp, vp, P_f32]),
    "ctsim_graph_begin": (C.c_int, [vp]),
    "ctsim_graph_end": (C.c_int, [vp, C.POINTER(vp)]),
    "ctsim_graph_launch": (C.c_int, [vp, vp]),
    "ctsim_graph_destroy": (C.c_int, [vp]),
    "ctsim_device_count": (C.c_int, [P_i32]),
    "ctsim_pwls_residual": (C.c_int, [vp, i32, vp, i32, vp, i32, vp, i32, sz, f64, vp, vp]),
    "ctsim_pwls_residual_noalloc": (C.c_int, [vp, i32, vp, i32, vp, i32, vp, i32, sz, f64, vp, vp]),
    "ctsim_huber": (C.c_int, [vp, i32, i32, i32, f64, i32, vp, vp, vp]),
    "ctsim_pwls_nesterov_update": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, i32, vp, i32, i32, i32, f64, f64, i32, f64, f64, vp]),
    "ctsim_pwls_nesterov_update_seq": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, i32, vp, i32, i32, i32, f64, f64, i32, vp, vp, i32, vp]),
    "ctsim_add_i32": (C.c_int, [vp, i32, vp]),
    "ctsim_pwls_sqs_update": (C.c_int, [vp, vp, vp, vp, i32, i32, i32, f64, f64, i32, vp]),
    "ctsim_sqdiff_sum": (C.c_int, [vp, vp, sz, vp, vp]),
    "ctsim_huber_f64": (C.c_int, [vp, i32, i32, i32, f64, i32, vp, vp, vp]),
    "ctsim_pwls_nesterov_update_f64": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, f64, f64, i32, f64, f64, vp]),
    "ctsim_sqdiff_sum_f64": (C.c_int, [vp, vp, sz, vp, vp]),
    "ctsim_fdk_create": (C.c_int, [C.POINTER(vp), C.POINTER(CtsimGeom), i32, P_f64, i32, P_f32, i32]),
    "ctsim_fdk_destroy": (C.c_int, [vp]),
    "ctsim_fbp2d_create": (C.c_int, [C.POINTER(vp), C.POINTER(CtsimGeom), i32, P_f64, i32, P_f32, i32]),
    "ctsim_fdk_filter": (C.c_int, [vp, i32, i32, vp, i32, vp, i32, vp]),
    "ctsim_poisson_counts": (C.c_int, [vp, i32, sz, f64, C.c_uint64, vp, vp]),
    "ctsim_log_attenuation": (C.c_int, [vp, sz, f64, f64, vp, i32, vp]),
    "ctsim_log_attenuation_ex": (C.c_int, [vp, sz, i32, f64, f64, i32, vp, vp, vp, i32, vp]),
    "ctsim_water_correction": (C.c_int, [vp, i32, vp, i32, i32, i32, i32, vp, i32, i32, vp]),
    "ctsim_pwls_weights": (C.c_int, [vp, sz, f64, f64, vp, i32, vp]),
    "ctsim_penalty": (C.c_int, [vp, i32, i32, i32, i32, f64, i32, vp, vp, vp]),
    "ctsim_penalty_f64": (C.c_int, [vp, i32, i32, i32, i32, f64, i32, vp, vp, vp]),
    "ctsim_pwls_lalm_sg": (C.c_int, [vp, vp, vp, sz, f64, vp]),
    "ctsim_pwls_lalm_inner": (C.c_int, [vp, vp, vp, vp, vp, i32, i32, i32, f64, f64, i32, f64, vp]),
    "ctsim_mul": (C.c_int, [vp, i32, vp, i32, vp, i32, sz, vp]),
    "ctsim_sub": (C.c_int, [vp, vp, vp, i32, sz, vp]),
    "ctsim_nusqs_seed": (C.c_int, [vp, vp, i32, i32, i32, f64, vp]),
    "ctsim_div": (C.c_int, [vp, vp, vp, sz, vp]),
    "ctsim_absdiff_floor": (C.c_int, [vp, vp, vp, sz, f64, vp]),
    "ctsim_dynamic_range": (C.c_int, [vp, sz, f64, f64, vp, vp]),
    "ctsim_pwls_nusqs_update": (C.c_int, [vp, vp, vp, vp, vp, i32, i32, i32, f64, f64, i32, vp]),
}


class CtsimError(RuntimeError):
    """A nonzero status from the C ABI, with the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[ctsim status {status}] {message}")
        self.status = status


_lib = None


def lib():
    """Load the shared library (once) and declare every signature."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or python -m ctsim_amd.build; there is no non-HIP fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise CtsimError(status, lib().ctsim_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
