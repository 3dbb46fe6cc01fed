"""The allocation-free residual (ctsim_pwls_residual_noalloc) against ctsim_pwls_residual.

Both run the same kernels over the same fixed reduction grid, so the weighted residual
and the FP64 cost must agree bit for bit.  Three calls in a row on one GPU, each smaller
than the one before (the scratch array keeps the previous call's partial sums; the
counter must have returned to zero), with the dtype mixes the PWLS forms use; each is
also checked against an FP64 numpy evaluation of w (p - y) and phi + wscale sum w (p - y)^2 / 2.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    (3_000_001, "f32", "bf16", "bf16", "bf16", 2.0e4),  # mixed PWLS form: FP32 projection, BF16 data/weights/residual
    (100_003, "f32", "f16", "f16", "f16", 1.0),
    (777, "f64", "f64", "f64", "f64", 3.5),              # the FP64 form
]


def _run(entry, p, y, w, r, n, ws, phi, st):
    from ctsim_amd._lib import call
    call(entry, p.vp, p.dtype, y.vp, y.dtype, w.vp, w.dtype, r.vp, r.dtype, n, ws, C.c_void_p(phi.ptr),
         C.c_void_p(st.handle))


def _quantise(a, dt):
    from ctsim_amd.device import DeviceArray
    return DeviceArray.from_numpy(a, dt)


def test_noalloc_residual_equals_residual_on_three_calls():
    from ctsim_amd.device import DeviceArray, Stream, bf16_from_f32, device_count, f32_from_bf16
    assert device_count() >= 1
    st = Stream()
    rng = np.random.default_rng(2024)
    for n, pdt, ydt, wdt, rdt, ws in CASES:
        print(f"case n={n} p={pdt} y={ydt} w={wdt} r={rdt}", flush=True)
        p = rng.normal(1.0, 0.5, n)
        y = p + rng.normal(0.0, 0.05, n)
        w = rng.uniform(0.0, 1.0, n)
        dp, dy, dw = _quantise(p, pdt), _quantise(y, ydt), _quantise(w, wdt)
        r_ref, r_na = DeviceArray(n, rdt).fill(7.0, st), DeviceArray(n, rdt).fill(-3.0, st)
        phi_ref, phi_na = DeviceArray(1, "f64").fill(1.25, st), DeviceArray(1, "f64").fill(1.25, st)
        _run("ctsim_pwls_residual_noalloc", dp, dy, dw, r_na, n, ws, phi_na, st)
        _run("ctsim_pwls_residual", dp, dy, dw, r_ref, n, ws, phi_ref, st)
        st.synchronize()
        a, b = r_ref.to_numpy(), r_na.to_numpy()
        assert np.array_equal(a, b), f"residual differs at {np.flatnonzero(a != b)[:5]}"
        f_ref, f_na = float(phi_ref.to_numpy()[0]), float(phi_na.to_numpy()[0])
        assert f_na == f_ref, (f_na, f_ref)
        # against FP64 numpy on the stored (rounded) operands
        hp, hy, hw = dp.to_numpy().astype(np.float64), dy.to_numpy().astype(np.float64), dw.to_numpy().astype(
            np.float64)
        d = hp - hy
        want = 1.25 + 0.5 * ws * np.sum(hw * d * d)
        assert abs(f_na - want) <= 1e-6 * abs(want), (f_na, want)
        # r = w (p - y) in FP32 (FP64 when all four are FP64), rounded to nearest even into
        # its storage dtype: numpy's IEEE arithmetic gives the same words
        if rdt == "f64":
            r_want = hw * (hp - hy)
        else:
            r32 = hw.astype(np.float32) * (hp.astype(np.float32) - hy.astype(np.float32))
            r_want = (f32_from_bf16(bf16_from_f32(r32)) if rdt == "bf16" else r32.astype(np.float16)).astype(
                np.float32)
        bad = np.flatnonzero(b.astype(np.float64) != r_want.astype(np.float64))
        if bad.size:
            print("words one storage ulp apart (device, numpy):",
                  list(zip(b[bad[:4]].tolist(), r_want[bad[:4]].tolist())), flush=True)
        # the device may round w (p - y) to the storage dtype once from the exact product
        # (a fused multiply with half-precision output) where numpy rounds to FP32 first:
        # near-ties then land one storage ulp apart; everything else is exact
        ulp = {"f64": 0.0, "f16": 2.0 ** -10, "bf16": 2.0 ** -7}[rdt]
        assert bad.size <= max(8, n // 1000), f"{bad.size} words differ"
        assert np.all(np.abs(b[bad].astype(np.float64) - r_want[bad].astype(np.float64))
                      <= ulp * np.abs(r_want[bad].astype(np.float64)) + 2.0 ** -24)
