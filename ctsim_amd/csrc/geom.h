// Device-side geometry of the CTSim projector family, shared by every kernel.
//
// The arithmetic restates the Linux CPU MEX kernels (geom/fbct_os/fbct_proj.c:174-260,
// geom/cbct_os/cbct_proj.c:454-623) so that a forward projector and its matched
// backprojector evaluate the very same footprint function: the pair is an exact
// transpose by construction.  T is the compute type (float, or double for the
// fp64 path); storage types are converted on load.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctsim {

// ---------------------------------------------------------------------------
// storage conversions
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ float ld_f(const T* p) { return (float)(*p); }
template <>
__device__ __forceinline__ float ld_f<__half>(const __half* p) { return __half2float(*p); }
template <>
__device__ __forceinline__ float ld_f<__hip_bfloat16>(const __hip_bfloat16* p) { return __bfloat162float(*p); }

template <class C, class S>
__device__ __forceinline__ C ld(const S* p) { return (C)ld_f(p); }

// value conversions (no address taken, so values stay in registers)
__device__ __forceinline__ float to_f(float x) { return x; }
__device__ __forceinline__ float to_f(__half x) { return __half2float(x); }
__device__ __forceinline__ float to_f(__hip_bfloat16 x) { return __bfloat162float(x); }
__device__ __forceinline__ double to_f(double x) { return x; }
template <>
__device__ __forceinline__ double ld<double, double>(const double* p) { return *p; }

// Range-checked loads through a buffer resource (base in scalar registers, byte offsets
// per lane): an access at or past nbytes from the base reads zero instead of memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, int nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, nbytes, 0x00020000);
}
template <class S>
__device__ __forceinline__ S buf_ld(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ __forceinline__ float buf_ld<float>(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ __forceinline__ double buf_ld<double>(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <>
__device__ __forceinline__ __half buf_ld<__half>(__amdgpu_buffer_rsrc_t r, int off) {
    return __ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
}
template <>
__device__ __forceinline__ __hip_bfloat16 buf_ld<__hip_bfloat16>(__amdgpu_buffer_rsrc_t r, int off) {
    const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
    return __builtin_bit_cast(__hip_bfloat16, b);
}

// clamp to [0, 1] (one v_med3 in FP32)
__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.f, 1.f); }
__device__ __forceinline__ double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

template <class S, class C>
__device__ __forceinline__ S cvt_store(C x) { return (S)x; }
template <>
__device__ __forceinline__ __half cvt_store<__half, float>(float x) { return __float2half_rn(x); }
template <>
__device__ __forceinline__ __hip_bfloat16 cvt_store<__hip_bfloat16, float>(float x) { return __float2bfloat16(x); }
template <>
__device__ __forceinline__ __half cvt_store<__half, double>(double x) { return __float2half_rn((float)x); }
template <>
__device__ __forceinline__ __hip_bfloat16 cvt_store<__hip_bfloat16, double>(double x) { return __float2bfloat16((float)x); }

// ---------------------------------------------------------------------------
// geometry records
// ---------------------------------------------------------------------------

// Per-view record.  sin/cos are evaluated in double on the host and rounded to T,
// as the reference does ((float) sin(beta)); the half-pixel projections follow
// geom/cbct_os/cbct_proj.c:471-474 and wz folds in the couch position (:465).
template <class T>
struct View {
    T s, c, dxs2, dxc2, dys2, ndyc2, wz, pad;
};

// Scalar geometry of one operator.  w* are the MATLAB-centred index offsets
// (-(n-1)/2 + offset, in cell units), sdd_du = SDD/du and zk = (dz/dv)*SDD so that
// zinc = zk / dsxy0 rounds exactly like the reference's dz / dv * sdd / dsxy0.
template <class T>
struct Geom {
    int nx, ny, nz, nu, nv, dettype;
    T dx, dy, dz, du, dv;
    T sad, sdd, sdd_du, zk;
    T wx, wy, wu, wv;
    T ou;  // arc pd offset term uses offset_u directly (geom/cbct_os/cbct_back.c:204)
    // backprojection of a z-slab of a taller volume: the column stride of the output
    // buffer (nz of the whole volume) and the slab's first slice, added to each view's wz;
    // zstride = nz and zoff = 0 for a whole volume
    int zstride;
    T zoff;
    // launch schedule (host side only; set by ctsim_plan_set_schedule / the autotuner):
    // sched_proj = form | split << 2 with form 0 auto, 1 LDS ring, 2 register pipeline and
    // split 0 auto, 1 one wave per work item, 2 a block's waves share an item;
    // sched_back bit 0: 0 consecutive columns per block, 1 2 x 2 column patches in 32 x 32
    // super-tiles (DD / SF); bit 1: the FDK (pd) backprojection's direct form instead of the
    // LDS-staged one
    int sched_proj;
    int sched_back;
};

// In-plane column quantities (cbct_proj.c:491-533).
template <class T>
struct Column {
    T t, d, r, amp;
    bool xdom;
};

template <class T>
__device__ __forceinline__ T fabs_t(T x) { return x < T(0) ? -x : x; }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }

// Division and square root of the footprint arithmetic.  The FP32 path uses the
// hardware's 1-ulp v_rcp_f32 / v_sqrt_f32: every weight is a continuous function of
// the quantities they produce, so the result moves by O(1e-7) relative, far inside the
// 1e-5 parity budget.  The FP64 path stays IEEE-exact.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double frcp(double b) { return 1.0 / b; }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double fsqrt(double a) { return sqrt(a); }

template <class T>
__device__ __forceinline__ Column<T> column_geom(T xc, T yc, const View<T>& v, T sad, T dx) {
    Column<T> k;
    k.t = xc * v.c - yc * v.s;
    k.d = sad + xc * v.s + yc * v.c;
    k.r = fsqrt(k.t * k.t + k.d * k.d);
    const T ir = frcp(k.r);
    T cg = k.d * ir;
    T sg = k.t * ir;
    T cphi = fabs_t(v.c * cg - v.s * sg);
    T sphi = fabs_t(v.s * cg + v.c * sg);
    k.xdom = cphi > sphi;
    k.amp = fdiv(dx, k.xdom ? cphi : sphi) * T(0.1);
    return k;
}

template <class T>
__device__ __forceinline__ T det_map(T num, T den, T sdd_du, T wu, int dettype) {
    return dettype == 1 ? sdd_du * fdiv(num, den) - wu : sdd_du * atan(fdiv(num, den)) - wu;
}

// Distance-driven boundaries (unshifted), sorted (fbct_proj.c:216-248).
template <class T>
__device__ __forceinline__ void dd_bounds(const Column<T>& k, const View<T>& v, const Geom<T>& g, T& umin,
                                          T& umax) {
    T a0, b0, a1, b1;
    if (k.xdom) {
        a0 = k.t + v.dxc2; b0 = k.d + v.dxs2;
        a1 = k.t - v.dxc2; b1 = k.d - v.dxs2;
    } else {
        a0 = k.t + v.dys2; b0 = k.d + v.ndyc2;
        a1 = k.t - v.dys2; b1 = k.d - v.ndyc2;
    }
    umin = det_map(a0, b0, g.sdd_du, g.wu, g.dettype);
    umax = det_map(a1, b1, g.sdd_du, g.wu, g.dettype);
    if (umin >= umax) {
        T tmp = umin; umin = umax; umax = tmp;
    }
}

template <class T>
__device__ __forceinline__ void cswap(T& a, T& b) {
    if (a > b) {
        T t = a; a = b; b = t;
    }
}

// Separable-footprint trapezoid corners, sorted by the reference's five
// compare-exchanges (fbct_proj.c:365-412).
template <class T>
__device__ __forceinline__ void sf_taus(const Column<T>& k, const View<T>& v, const Geom<T>& g, T tau[4]) {
    tau[0] = det_map(k.t + v.dxc2 + v.dys2, k.d + v.dxs2 + v.ndyc2, g.sdd_du, g.wu, g.dettype);
    tau[1] = det_map(k.t - v.dxc2 + v.dys2, k.d - v.dxs2 + v.ndyc2, g.sdd_du, g.wu, g.dettype);
    tau[2] = det_map(k.t + v.dxc2 - v.dys2, k.d + v.dxs2 - v.ndyc2, g.sdd_du, g.wu, g.dettype);
    tau[3] = det_map(k.t - v.dxc2 - v.dys2, k.d - v.dxs2 - v.ndyc2, g.sdd_du, g.wu, g.dettype);
    cswap(tau[0], tau[1]);
    cswap(tau[2], tau[3]);
    cswap(tau[0], tau[2]);
    cswap(tau[1], tau[3]);
    cswap(tau[1], tau[2]);
}

// Trapezoid integral over cell [iu - 1/2, iu + 1/2] (fbct_proj.c:428-453).
template <class T>
__device__ __forceinline__ T sf_weight(int iu, T t0, T t1, T t2, T t3) {
    T w = T(0);
    const T bl = T(iu) - T(0.5), br = T(iu) + T(0.5);
    T xl = bl > t0 ? bl : t0, xr = br < t1 ? br : t1;
    if (xr > xl) w += fdiv((xr - t0) * (xr - t0) - (xl - t0) * (xl - t0), t1 - t0) * T(0.5);
    xl = bl > t1 ? bl : t1;
    xr = br < t2 ? br : t2;
    if (xr > xl) w += (xr - xl);
    xl = bl > t2 ? bl : t2;
    xr = br < t3 ? br : t3;
    if (xr > xl) w += fdiv((xl - t3) * (xl - t3) - (xr - t3) * (xr - t3), t3 - t2) * T(0.5);
    return w;
}

// Clamp of an unshifted [lo, hi] footprint to a detector of n cells after the
// +0.5 edge shift; cap > 0 limits the footprint to cap cells (3D: 16).
// Returns false when the footprint misses the detector.
template <class T>
__device__ __forceinline__ bool clamp_cells(T lo, T hi, int n, int cap, T& clo, T& chi, int& ilo, int& ihi) {
    T a = lo + T(0.5), b = hi + T(0.5);
    clo = a > T(0) ? a : T(0);
    chi = b < T(n) ? b : T(n);
    if (clo > chi) return false;
    ilo = (int)floor(clo);
    ihi = (int)floor(chi);
    ihi = ihi < n - 1 ? ihi : n - 1;
    if (cap > 0) ihi = ihi < ilo + cap - 1 ? ihi : ilo + cap - 1;
    return true;
}

// DD box weight of cell i in the clamped footprint [clo, chi] covering cells
// ilo..ihi (cbct_proj.c:553-567); zero outside.
template <class T>
__device__ __forceinline__ T dd_weight(int i, T clo, T chi, int ilo, int ihi) {
    if (i < ilo || i > ihi) return T(0);
    if (ilo == ihi) return chi - clo;
    if (i == ilo) return T(ilo + 1) - clo;
    if (i == ihi) return chi - T(ihi);
    return T(1);
}

// Branch-free DD weight of detector cell i for the unshifted footprint [lo, hi] on a
// detector of n cells with the 16-cell cap: the reference's first / middle / last /
// single-cell rules (cbct_proj.c:597-623) are all the overlap min(chi, i+1) - max(clo, i)
// of the clamped interval (same float operations); beyond the cap the cells get nothing
// and the capped last cell gets chi - i.  Returns 0 for an empty footprint.
template <class T>
__device__ __forceinline__ T dd_overlap16(int i, T lo, T hi, int n) {
    const T a = lo + T(0.5), b = hi + T(0.5);
    const T clo = a > T(0) ? a : T(0);
    const T chi = b < T(n) ? b : T(n);
    const T fi = T(i);
    T w = (chi < fi + T(1) ? chi : fi + T(1)) - (clo > fi ? clo : fi);
    w = w > T(0) ? w : T(0);
    const T capc = floor(clo) + T(15);
    w = fi > capc ? T(0) : (fi == capc ? chi - fi : w);
    return clo > chi ? T(0) : w;
}

// dd_overlap16 for footprints at most 14 cells wide: cell i in [0, n) lies inside the
// clamp range [0, n], so clamping cannot change the overlap, and a footprint narrower
// than 15 cells never reaches the cap (its last cell, floor(chi), is at most
// floor(clo) + 15, where the capped rule chi - i equals the plain overlap).
template <class T>
__device__ __forceinline__ T dd_overlap(int i, T lo, T hi) {
    const T a = lo + T(0.5), b = hi + T(0.5);
    const T fi = T(i);
    const T w = (b < fi + T(1) ? b : fi + T(1)) - (a > fi ? a : fi);
    return w > T(0) ? w : T(0);
}

// v window of a column (cbct_proj.c:569-580): magnified slice pitch zinc,
// first voxel edge vstart and the voxel range [iz0, iz1).
template <class T>
__device__ __forceinline__ bool v_window(const Geom<T>& g, T d, T wz, T& zinc, T& vstart, int& iz0, int& iz1) {
    zinc = fdiv(g.zk, d);
    vstart = (T)(((double)wz - 0.5) * (double)zinc - (double)g.wv);
    const T rz = frcp(zinc);
    iz0 = (int)floor((T(-1) - vstart) * rz);
    iz1 = (int)ceil((T(g.nv + 1) - vstart) * rz);
    iz0 = iz0 > 0 ? iz0 : 0;
    iz1 = iz1 < g.nz ? iz1 : g.nz;
    return iz0 <= iz1;
}

// ---------------------------------------------------------------------------
// strip of image columns that can reach a detector tile (forward gathers)
// ---------------------------------------------------------------------------

// minor-index candidate range of major line m for the tile whose unshifted u
// coordinates span [uL, uR] (cell units).  Returns false for "no candidates".
template <class T>
__device__ __forceinline__ bool strip_range(const Geom<T>& g, const View<T>& v, bool row_major, int m, T kL,
                                            T kR, bool exhaustive, int& lo, int& hi) {
    const int nminor = row_major ? g.nx : g.ny;
    if (exhaustive) {
        lo = 0;
        hi = nminor - 1;
        return true;
    }
    T x0, x1, mg;
    if (row_major) {
        // row y = y0: x = (k (sad + y0 c) + y0 s) / (c - k s)
        const T y0 = (T(m) + g.wy) * g.dy;
        const T d0 = v.c - kL * v.s, d1 = v.c - kR * v.s;
        if (d0 * v.c <= T(0) || d1 * v.c <= T(0)) {
            lo = 0; hi = nminor - 1;
            return true;
        }
        x0 = (kL * (g.sad + y0 * v.c) + y0 * v.s) / d0;
        x1 = (kR * (g.sad + y0 * v.c) + y0 * v.s) / d1;
        x0 = x0 / g.dx - g.wx;
        x1 = x1 / g.dx - g.wx;
        mg = T(2) + g.dy / g.dx;
    } else {
        // column x = x0: y = (x0 c - k (sad + x0 s)) / (s + k c)
        const T xx = (T(m) + g.wx) * g.dx;
        const T d0 = v.s + kL * v.c, d1 = v.s + kR * v.c;
        if (d0 * v.s <= T(0) || d1 * v.s <= T(0)) {
            lo = 0; hi = nminor - 1;
            return true;
        }
        x0 = (xx * v.c - kL * (g.sad + xx * v.s)) / d0;
        x1 = (xx * v.c - kR * (g.sad + xx * v.s)) / d1;
        x0 = x0 / g.dy - g.wy;
        x1 = x1 / g.dy - g.wy;
        mg = T(2) + g.dx / g.dy;
    }
    T a = x0 < x1 ? x0 : x1, b = x0 < x1 ? x1 : x0;
    a -= mg;
    b += mg;
    if (!(b >= T(0)) || !(a <= T(nminor - 1))) return false;  // also rejects NaN
    lo = a < T(0) ? 0 : (int)floor(a);
    hi = b > T(nminor - 1) ? nminor - 1 : (int)ceil(b);
    return lo <= hi;
}

// detector u (unshifted cell coordinate) -> ray slope k = t / D
template <class T>
__device__ __forceinline__ T u_to_slope(const Geom<T>& g, T u) {
    const T um = (u + g.wu) * g.du;
    return g.dettype == 1 ? um / (g.sdd) : tan(um / g.sdd);
}

// ---------------------------------------------------------------------------
// wave helpers (64-wide wavefronts)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int rdl_i(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ __forceinline__ float rdl_f(float x, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
__device__ __forceinline__ double rdl_d(double x, int lane) {
    long long b = __double_as_longlong(x);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float rdl(float x, int lane) { return rdl_f(x, lane); }
__device__ __forceinline__ double rdl(double x, int lane) { return rdl_d(x, lane); }

// a value every lane of the wave holds (e.g. read from one LDS address), moved to a
// scalar register so that comparisons on it become uniform branches
// Tile of rank r in the order c, c+1, c-1, c+2, c-2, ... (then the rest of the longer side)
// for nut tiles around the tile c of the central ray.
__device__ __forceinline__ int tile_by_rank(int r, int c, int nut) {
    const int left = c, right = nut - 1 - c, m = min(left, right);
    if (r == 0) return c;
    if (r <= 2 * m) {
        const int k = (r + 1) >> 1;
        return (r & 1) ? c + k : c - k;
    }
    const int extra = r - 2 * m;
    return right > left ? c + m + extra : c - m - extra;
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
__device__ __forceinline__ double uni(double x) {
    long long b = __double_as_longlong(x);
    int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// inclusive prefix sum across the 64 lanes
__device__ __forceinline__ int wave_incl_scan(int x) {
    const int l = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (l >= off) x += y;
    }
    return x;
}

// XCD-aware block remap: the dispatcher places block i on XCD (i mod 8); give each
// XCD a contiguous range of logical blocks so that neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int nxcd = 8;
    const int per = nblocks / nxcd;
    const int rem = nblocks % nxcd;
    const int xcd = bid % nxcd;
    const int idx = bid / nxcd;
    // XCDs below rem own per+1 logical blocks
    const int base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + idx;
}

}  // namespace ctsim
