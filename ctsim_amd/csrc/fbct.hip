// 2D fan-beam projector family for gfx950: distance-driven and separable-footprint
// forward projection (geom/fbct_os/fbct_proj.c fbct_proj_dd :155-291, fbct_proj_tf
// :294-466), their matched backprojections (geom/fbct_os/fbct_back.c fbct_back_dd
// :82-225, fbct_back_tf :227-399) and the pixel-driven FBP backprojection
// (fbct_back_pd :3-80).  2D footprints are not capped (the 3D code caps at 16 cells).
//
// Layouts as in MATLAB: image [nx ny] (x fastest), sinogram [nu nviews].
// Batched forms take a table of independent slices, each with its own geometry,
// and walk a flat work list of (slice, view, tile) or (slice, pixel block) items,
// so slices of uneven size share one launch.
#include "internal.h"

namespace ctsim {

constexpr int kW2 = 64;
// image lines whose strip holds at least this many pixels on average take a batch each in
// the forward projection (fewer: several lines share a batch)
constexpr int kLineBatch = 12;

enum Foot2 { F2_DD = 0, F2_TF = 1 };

// ---------------------------------------------------------------------------------
// forward: one wavefront per (view, 64 detector cells); lanes own cells.  Candidate
// pixels of the strip are evaluated 64 at a time (one per lane) into LDS entries, and
// every tile cell notes the first and the last entry whose footprint covers it (LDS
// min / max, order-free).  Each lane then walks the entries of its own cell range in
// candidate order, so a cell adds the same products in the same order as a walk over
// all candidates would, and a pixel costs the few lanes it covers, not the whole wave.
// ---------------------------------------------------------------------------------
// share `share` of `nshares` of the strip (its rows or columns along the major axis):
// returns this lane's cell sum.  Called by whole 256-thread blocks (four waves).
template <class Tin, class T, int FOOT>
__device__ __forceinline__ T fbct_proj_acc(const Geom<T>& g, const View<T>& v, int ut, const Tin* __restrict__ img,
                                           int exhaustive, int share, int nshares) {
    __shared__ T ep[4][FOOT == F2_DD ? 2 : 4][kW2];  // footprint parameters of the entries
    __shared__ T eval_[4][kW2];                        // amp * pixel value
    __shared__ int ecell[4][kW2];                      // ilo | ihi << 16 (empty: ilo 1 > ihi 0)
    __shared__ int cfirst[4][kW2], clast[4][kW2];      // per tile cell: first and last entry
    const int wv = (int)(threadIdx.x / kW2);
    const int lane = lane_id();
    const int iu_a = ut * kW2;
    const int iu_b = min(iu_a + kW2, g.nu) - 1;
    const int iu = iu_a + lane;
    const bool row_major = fabs_t(v.c) >= fabs_t(v.s);
    const int nmajor = row_major ? g.ny : g.nx;
    const int m_begin = nmajor * share / nshares, m_end = nmajor * (share + 1) / nshares;
    const T kL = u_to_slope(g, T(iu_a) - T(0.5));
    const T kR = u_to_slope(g, T(iu_b) + T(0.5));
    T acc = T(0);
    for (int m0 = m_begin; m0 < m_end; m0 += kW2) {
        const int m = m0 + lane;
        int lo = 0, hi = -1;
        if (m < m_end) {
            if (!strip_range(g, v, row_major, m, kL, kR, exhaustive != 0, lo, hi)) hi = lo - 1;
        }
        const int cnt = hi - lo + 1 > 0 ? hi - lo + 1 : 0;
        const int incl = wave_incl_scan(cnt);
        const int excl = incl - cnt;
        const int total = rdl_i(incl, kW2 - 1);
        // one batch of 64 candidates: entries in LDS, the covering range of every tile cell,
        // then each lane's walk over that range
        auto batch = [&](int major, int minor, bool ok) {
            const int ix = row_major ? minor : major;
            const int iy = row_major ? major : minor;
            int ilo = 0, ihi = -1;
            T p0 = T(0), p1 = T(0), p2 = T(0), p3 = T(0), val = T(0);
            // the pixel is loaded before its footprint is known (every candidate lies inside
            // the image), so that the load's latency overlaps the footprint arithmetic
            T pix = T(0);
            if (ok) pix = ld<T>(img + (size_t)iy * g.nx + ix);
            if (ok) {
                const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
                const Column<T> cg = column_geom(xc, yc, v, g.sad, g.dx);
                T clo, chi;
                if (FOOT == F2_DD) {
                    T umin, umax;
                    dd_bounds(cg, v, g, umin, umax);
                    ok = clamp_cells(umin, umax, g.nu, 0, clo, chi, ilo, ihi);
                    p0 = clo;
                    p1 = chi;
                } else {
                    T tau[4];
                    sf_taus(cg, v, g, tau);
                    ok = clamp_cells(tau[0], tau[3], g.nu, 0, clo, chi, ilo, ihi);
                    p0 = tau[0]; p1 = tau[1]; p2 = tau[2]; p3 = tau[3];
                }
                ok = ok && ihi >= iu_a && ilo <= iu_b;
                if (ok) val = cg.amp * pix;
            }
            ecell[wv][lane] = ok ? (ilo | (ihi << 16)) : 1;
            ep[wv][0][lane] = p0;
            ep[wv][1][lane] = p1;
            if constexpr (FOOT == F2_TF) {
                ep[wv][2][lane] = p2;
                ep[wv][3][lane] = p3;
            }
            eval_[wv][lane] = val;
            cfirst[wv][lane] = kW2;
            clast[wv][lane] = -1;
            __builtin_amdgcn_wave_barrier();
            if (ok) {
                const int c1 = min(ihi, iu_b) - iu_a;
                for (int c = max(ilo, iu_a) - iu_a; c <= c1; ++c) {
                    atomicMin(&cfirst[wv][c], lane);
                    atomicMax(&clast[wv][c], lane);
                }
            }
            __builtin_amdgcn_wave_barrier();
            const int jl = clast[wv][lane];
            for (int e = cfirst[wv][lane]; e <= jl; ++e) {
                const int cl = ecell[wv][e];
                const int cilo = cl & 0xffff, cihi = cl >> 16;
                if (iu >= cilo && iu <= cihi) {
                    const T cp0 = ep[wv][0][e], cp1 = ep[wv][1][e];
                    T w;
                    if constexpr (FOOT == F2_DD) {
                        w = dd_weight(iu, cp0, cp1, cilo, cihi);
                    } else {
                        w = sf_weight(iu, cp0, cp1, ep[wv][2][e], ep[wv][3][e]);
                    }
                    acc += w * eval_[wv][e];
                }
            }
            __builtin_amdgcn_wave_barrier();
        };
        const unsigned long long nonempty = __ballot(cnt > 0);
        const int nlines = __popcll(nonempty);
        if (total >= kLineBatch * nlines) {
            // wide strips: one image line per batch.  The pixels of a line project in order,
            // so the ones covering a cell are consecutive entries and each lane's walk is as
            // long as its cell's coverage, where a batch spanning several lines walks the
            // entries of every line in between
            unsigned long long left = nonempty;
            while (left) {
                const int j = __ffsll((long long)left) - 1;
                left &= left - 1;
                const int lo_j = rdl_i(lo, j), cnt_j = rdl_i(cnt, j);
                for (int k0 = 0; k0 < cnt_j; k0 += kW2) batch(m0 + j, lo_j + k0 + lane, k0 + lane < cnt_j);
            }
        } else {
            // narrow strips: consecutive candidates of several lines share a batch
            for (int k0 = 0; k0 < total; k0 += kW2) {
                const int k = k0 + lane;
                int j = 0;
#pragma unroll
                for (int step = 32; step >= 1; step >>= 1) {
                    const int e = __shfl(excl, j + step, 64);
                    if (e <= k) j += step;
                }
                const int lo_j = __shfl(lo, j, 64);
                const int ex_j = __shfl(excl, j, 64);
                batch(m0 + j, lo_j + (k - ex_j), k < total);
            }
        }
    }
    return acc;
}

template <class T, int S>
__device__ __forceinline__ T combine_shares(T (*part)[kW2], int wv, int lane) {
    const int w0 = wv - wv % S;
    if (S == 1) return part[w0][lane];
    if (S == 2) return part[w0][lane] + part[w0 + 1][lane];
    return (part[w0][lane] + part[w0 + 1][lane]) + (part[w0 + 2][lane] + part[w0 + 3][lane]);
}

// S waves share a (view, tile) item, each walking 1/S of the strip's lines; the partial
// sums are added in a fixed order.  S = 4 when the items alone would leave the chip
// with few waves.
template <class Tin, class Tout, class T, int FOOT, int S>
__global__ __launch_bounds__(256) void fbct_proj_kernel(Geom<T> g, const View<T>* __restrict__ views, int vstride,
                                                        int nviews, const Tin* __restrict__ img,
                                                        Tout* __restrict__ proj, int exhaustive) {
    __shared__ T part[4][kW2];
    const int nut = (g.nu + kW2 - 1) / kW2;
    const int wv = (int)(threadIdx.x / kW2), lane = lane_id();
    const int wid = blockIdx.x * (4 / S) + wv / S;
    const bool valid = wid < nviews * nut;
    // longest work first: every view's tile of the central ray, then the tiles beside it
    const int ut = valid ? tile_by_rank(wid / nviews, min(max((int)floor(-g.wu / T(kW2)), 0), nut - 1), nut) : 0;
    const int b = valid ? wid % nviews : 0;
    T acc = T(0);
    if (valid) {
        // the view in registers (a wave-uniform index: scalar loads, once per item), so that
        // the walk's LDS atomics and barriers do not make the compiler reload it every batch
        const View<T> vw = views[(size_t)uni(b) * vstride];
        acc = fbct_proj_acc<Tin, T, FOOT>(g, vw, ut, img, exhaustive, wv % S, S);
    }
    if (S > 1) {
        part[wv][lane] = acc;
        __syncthreads();
        if (wv % S) return;
        acc = combine_shares<T, S>(part, wv, lane);
    }
    const int iu = ut * kW2 + lane;
    if (valid && iu < g.nu) proj[(size_t)b * g.nu + iu] = cvt_store<Tout>(acc);
}

// ---------------------------------------------------------------------------------
// matched backprojection (DD / SF) and pixel-driven backprojection: lanes own 64
// consecutive pixels of a row; the 4 waves of a block split the views into four
// contiguous ranges and their partial sums are added in a fixed order.
// ---------------------------------------------------------------------------------
template <class Tin, class T, int KIND>  // KIND 0 dd, 1 tf, 2 pd
__device__ __forceinline__ T fbct_back_pixel(const Geom<T>& g, const View<T>* __restrict__ views, int vstride, int b0,
                                             int b1, int ix, int iy, const Tin* __restrict__ proj) {
    T img = T(0);
    if (ix >= g.nx) return img;
    const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
    for (int b = b0; b < b1; ++b) {
        const View<T> v = views[(size_t)b * vstride];
        const Tin* row = proj + (size_t)b * g.nu;
        if (KIND == 2) {
            const T xr = xc * v.c - yc * v.s;
            const T yr = xc * v.s + yc * v.c;
            T u, w2s;
            if (g.dettype == 1) {
                const T w2 = g.sdd / (g.sad + yr);
                w2s = w2 * w2;
                u = xr * w2 / g.du - g.wu;
            } else {
                const T roff = g.ou * g.sad / g.sdd;
                u = atan((xr - roff) / (g.sad + yr)) * g.sdd / g.du - g.wu;
                w2s = g.sdd * g.sdd / ((g.sad + yr) * (g.sad + yr) + (xr - roff) * (xr - roff));
            }
            if (!((double)u >= 1e-6 && (double)u <= (double)(g.nu - 1) - 1e-6)) continue;  // also skips NaN
            const int iu = (int)floor(u);
            const T wl = u - T(iu);
            img += w2s * ((T(1) - wl) * ld<T>(row + iu) + wl * ld<T>(row + iu + 1));
            continue;
        }
        const Column<T> cg = column_geom(xc, yc, v, g.sad, g.dx);
        T clo, chi, p0, p1, p2 = T(0), p3 = T(0);
        int ilo, ihi;
        bool ok;
        if (KIND == 0) {
            T umin, umax;
            dd_bounds(cg, v, g, umin, umax);
            ok = clamp_cells(umin, umax, g.nu, 0, clo, chi, ilo, ihi);
            p0 = clo;
            p1 = chi;
        } else {
            T tau[4];
            sf_taus(cg, v, g, tau);
            ok = clamp_cells(tau[0], tau[3], g.nu, 0, clo, chi, ilo, ihi);
            p0 = tau[0]; p1 = tau[1]; p2 = tau[2]; p3 = tau[3];
        }
        if (!ok) continue;
        T s = T(0);
        for (int iu = ilo; iu <= ihi; ++iu) {
            const T w = KIND == 0 ? dd_weight(iu, p0, p1, ilo, ihi) : sf_weight(iu, p0, p1, p2, p3);
            s += w * ld<T>(row + iu);
        }
        img += cg.amp * s;
    }
    return img;
}

template <class Tin, class Tout, class T, int KIND>
__global__ __launch_bounds__(256) void fbct_back_kernel(Geom<T> g, const View<T>* __restrict__ views, int vstride,
                                                        int nviews, const Tin* __restrict__ proj,
                                                        Tout* __restrict__ img, T scale) {
    __shared__ T part[3][kW2];
    const int nxb = (g.nx + kW2 - 1) / kW2;
    const int iy = blockIdx.x / nxb;
    const int ix = (blockIdx.x % nxb) * kW2 + lane_id();
    const int w = threadIdx.x / kW2;
    const int per = (nviews + 3) / 4;
    const int b0 = min(w * per, nviews), b1 = min(b0 + per, nviews);
    const T s = fbct_back_pixel<Tin, T, KIND>(g, views, vstride, b0, b1, ix, iy, proj);
    if (w > 0) part[w - 1][lane_id()] = s;
    __syncthreads();
    if (w == 0 && ix < g.nx) {
        const T tot = (s + part[0][lane_id()]) + (part[1][lane_id()] + part[2][lane_id()]);
        img[(size_t)iy * g.nx + ix] = cvt_store<Tout>(scale * tot);
    }
}

// ---------------------------------------------------------------------------------
// Batched uneven slices in one launch.  items are (slice, view, u-tile) for the
// forward and (slice, row, 64-pixel block) for the back pass; a wave finds its slice
// by binary search over the slices' first item.  One launch covers the whole batch.
// ---------------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ int find_slice(const Slice2<T>* __restrict__ sl, int nslices, int item) {
    int lo = 0, hi = nslices - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sl[mid].item0 <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Schedules: the forward pass splits each (slice, view, tile) item's strip over S waves
// of a block (S = 1, 2, 4; the shares are added in a fixed order); the back pass splits
// each (slice, row, 64-pixel) item's views over VS waves (VS = 1, 2, 4).  A block is
// always 4 waves and holds 4/S (4/VS) items.
template <class Tin, class Tout, class T, int FOOT, int S>
__global__ __launch_bounds__(256) void fbct_proj_batch_kernel(const Slice2<T>* __restrict__ sl, int nslices,
                                                              int nitems, const Tin* __restrict__ img,
                                                              Tout* __restrict__ proj) {
    __shared__ T part[4][kW2];
    // blocks in dispatch order (consecutive blocks on different XCDs): a slice's items, longest
    // first, spread over the chip (a contiguous range of the list per XCD, which would keep a
    // slice's 0.5 MB image in one L2, measured 18% slower: it leaves the XCDs unevenly loaded)
    const int bid = blockIdx.x;
    const int wv = (int)(threadIdx.x / kW2), lane = lane_id();
    const int item = bid * (4 / S) + wv / S;
    const bool valid = item < nitems;
    T acc = T(0);
    int si = 0, ut = 0, b = 0;
    if (valid) {
        // the item's slice geometry and view in registers (wave-uniform indices: scalar loads,
        // once per item), so that the walk's LDS atomics and barriers do not make the
        // compiler reload them from the slice table every batch
        si = uni(find_slice(sl, nslices, item));
        const Slice2<T>& s = sl[si];
        const Geom<T> g = s.g;
        const int local = item - s.item0;
        const int nut = (g.nu + kW2 - 1) / kW2;
        // within the slice, longest work first (the tiles around the central ray's)
        ut = uni(tile_by_rank(local / s.nviews, min(max((int)floor(-g.wu / T(kW2)), 0), nut - 1), nut));
        b = uni(local % s.nviews);
        const View<T> vw = s.views[b];
        acc = fbct_proj_acc<Tin, T, FOOT>(g, vw, ut, img + s.img_off, 0, wv % S, S);
    }
    if (S > 1) {
        part[wv][lane] = acc;
        __syncthreads();
        if (wv % S) return;
        acc = combine_shares<T, S>(part, wv, lane);
    }
    if (!valid) return;
    const Slice2<T>& s = sl[si];
    const int iu = ut * kW2 + lane;
    if (iu < s.g.nu) proj[s.sino_off + (size_t)b * s.g.nu + iu] = cvt_store<Tout>(acc);
}

template <class Tin, class Tout, class T, int KIND, int VS>
__global__ __launch_bounds__(256) void fbct_back_batch_kernel(const Slice2<T>* __restrict__ sl, int nslices,
                                                              int nitems, const Tin* __restrict__ proj,
                                                              Tout* __restrict__ img) {
    __shared__ T part[4][kW2];
    // blocks in dispatch order, as in the forward (a contiguous range per XCD measured 19%
    // slower here too)
    const int bid = blockIdx.x;
    const int wv = (int)(threadIdx.x / kW2), lane = lane_id();
    const int item = bid * (4 / VS) + wv / VS;
    const bool valid = item < nitems;
    T v = T(0);
    int si = 0, ix = 0, iy = 0;
    if (valid) {
        // the item's slice geometry in registers (a wave-uniform index: scalar loads, once
        // per item) rather than read from the slice table inside the view loop
        si = uni(find_slice(sl, nslices, item));
        const Slice2<T>& s = sl[si];
        const Geom<T> g = s.g;
        const View<T>* views = s.views;
        const int nviews = s.nviews;
        const int local = item - s.item0;
        const int nxb = (g.nx + kW2 - 1) / kW2;
        iy = local / nxb;
        ix = (local % nxb) * kW2 + lane;
        const int share = wv % VS;
        const int per = (nviews + VS - 1) / VS;
        const int b0 = min(share * per, nviews), b1 = min(b0 + per, nviews);
        v = fbct_back_pixel<Tin, T, KIND>(g, views, 1, b0, b1, ix, iy, proj + s.sino_off);
    }
    if (VS > 1) {
        part[wv][lane] = v;
        __syncthreads();
        if (wv % VS) return;
        v = combine_shares<T, VS>(part, wv, lane);
    }
    if (!valid) return;
    const Slice2<T>& s = sl[si];
    if (ix < s.g.nx) img[s.img_off + (size_t)iy * s.g.nx + ix] = cvt_store<Tout>(v);
}

}  // namespace ctsim

// ===================================================================================
// host launchers
// ===================================================================================
namespace ctsim {

template <class Tin, class Tout, class T>
static hipError_t proj2(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* img, void* proj,
                        int kind, int exhaustive, hipStream_t st) {
    const long long nut = (g.nu + kW2 - 1) / kW2;
    const long long items = (long long)nviews * nut;
    if (items == 0) return hipSuccess;
    if (items < kFewItems2D) {
        const unsigned blocks = (unsigned)items;
        if (kind == F2_DD)
            hipLaunchKernelGGL((fbct_proj_kernel<Tin, Tout, T, F2_DD, 4>), dim3(blocks), dim3(256), 0, st, g, views,
                               vstride, nviews, (const Tin*)img, (Tout*)proj, exhaustive);
        else
            hipLaunchKernelGGL((fbct_proj_kernel<Tin, Tout, T, F2_TF, 4>), dim3(blocks), dim3(256), 0, st, g, views,
                               vstride, nviews, (const Tin*)img, (Tout*)proj, exhaustive);
        return hipGetLastError();
    }
    const unsigned blocks = (unsigned)((items + 3) / 4);
    if (kind == F2_DD)
        hipLaunchKernelGGL((fbct_proj_kernel<Tin, Tout, T, F2_DD, 1>), dim3(blocks), dim3(256), 0, st, g, views, vstride,
                           nviews, (const Tin*)img, (Tout*)proj, exhaustive);
    else
        hipLaunchKernelGGL((fbct_proj_kernel<Tin, Tout, T, F2_TF, 1>), dim3(blocks), dim3(256), 0, st, g, views, vstride,
                           nviews, (const Tin*)img, (Tout*)proj, exhaustive);
    return hipGetLastError();
}

template <class Tin, class Tout, class T>
static hipError_t back2(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* proj, void* img,
                        int kind, T scale, hipStream_t st) {
    const unsigned blocks = (unsigned)(((g.nx + kW2 - 1) / kW2) * g.ny);
    if (blocks == 0) return hipSuccess;
    if (kind == 0)
        hipLaunchKernelGGL((fbct_back_kernel<Tin, Tout, T, 0>), dim3(blocks), dim3(256), 0, st, g, views, vstride,
                           nviews, (const Tin*)proj, (Tout*)img, scale);
    else if (kind == 1)
        hipLaunchKernelGGL((fbct_back_kernel<Tin, Tout, T, 1>), dim3(blocks), dim3(256), 0, st, g, views, vstride,
                           nviews, (const Tin*)proj, (Tout*)img, scale);
    else
        hipLaunchKernelGGL((fbct_back_kernel<Tin, Tout, T, 2>), dim3(blocks), dim3(256), 0, st, g, views, vstride,
                           nviews, (const Tin*)proj, (Tout*)img, scale);
    return hipGetLastError();
}

#define CT_DISPATCH_IN(dt, MACRO)                   \
    switch (dt) {                                   \
        case DT_F32: MACRO(float);                  \
        case DT_F16: MACRO(__half);                 \
        case DT_BF16: MACRO(__hip_bfloat16);        \
        default: return hipErrorInvalidValue;       \
    }

template <>
hipError_t fbct_proj_launch<float>(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                   const void* img, int img_dt, void* proj, int proj_dt, int kind, int exhaustive,
                                   hipStream_t st) {
#define OUT_F32(TI) return proj2<TI, float, float>(g, views, vstride, nviews, img, proj, kind, exhaustive, st)
#define OUT_F16(TI) return proj2<TI, __half, float>(g, views, vstride, nviews, img, proj, kind, exhaustive, st)
#define OUT_BF16(TI) return proj2<TI, __hip_bfloat16, float>(g, views, vstride, nviews, img, proj, kind, exhaustive, st)
    switch (proj_dt) {
        case DT_F32: CT_DISPATCH_IN(img_dt, OUT_F32);
        case DT_F16: CT_DISPATCH_IN(img_dt, OUT_F16);
        case DT_BF16: CT_DISPATCH_IN(img_dt, OUT_BF16);
        default: return hipErrorInvalidValue;
    }
#undef OUT_F32
#undef OUT_F16
#undef OUT_BF16
}

template <>
hipError_t fbct_proj_launch<double>(const Geom<double>& g, const View<double>* views, int vstride, int nviews,
                                    const void* img, int img_dt, void* proj, int proj_dt, int kind, int exhaustive,
                                    hipStream_t st) {
    if (img_dt != DT_F64 || proj_dt != DT_F64) return hipErrorInvalidValue;
    return proj2<double, double, double>(g, views, vstride, nviews, img, proj, kind, exhaustive, st);
}

template <>
hipError_t fbct_back_launch<float>(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                   const void* proj, int proj_dt, void* img, int img_dt, int kind, float scale,
                                   hipStream_t st) {
    if (img_dt != DT_F32) return hipErrorInvalidValue;
#define B2(TI) return back2<TI, float, float>(g, views, vstride, nviews, proj, img, kind, scale, st)
    CT_DISPATCH_IN(proj_dt, B2);
#undef B2
}

template <>
hipError_t fbct_back_launch<double>(const Geom<double>& g, const View<double>* views, int vstride, int nviews,
                                    const void* proj, int proj_dt, void* img, int img_dt, int kind, double scale,
                                    hipStream_t st) {
    if (img_dt != DT_F64 || proj_dt != DT_F64) return hipErrorInvalidValue;
    return back2<double, double, double>(g, views, vstride, nviews, proj, img, kind, scale, st);
}

size_t slice2_size(bool dbl) { return dbl ? sizeof(Slice2<double>) : sizeof(Slice2<float>); }

template <class Tin, class Tout, class T, int FOOT, int S>
static void proj2b_s(const void* sl, int nslices, int nitems, const void* img, void* proj, hipStream_t st) {
    const unsigned blocks = (unsigned)(((long long)nitems * S + 3) / 4);
    hipLaunchKernelGGL((fbct_proj_batch_kernel<Tin, Tout, T, FOOT, S>), dim3(blocks), dim3(256), 0, st,
                       (const Slice2<T>*)sl, nslices, nitems, (const Tin*)img, (Tout*)proj);
}

template <class Tin, class Tout, class T>
static hipError_t proj2b(const void* sl, int nslices, int nitems, const void* img, void* proj, int kind, int split,
                         hipStream_t st) {
    if (nitems == 0) return hipSuccess;
    if (split != 1 && split != 2 && split != 4) return hipErrorInvalidValue;
#define CT_P2B(F, S) proj2b_s<Tin, Tout, T, F, S>(sl, nslices, nitems, img, proj, st)
    if (kind == F2_DD) {
        if (split == 1) CT_P2B(F2_DD, 1); else if (split == 2) CT_P2B(F2_DD, 2); else CT_P2B(F2_DD, 4);
    } else {
        if (split == 1) CT_P2B(F2_TF, 1); else if (split == 2) CT_P2B(F2_TF, 2); else CT_P2B(F2_TF, 4);
    }
#undef CT_P2B
    return hipGetLastError();
}

template <class Tin, class Tout, class T, int KIND, int VS>
static void back2b_s(const void* sl, int nslices, int nitems, const void* proj, void* img, hipStream_t st) {
    const unsigned blocks = (unsigned)(((long long)nitems * VS + 3) / 4);
    hipLaunchKernelGGL((fbct_back_batch_kernel<Tin, Tout, T, KIND, VS>), dim3(blocks), dim3(256), 0, st,
                       (const Slice2<T>*)sl, nslices, nitems, (const Tin*)proj, (Tout*)img);
}

template <class Tin, class Tout, class T>
static hipError_t back2b(const void* sl, int nslices, int nitems, const void* proj, void* img, int kind, int split,
                         hipStream_t st) {
    if (nitems == 0) return hipSuccess;
    if (split != 1 && split != 2 && split != 4) return hipErrorInvalidValue;
#define CT_B2B(K)                                                                         \
    do {                                                                                  \
        if (split == 1) back2b_s<Tin, Tout, T, K, 1>(sl, nslices, nitems, proj, img, st); \
        else if (split == 2) back2b_s<Tin, Tout, T, K, 2>(sl, nslices, nitems, proj, img, st); \
        else back2b_s<Tin, Tout, T, K, 4>(sl, nslices, nitems, proj, img, st);            \
    } while (0)
    if (kind == 0) CT_B2B(0);
    else if (kind == 1) CT_B2B(1);
    else CT_B2B(2);
#undef CT_B2B
    return hipGetLastError();
}

template <>
hipError_t fbct_proj_batch_launch<float>(const void* sl, int nslices, int nitems, const void* img, int img_dt,
                                         void* proj, int proj_dt, int kind, int split, hipStream_t st) {
#define OUT_F32(TI) return proj2b<TI, float, float>(sl, nslices, nitems, img, proj, kind, split, st)
#define OUT_F16(TI) return proj2b<TI, __half, float>(sl, nslices, nitems, img, proj, kind, split, st)
#define OUT_BF16(TI) return proj2b<TI, __hip_bfloat16, float>(sl, nslices, nitems, img, proj, kind, split, st)
    switch (proj_dt) {
        case DT_F32: CT_DISPATCH_IN(img_dt, OUT_F32);
        case DT_F16: CT_DISPATCH_IN(img_dt, OUT_F16);
        case DT_BF16: CT_DISPATCH_IN(img_dt, OUT_BF16);
        default: return hipErrorInvalidValue;
    }
#undef OUT_F32
#undef OUT_F16
#undef OUT_BF16
}

template <>
hipError_t fbct_proj_batch_launch<double>(const void* sl, int nslices, int nitems, const void* img, int img_dt,
                                          void* proj, int proj_dt, int kind, int split, hipStream_t st) {
    if (img_dt != DT_F64 || proj_dt != DT_F64) return hipErrorInvalidValue;
    return proj2b<double, double, double>(sl, nslices, nitems, img, proj, kind, split, st);
}

template <>
hipError_t fbct_back_batch_launch<float>(const void* sl, int nslices, int nitems, const void* proj, int proj_dt,
                                         void* img, int img_dt, int kind, int split, hipStream_t st) {
#define OUT_F32(TI) return back2b<TI, float, float>(sl, nslices, nitems, proj, img, kind, split, st)
#define OUT_F16(TI) return back2b<TI, __half, float>(sl, nslices, nitems, proj, img, kind, split, st)
    switch (img_dt) {
        case DT_F32: CT_DISPATCH_IN(proj_dt, OUT_F32);
        case DT_F16: CT_DISPATCH_IN(proj_dt, OUT_F16);
        default: return hipErrorInvalidValue;
    }
#undef OUT_F32
#undef OUT_F16
}

template <>
hipError_t fbct_back_batch_launch<double>(const void* sl, int nslices, int nitems, const void* proj, int proj_dt,
                                          void* img, int img_dt, int kind, int split, hipStream_t st) {
    if (img_dt != DT_F64 || proj_dt != DT_F64) return hipErrorInvalidValue;
    return back2b<double, double, double>(sl, nslices, nitems, proj, img, kind, split, st);
}

}  // namespace ctsim
