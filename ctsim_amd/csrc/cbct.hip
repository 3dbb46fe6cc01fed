// 3D cone-beam projector family for gfx950: distance-driven (DD) and separable-
// footprint (SF, "tf") forward projection, their matched backprojections, and the
// voxel-driven FDK backprojection (pd / hpd).
//
// Semantics: geom/cbct_os/cbct_proj.c (cbct_proj1_dd :427-642, cbct_proj1_tf :912-1169)
// and geom/cbct_os/cbct_back.c (cbct_back1_pd :135-262, cbct_back1_dd :479-692,
// cbct_back1_tf :945-1197).  The reference scatters voxel by voxel; here both
// directions are gathers with one writer per output element, so no atomics are used
// and results are deterministic.
//
// Device layouts: volume "ZXY" (z fastest: vol[(iy*nx + ix)*nz + iz]) so that a
// wavefront walks one voxel column; sinogram [view][u][v] (v fastest) as in MATLAB.
#include "internal.h"

namespace ctsim {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

enum Foot { FOOT_DD = 0, FOOT_TF = 1 };

// ---------------------------------------------------------------------------------
// Forward projection: one wavefront per (view, tile of TU detector columns, 64 detector
// rows).  Lanes own detector rows iv and TU accumulators.  The wave enumerates the
// image columns whose footprint can reach its tile (a strip, found by inverse-mapping
// the tile edges line by line), evaluates 64 candidate footprints at a time (one per
// lane), and then walks the surviving columns one by one with their parameters
// broadcast from the owning lane.  For each column every lane gathers the voxels whose
// v footprint covers its row, with the exact clamps of the reference.
// ---------------------------------------------------------------------------------

template <class Tin, class Tout, class T, int FOOT, int TU>
__global__ __launch_bounds__(kBlock) void cbct_proj_kernel(Geom<T> g, const View<T>* __restrict__ views,
                                                           int vstride, int nviews, const Tin* __restrict__ vol,
                                                           const uint8_t* __restrict__ map,
                                                           Tout* __restrict__ proj, int exhaustive) {
    const int nut = (g.nu + TU - 1) / TU;
    const int nvc = (g.nv + kWave - 1) / kWave;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int wid = bid * kWavesPerBlock + (int)(threadIdx.x / kWave);
    if (wid >= nviews * nut * nvc) return;  // whole wave exits together
    const int vc = wid % nvc;
    const int ut = (wid / nvc) % nut;
    const int b = wid / (nvc * nut);
    const int lane = lane_id();
    const int iv = vc * kWave + lane;
    const int iu_a = ut * TU;
    const int iu_b = min(iu_a + TU, g.nu) - 1;  // last cell of the tile
    const View<T> v = views[(size_t)b * vstride];
    const bool row_major = fabs_t(v.c) >= fabs_t(v.s);
    const int nmajor = row_major ? g.ny : g.nx;
    const T kL = u_to_slope(g, T(iu_a) - T(0.5));
    const T kR = u_to_slope(g, T(iu_b) + T(0.5));
    const T v_lo_chunk = T(vc * kWave), v_hi_chunk = T(vc * kWave + kWave);

    T acc[TU];
#pragma unroll
    for (int t = 0; t < TU; ++t) acc[t] = T(0);

    for (int m0 = 0; m0 < nmajor; m0 += kWave) {
        const int m = m0 + lane;
        int lo = 0, hi = -1;
        if (m < nmajor) {
            if (!strip_range(g, v, row_major, m, kL, kR, exhaustive != 0, lo, hi)) hi = lo - 1;
        }
        const int cnt = hi - lo + 1 > 0 ? hi - lo + 1 : 0;
        const int incl = wave_incl_scan(cnt);
        const int excl = incl - cnt;
        const int total = rdl_i(incl, kWave - 1);
        for (int k0 = 0; k0 < total; k0 += kWave) {
            const int k = k0 + lane;
            // owning line: the largest j with excl_j <= k
            int j = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1) {
                const int e = __shfl(excl, j + step, 64);
                if (e <= k) j += step;
            }
            const int lo_j = __shfl(lo, j, 64);
            const int ex_j = __shfl(excl, j, 64);
            const int major = m0 + j;
            const int minor = lo_j + (k - ex_j);
            const int ix = row_major ? minor : major;
            const int iy = row_major ? major : minor;

            // exact footprint of the candidate column
            bool ok = k < total;
            int ilo = 0, ihi = -1, iz0 = 0, iz1 = 0;
            T p0 = T(0), p1 = T(0), p2 = T(0), p3 = T(0), amp = T(0), r = T(1), zinc = T(1), vstart = T(0);
            if (ok && map != nullptr) ok = map[iy * g.nx + ix] != 0;
            if (ok) {
                const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
                const Column<T> cg = column_geom(xc, yc, v, g.sad, g.dx);
                T clo, chi;
                if (FOOT == FOOT_DD) {
                    T umin, umax;
                    dd_bounds(cg, v, g, umin, umax);
                    ok = clamp_cells(umin, umax, g.nu, 16, clo, chi, ilo, ihi);
                    p0 = clo;
                    p1 = chi;
                } else {
                    T tau[4];
                    sf_taus(cg, v, g, tau);
                    ok = clamp_cells(tau[0], tau[3], g.nu, 16, clo, chi, ilo, ihi);
                    p0 = tau[0]; p1 = tau[1]; p2 = tau[2]; p3 = tau[3];
                }
                ok = ok && ihi >= iu_a && ilo <= iu_b;
                if (ok) {
                    ok = v_window(g, cg.d, v.wz, zinc, vstart, iz0, iz1) && iz0 < iz1;
                    amp = cg.amp;
                    r = cg.r;
                }
                if (ok) {
                    // does the column's v extent reach this wave's rows at all?
                    const T e0 = vstart + T(iz0) * zinc + T(0.5);
                    const T e1 = vstart + T(iz1) * zinc + T(0.5);
                    ok = e1 > v_lo_chunk - T(1) && e0 < v_hi_chunk + T(1);
                }
            }
            unsigned long long mask = __ballot(ok);
            while (mask) {
                const int src = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const int cix = rdl_i(ix, src), ciy = rdl_i(iy, src);
                const int cilo = rdl_i(ilo, src), cihi = rdl_i(ihi, src);
                const int ciz0 = rdl_i(iz0, src), ciz1 = rdl_i(iz1, src);
                const T cp0 = rdl(p0, src), cp1 = rdl(p1, src);
                T cp2 = T(0), cp3 = T(0);
                if (FOOT == FOOT_TF) {
                    cp2 = rdl(p2, src);
                    cp3 = rdl(p3, src);
                }
                const T camp = rdl(amp, src), cr = rdl(r, src);
                const T czinc = rdl(zinc, src), cvstart = rdl(vstart, src);
                const Tin* col = vol + (size_t)(ciy * g.nx + cix) * g.nz;

                // v gather for this lane's detector row
                T q = T(0);
                if (iv < g.nv) {
                    const T inv = T(1) / czinc;
                    const T f = (T(iv) - cvstart) * inv;
                    int za = (int)floor(f - T(0.5) * inv) - 1;
                    int zb = (int)floor(f + T(0.5) * inv) + 1;
                    za = za > ciz0 ? za : ciz0;
                    zb = zb < ciz1 - 1 ? zb : ciz1 - 1;
                    const T inv_r = T(1) / cr;
                    for (int iz = za; iz <= zb; ++iz) {
                        const T vmin = cvstart + T(iz) * czinc;
                        T clo, chi;
                        int vlo, vhi;
                        if (!clamp_cells(vmin, vmin + czinc, g.nv, 16, clo, chi, vlo, vhi)) continue;
                        const T wv = dd_weight(iv, clo, chi, vlo, vhi);
                        if (wv == T(0)) continue;
                        const T zc = (T(iz) + v.wz) * g.dz;
                        const T ampv = sqrt(cr * cr + zc * zc) * inv_r;
                        q += wv * ampv * ld<T>(col + iz);
                    }
                }
                const T aq = camp * q;
#pragma unroll
                for (int t = 0; t < TU; ++t) {
                    const int iu = iu_a + t;
                    if (iu >= cilo && iu <= cihi) {
                        T wu;
                        if (FOOT == FOOT_DD)
                            wu = dd_weight(iu, cp0, cp1, cilo, cihi);
                        else
                            wu = sf_weight(iu, cp0, cp1, cp2, cp3);
                        acc[t] += wu * aq;
                    }
                }
            }
        }
    }
    if (iv < g.nv) {
        Tout* out = proj + (size_t)b * g.nu * g.nv;
#pragma unroll
        for (int t = 0; t < TU; ++t) {
            const int iu = iu_a + t;
            if (iu < g.nu) out[(size_t)iu * g.nv + iv] = cvt_store<Tout>(acc[t]);
        }
    }
}

// ---------------------------------------------------------------------------------
// Matched backprojection: one wavefront per (image column, 64 consecutive z).  The
// column's u footprint is wave-uniform; each lane gathers its voxel's v footprint
// from the sinogram rows of the footprint cells (v fastest, so loads coalesce) and
// accumulates all views in registers; one store per voxel at the end.
// ---------------------------------------------------------------------------------
template <class Tin, class Tout, class T, int FOOT>
__global__ __launch_bounds__(kBlock) void cbct_back_kernel(Geom<T> g, const View<T>* __restrict__ views, int vstride, int nviews,
                                                           const Tin* __restrict__ proj,
                                                           const uint8_t* __restrict__ map,
                                                           Tout* __restrict__ vol, T scale, int accumulate) {
    const int nzc = (g.nz + kWave - 1) / kWave;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int wid = bid * kWavesPerBlock + (int)(threadIdx.x / kWave);
    if (wid >= g.nx * g.ny * nzc) return;
    const int zc = wid % nzc;
    const int colid = wid / nzc;
    const int ix = colid % g.nx, iy = colid / g.nx;
    const int iz = zc * kWave + lane_id();
    if (map != nullptr && map[iy * g.nx + ix] == 0) {
        if (iz < g.nz && !accumulate) vol[(size_t)colid * g.nz + iz] = cvt_store<Tout>(T(0));
        return;
    }
    const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
    T img = T(0);
    for (int b = 0; b < nviews; ++b) {
        const View<T> v = views[(size_t)b * vstride];
        const Column<T> cg = column_geom(xc, yc, v, g.sad, g.dx);
        T clo, chi, p0, p1, p2 = T(0), p3 = T(0);
        int ilo, ihi;
        bool ok;
        if (FOOT == FOOT_DD) {
            T umin, umax;
            dd_bounds(cg, v, g, umin, umax);
            ok = clamp_cells(umin, umax, g.nu, 16, clo, chi, ilo, ihi);
            p0 = clo;
            p1 = chi;
        } else {
            T tau[4];
            sf_taus(cg, v, g, tau);
            ok = clamp_cells(tau[0], tau[3], g.nu, 16, clo, chi, ilo, ihi);
            p0 = tau[0]; p1 = tau[1]; p2 = tau[2]; p3 = tau[3];
        }
        if (!ok) continue;
        T zinc, vstart;
        int iz0, iz1;
        if (!v_window(g, cg.d, v.wz, zinc, vstart, iz0, iz1)) continue;
        if (iz < iz0 || iz >= iz1) continue;
        const T vmin = vstart + T(iz) * zinc;
        int vlo, vhi;
        if (!clamp_cells(vmin, vmin + zinc, g.nv, 16, clo, chi, vlo, vhi)) continue;
        const T zcm = (T(iz) + v.wz) * g.dz;
        const T ampv = sqrt(cg.r * cg.r + zcm * zcm) / cg.r;
        const Tin* pv = proj + (size_t)b * g.nu * g.nv;
        T s = T(0);
        for (int iu = ilo; iu <= ihi; ++iu) {
            const T wu = FOOT == FOOT_DD ? dd_weight(iu, p0, p1, ilo, ihi) : sf_weight(iu, p0, p1, p2, p3);
            const Tin* row = pv + (size_t)iu * g.nv;
            for (int ivv = vlo; ivv <= vhi; ++ivv) s += ld<T>(row + ivv) * wu * dd_weight(ivv, clo, chi, vlo, vhi);
        }
        img += cg.amp * ampv * s;
    }
    if (iz < g.nz) {
        Tout* o = vol + (size_t)colid * g.nz + iz;
        const T val = scale * img;
        *o = cvt_store<Tout>(accumulate ? ld<T>(o) + val : val);
    }
}

// ---------------------------------------------------------------------------------
// FDK voxel-driven bilinear backprojection (cbct_back1_pd).  off_boundary = 1 is
// 'back,pd' (skip below the detector, stop above it), 0 is 'back,hpd' (clamp v).
// ---------------------------------------------------------------------------------
template <class Tin, class Tout, class T>
__global__ __launch_bounds__(kBlock) void cbct_back_pd_kernel(Geom<T> g, const View<T>* __restrict__ views,
                                                              int vstride, int nviews, const Tin* __restrict__ proj,
                                                              const uint8_t* __restrict__ map,
                                                              Tout* __restrict__ vol, T scale, int accumulate,
                                                              int off_boundary) {
    const int nzc = (g.nz + kWave - 1) / kWave;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int wid = bid * kWavesPerBlock + (int)(threadIdx.x / kWave);
    if (wid >= g.nx * g.ny * nzc) return;
    const int zc = wid % nzc;
    const int colid = wid / nzc;
    const int ix = colid % g.nx, iy = colid / g.nx;
    const int iz = zc * kWave + lane_id();
    if (map != nullptr && map[iy * g.nx + ix] == 0) {
        if (iz < g.nz && !accumulate) vol[(size_t)colid * g.nz + iz] = cvt_store<Tout>(T(0));
        return;
    }
    const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
    const double lim_u = (double)(g.nu - 1) - 1e-6;
    const double lim_v = (double)(g.nv - 1) - 1e-6;
    T img = T(0);
    for (int b = 0; b < nviews; ++b) {
        const View<T> v = views[(size_t)b * vstride];
        const T xr = xc * v.c - yc * v.s;
        const T yr = xc * v.s + yc * v.c;
        T u, w2, w2s;
        if (g.dettype == 1) {
            w2 = g.sdd / (g.sad + yr);
            w2s = w2 * w2;
            u = xr * w2 / g.du - g.wu;
        } else {
            const T roff = g.ou * g.sad / g.sdd;
            u = atan((xr - roff) / (g.sad + yr)) * g.sdd / g.du - g.wu;
            w2s = g.sdd * g.sdd / ((g.sad + yr) * (g.sad + yr) + (xr - roff) * (xr - roff));
            w2 = g.sdd / (g.sad + yr);
        }
        if (!((double)u >= 1e-6 && (double)u <= lim_u)) continue;  // also skips NaN
        const int iu = (int)floor(u);
        const T wul = u - T(iu);
        if (iz >= g.nz) continue;
        T vv = (T(iz) + v.wz) * g.dz * w2 / g.dv - g.wv;
        if (off_boundary) {
            // the reference walks iz upwards and stops at the first voxel above the
            // detector; v is monotone in iz, so that is a per-voxel test when v grows,
            // and "nothing if voxel 0 is already above" when v shrinks
            if (!((double)vv >= 1e-6 && (double)vv <= lim_v)) continue;
            if (!(g.dz * w2 > T(0))) {
                const T v0 = v.wz * g.dz * w2 / g.dv - g.wv;
                if ((double)v0 > lim_v) continue;
            }
        } else {
            if ((double)vv < 1e-6) vv = T(0.5);
            if ((double)vv > lim_v) vv = T(g.nv) - T(1.5);
            if (vv != vv) continue;  // NaN
        }
        int iv = (int)floor(vv);
        iv = iv < 0 ? 0 : (iv > g.nv - 2 ? g.nv - 2 : iv);  // no-op for finite v; keeps NaN-free reads in range
        const T wvl = vv - T(iv);
        const Tin* p = proj + ((size_t)b * g.nu + iu) * g.nv + iv;
        const T a0 = ld<T>(p), a1 = ld<T>(p + 1), a2 = ld<T>(p + g.nv), a3 = ld<T>(p + g.nv + 1);
        img += w2s * ((T(1) - wul) * ((T(1) - wvl) * a0 + wvl * a1) + wul * ((T(1) - wvl) * a2 + wvl * a3));
    }
    if (iz < g.nz) {
        Tout* o = vol + (size_t)colid * g.nz + iz;
        const T val = scale * img;
        *o = cvt_store<Tout>(accumulate ? ld<T>(o) + val : val);
    }
}

// ---------------------------------------------------------------------------------
// Layout changes between MATLAB order (x fastest) and the device ZXY order.
// 32x32 tiles through LDS so that both sides are coalesced.
// ---------------------------------------------------------------------------------
template <class S, class D>
__global__ __launch_bounds__(256) void xyz_to_zxy_kernel(const S* __restrict__ in, D* __restrict__ out, int nxy,
                                                         int nz) {
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, bz = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int z = bz + k, c = bx + tx;
        tile[k][tx] = (z < nz && c < nxy) ? ld_f(in + (size_t)z * nxy + c) : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = bx + k, z = bz + tx;
        if (c < nxy && z < nz) out[(size_t)c * nz + z] = cvt_store<D>(tile[tx][k]);
    }
}

template <class S, class D>
__global__ __launch_bounds__(256) void zxy_to_xyz_kernel(const S* __restrict__ in, D* __restrict__ out, int nxy,
                                                         int nz) {
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, bz = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int c = bx + k, z = bz + tx;
        tile[k][tx] = (c < nxy && z < nz) ? ld_f(in + (size_t)c * nz + z) : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int z = bz + k, c = bx + tx;
        if (z < nz && c < nxy) out[(size_t)z * nxy + c] = cvt_store<D>(tile[tx][k]);
    }
}

// double variants keep full precision
__global__ __launch_bounds__(256) void xyz_to_zxy_f64(const double* __restrict__ in, double* __restrict__ out,
                                                      int nxy, int nz) {
    __shared__ double tile[32][33];
    const int bx = blockIdx.x * 32, bz = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int z = bz + k, c = bx + tx;
        tile[k][tx] = (z < nz && c < nxy) ? in[(size_t)z * nxy + c] : 0.0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = bx + k, z = bz + tx;
        if (c < nxy && z < nz) out[(size_t)c * nz + z] = tile[tx][k];
    }
}

__global__ __launch_bounds__(256) void zxy_to_xyz_f64(const double* __restrict__ in, double* __restrict__ out,
                                                      int nxy, int nz) {
    __shared__ double tile[32][33];
    const int bx = blockIdx.x * 32, bz = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int c = bx + k, z = bz + tx;
        tile[k][tx] = (c < nxy && z < nz) ? in[(size_t)c * nz + z] : 0.0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int z = bz + k, c = bx + tx;
        if (z < nz && c < nxy) out[(size_t)z * nxy + c] = tile[tx][k];
    }
}

}  // namespace ctsim

// ===================================================================================
// host launchers
// ===================================================================================
namespace ctsim {

constexpr int kTU = 16;  // detector columns per forward-projection wave

template <class Tin, class Tout, class T>
static hipError_t proj3(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* vol,
                        const uint8_t* map, void* proj, int foot, int exhaustive, hipStream_t st) {
    const long long nut = (g.nu + kTU - 1) / kTU, nvc = (g.nv + kWave - 1) / kWave;
    const long long waves = (long long)nviews * nut * nvc;
    const unsigned blocks = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    if (blocks == 0) return hipSuccess;
    if (foot == FOOT_DD)
        hipLaunchKernelGGL((cbct_proj_kernel<Tin, Tout, T, FOOT_DD, kTU>), dim3(blocks), dim3(kBlock), 0, st, g, views,
                           vstride, nviews, (const Tin*)vol, map, (Tout*)proj, exhaustive);
    else
        hipLaunchKernelGGL((cbct_proj_kernel<Tin, Tout, T, FOOT_TF, kTU>), dim3(blocks), dim3(kBlock), 0, st, g, views,
                           vstride, nviews, (const Tin*)vol, map, (Tout*)proj, exhaustive);
    return hipGetLastError();
}

template <class Tin, class Tout, class T>
static hipError_t back3(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* proj,
                        const uint8_t* map, void* vol, int foot, int pd_mode, T scale, int accumulate,
                        hipStream_t st) {
    const long long nzc = (g.nz + kWave - 1) / kWave;
    const long long waves = (long long)g.nx * g.ny * nzc;
    const unsigned blocks = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    if (blocks == 0) return hipSuccess;
    if (foot == 2 && (g.nu < 2 || g.nv < 2)) return hipErrorInvalidValue;  // bilinear needs 2x2 cells
    if (foot == 2)
        hipLaunchKernelGGL((cbct_back_pd_kernel<Tin, Tout, T>), dim3(blocks), dim3(kBlock), 0, st, g, views, vstride,
                           nviews, (const Tin*)proj, map, (Tout*)vol, scale, accumulate, pd_mode);
    else if (foot == FOOT_DD)
        hipLaunchKernelGGL((cbct_back_kernel<Tin, Tout, T, FOOT_DD>), dim3(blocks), dim3(kBlock), 0, st, g, views,
                           vstride, nviews, (const Tin*)proj, map, (Tout*)vol, scale, accumulate);
    else
        hipLaunchKernelGGL((cbct_back_kernel<Tin, Tout, T, FOOT_TF>), dim3(blocks), dim3(kBlock), 0, st, g, views,
                           vstride, nviews, (const Tin*)proj, map, (Tout*)vol, scale, accumulate);
    return hipGetLastError();
}

template <>
hipError_t cbct_proj_launch<float>(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                   const void* vol, int vol_dt, const uint8_t* map, void* proj, int proj_dt, int foot,
                                   int exhaustive, hipStream_t st) {
#define P3(TI, TO) return proj3<TI, TO, float>(g, views, vstride, nviews, vol, map, proj, foot, exhaustive, st)
#define P3_OUT(TI)                                   \
    switch (proj_dt) {                               \
        case DT_F32: P3(TI, float);                  \
        case DT_F16: P3(TI, __half);                 \
        case DT_BF16: P3(TI, __hip_bfloat16);        \
        default: return hipErrorInvalidValue;        \
    }
    switch (vol_dt) {
        case DT_F32: P3_OUT(float);
        case DT_F16: P3_OUT(__half);
        case DT_BF16: P3_OUT(__hip_bfloat16);
        default: return hipErrorInvalidValue;
    }
#undef P3_OUT
#undef P3
}

template <>
hipError_t cbct_proj_launch<double>(const Geom<double>& g, const View<double>* views, int vstride, int nviews,
                                    const void* vol, int vol_dt, const uint8_t* map, void* proj, int proj_dt, int foot,
                                    int exhaustive, hipStream_t st) {
    if (vol_dt != DT_F64 || proj_dt != DT_F64) return hipErrorInvalidValue;
    return proj3<double, double, double>(g, views, vstride, nviews, vol, map, proj, foot, exhaustive, st);
}

template <>
hipError_t cbct_back_launch<float>(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                   const void* proj, int proj_dt, const uint8_t* map, void* vol, int vol_dt, int foot,
                                   float scale, int accumulate, hipStream_t st) {
    if (vol_dt != DT_F32) return hipErrorInvalidValue;
    switch (proj_dt) {
        case DT_F32: return back3<float, float, float>(g, views, vstride, nviews, proj, map, vol, foot, 0, scale, accumulate, st);
        case DT_F16: return back3<__half, float, float>(g, views, vstride, nviews, proj, map, vol, foot, 0, scale, accumulate, st);
        case DT_BF16: return back3<__hip_bfloat16, float, float>(g, views, vstride, nviews, proj, map, vol, foot, 0, scale, accumulate, st);
        default: return hipErrorInvalidValue;
    }
}

template <>
hipError_t cbct_back_launch<double>(const Geom<double>& g, const View<double>* views, int vstride, int nviews,
                                    const void* proj, int proj_dt, const uint8_t* map, void* vol, int vol_dt, int foot,
                                    double scale, int accumulate, hipStream_t st) {
    if (vol_dt != DT_F64 || proj_dt != DT_F64) return hipErrorInvalidValue;
    return back3<double, double, double>(g, views, vstride, nviews, proj, map, vol, foot, 0, scale, accumulate, st);
}

template <>
hipError_t cbct_back_pd_launch<float>(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                      const void* proj, int proj_dt, const uint8_t* map, void* vol, int vol_dt,
                                      int off_boundary, float scale, int accumulate, hipStream_t st) {
    if (vol_dt != DT_F32) return hipErrorInvalidValue;
    switch (proj_dt) {
        case DT_F32: return back3<float, float, float>(g, views, vstride, nviews, proj, map, vol, 2, off_boundary, scale, accumulate, st);
        case DT_F16: return back3<__half, float, float>(g, views, vstride, nviews, proj, map, vol, 2, off_boundary, scale, accumulate, st);
        case DT_BF16: return back3<__hip_bfloat16, float, float>(g, views, vstride, nviews, proj, map, vol, 2, off_boundary, scale, accumulate, st);
        default: return hipErrorInvalidValue;
    }
}

template <>
hipError_t cbct_back_pd_launch<double>(const Geom<double>& g, const View<double>* views, int vstride, int nviews,
                                       const void* proj, int proj_dt, const uint8_t* map, void* vol, int vol_dt,
                                       int off_boundary, double scale, int accumulate, hipStream_t st) {
    if (vol_dt != DT_F64 || proj_dt != DT_F64) return hipErrorInvalidValue;
    return back3<double, double, double>(g, views, vstride, nviews, proj, map, vol, 2, off_boundary, scale, accumulate, st);
}

template <class S, class D>
static hipError_t xz(const void* src, void* dst, int nxy, int nz, bool to_zxy, hipStream_t st) {
    dim3 grid((nxy + 31) / 32, (nz + 31) / 32);
    if (to_zxy)
        hipLaunchKernelGGL((xyz_to_zxy_kernel<S, D>), grid, dim3(256), 0, st, (const S*)src, (D*)dst, nxy, nz);
    else
        hipLaunchKernelGGL((zxy_to_xyz_kernel<S, D>), grid, dim3(256), 0, st, (const S*)src, (D*)dst, nxy, nz);
    return hipGetLastError();
}

hipError_t xyz_zxy_launch(const void* src, int sdt, void* dst, int ddt, int nxy, int nz, bool to_zxy, hipStream_t st) {
    if (nxy <= 0 || nz <= 0) return hipSuccess;
    if (sdt == DT_F64 || ddt == DT_F64) {
        if (sdt != DT_F64 || ddt != DT_F64) return hipErrorInvalidValue;
        dim3 grid((nxy + 31) / 32, (nz + 31) / 32);
        if (to_zxy)
            hipLaunchKernelGGL(xyz_to_zxy_f64, grid, dim3(256), 0, st, (const double*)src, (double*)dst, nxy, nz);
        else
            hipLaunchKernelGGL(zxy_to_xyz_f64, grid, dim3(256), 0, st, (const double*)src, (double*)dst, nxy, nz);
        return hipGetLastError();
    }
#define XZ_D(S)                                                   \
    switch (ddt) {                                                \
        case DT_F32: return xz<S, float>(src, dst, nxy, nz, to_zxy, st);  \
        case DT_F16: return xz<S, __half>(src, dst, nxy, nz, to_zxy, st); \
        case DT_BF16: return xz<S, __hip_bfloat16>(src, dst, nxy, nz, to_zxy, st); \
        default: return hipErrorInvalidValue;                     \
    }
    switch (sdt) {
        case DT_F32: XZ_D(float);
        case DT_F16: XZ_D(__half);
        case DT_BF16: XZ_D(__hip_bfloat16);
        default: return hipErrorInvalidValue;
    }
#undef XZ_D
}

}  // namespace ctsim
