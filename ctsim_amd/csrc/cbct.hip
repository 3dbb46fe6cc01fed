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
#include <type_traits>

#include "internal.h"
#include "penalty.h"

namespace ctsim {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

enum Foot { FOOT_DD = 0, FOOT_TF = 1 };

// ---------------------------------------------------------------------------------
// Forward projection.  Work item = (view, tile of TU detector columns, RPL*64 detector
// rows); SPLIT waves share an item, each walking a contiguous share of the image lines,
// and combine their partial sums through LDS in a fixed order (deterministic).  Lanes
// own detector rows (RPL of them) and TU accumulators each.
//
// The wave enumerates the image columns whose footprint can reach its tile (the strip,
// found by inverse-mapping the tile edges line by line), evaluates 64 candidate
// footprints at a time (one per lane, including the tile's u weights); each lane keeps
// its candidate's parameters in registers.  The wave then walks the survivors in a
// software pipeline, reading a column's parameters into scalar registers with readlane:
// the voxel loads of the next column are issued once the current column is staged
// (value times the cone factor ampv, in LDS), and fly while it is gathered: every row
// sums the voxels whose v footprint covers it, with the exact clamps of the reference.
// Columns of up to kFastSpan voxels (the common case) take one straight-line pass with
// no window loop, staging only the 64-voxel slots their range needs; longer ones are
// staged in windows of kZWin voxels.  FP32 volumes of at most kFastSpan slices take the
// ring form below (cbct_proj_ring_kernel) instead.
// ---------------------------------------------------------------------------------
constexpr int kZWin = 256;  // voxels staged per LDS window and wave
constexpr int kPre = 3;     // voxel loads per lane issued one column ahead (3 x 64 voxels)
static_assert(kPre == 3, "the pipeline below holds three prefetch registers");
constexpr int kZPad = 4;    // zero words kept after the staged voxels (the gather reads K - 1 <= 3 past them)
// columns of at most this many voxels stage the prefetched slots they need (lanes past zb
// read zeros) and four zero words after the last slot
constexpr int kFastSpan = kPre * kWave;
static_assert(kPre * kWave + kZPad <= kZWin + kZPad, "the prefetched slots and pads fit the staging window");

// voxel range and gather class of a candidate column, packed in one word:
// za | zb << 13 | class << 26 (nz <= 8192).  Class 0..2: the fixed-count gathers of
// 2, 3, 4 voxels per row (zinc in [1, 14], [1/2, 1), [1/3, 1/2)); 3: the counted loop.
constexpr int kZBits = 13;
constexpr int kZMask = (1 << kZBits) - 1;
constexpr int kMaxNz = 1 << kZBits;
constexpr int kBackSlab = 8 * kWave;  // slices per backprojector launch (8 voxels per lane)
constexpr int kSuperTile = 32;        // super-tile edge in columns of the backprojectors' patch order

template <class T>
__device__ __forceinline__ int gather_class(T zinc) {
    if (zinc > T(14) || zinc < T(1) / T(3)) return 3;
    return zinc >= T(1) ? 0 : (zinc >= T(0.5) ? 1 : 2);
}

// Ten registers per lane (no LDS).  1/zinc is recomputed from zinc (the same v_rcp, the
// same value); the footprint corners, needed only for footprints wider than four tile
// cells, are recomputed from the column on that path.  The tile weights carry the
// column's amp / r, so the staged voxel value is sqrt(r^2 + zc^2) * img.
template <class T>
struct ColParams {
    int col;    // image column iy * nx + ix
    int cells;  // footprint cells ilo | ihi << 16 (detectors narrower than 32768 cells)
    int zab;    // voxel range and gather class (see kZBits)
    T r2;       // r^2
    T zinc, vstart, w0, w1, w2, w3;
};

// footprint corners of column (ix, iy) in view v: DD (clo, chi) in p0, p1; SF the sorted taus
template <class T, int FOOT>
__device__ __forceinline__ bool col_corners(const Geom<T>& g, const View<T>& v, int ix, int iy, T& p0, T& p1, T& p2,
                                           T& p3, int& ilo, int& ihi, T& amp, T& r, T& d) {
    const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
    const Column<T> cg = column_geom(xc, yc, v, g.sad, g.dx);
    T clo, chi;
    bool ok;
    if (FOOT == FOOT_DD) {
        T umin, umax;
        dd_bounds(cg, v, g, umin, umax);
        ok = clamp_cells(umin, umax, g.nu, 16, clo, chi, ilo, ihi);
        p0 = clo;
        p1 = chi;
        p2 = p3 = T(0);
    } else {
        T tau[4];
        sf_taus(cg, v, g, tau);
        ok = clamp_cells(tau[0], tau[3], g.nu, 16, clo, chi, ilo, ihi);
        p0 = tau[0]; p1 = tau[1]; p2 = tau[2]; p3 = tau[3];
    }
    amp = cg.amp;
    r = cg.r;
    d = cg.d;
    return ok;
}

// weight of tile cell iu for a column whose u footprint is (ilo, ihi, p*): DD uses
// (clo, chi) in p0, p1; SF the sorted corners p0..p3
template <class T, int FOOT>
__device__ __forceinline__ T u_weight(int iu, int ilo, int ihi, T p0, T p1, T p2, T p3) {
    if (iu < ilo || iu > ihi) return T(0);
    return FOOT == FOOT_DD ? dd_weight(iu, p0, p1, ilo, ihi) : sf_weight(iu, p0, p1, p2, p3);
}

// Staged voxels of a window [zw0, zw1] sit at zwin[1 + j - zw0]; zwin[0] and the
// kZPad words after the last voxel are zero.  gather_row sums, over the K voxels from
// ja on, (overlap with the row) * staged value, where voxel j covers detector rows
// [x_j, x_j + zinc] relative to the row's lower edge, x_j = j zinc - tt, and the row is
// [0, 1].  The DD rule of dd_overlap, max(min(x_j + zinc, 1) - max(x_j, 0), 0), equals
// clamp01(x_{j+1}) - clamp01(x_j) bit for bit when x_j < x_{j+1}: where the overlap is
// positive the clamps reduce to the same min and max, elsewhere both ends clamp to the
// same bound.  ja is clamped to [zw0 - 1, zw1 + 1] so that voxels outside the window
// read zero words.

template <class T, int K>
__device__ __forceinline__ T gather_row(const T* zwin, int zw0, int zw1, int ja, T tt, T zinc) {
    static_assert(K - 1 <= kZPad, "pad covers the voxels past the window");
    ja = min(max(ja, zw0 - 1), zw1 + 1);
    const T* pz = zwin + (ja - zw0 + 1);
    // voxel edges x_k = x_0 + k zinc, each one FMA from x_0 with the clamp folded in
    const T x0 = T(ja) * zinc - tt;
    T cl = clamp01(x0);
    T s = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const T ch = clamp01(fma_t(T(k + 1), zinc, x0));
        s += (ch - cl) * pz[k];
        cl = ch;
    }
    return s;
}

template <class Tin, class Tout, class T, int FOOT, int TU, int RPL, int SPLIT>
__global__ __launch_bounds__(kBlock, sizeof(T) == 4 ? 6 : 1) void cbct_proj_kernel(Geom<T> g, const View<T>* __restrict__ views,
                                                           int vstride, int nviews, const Tin* __restrict__ vol,
                                                           const uint8_t* __restrict__ map,
                                                           Tout* __restrict__ proj, int exhaustive) {
    static_assert(SPLIT == 1 || SPLIT == kWavesPerBlock, "SPLIT is 1 or the block's wave count");
    static_assert(RPL == 1, "one detector row per lane");
    __shared__ T zbuf[kWavesPerBlock][kZWin + 1 + kZPad];
    // each lane's row of the tile, [cell][lane] so that a wave's access is one word per
    // bank.  Every word has one writer (its lane).
    __shared__ T accs[kWavesPerBlock][TU][kWave];
    const int rows = RPL * kWave;
    const int nut = (g.nu + TU - 1) / TU;
    const int nvc = (g.nv + rows - 1) / rows;
    const int items = nviews * nut * nvc;
    // longest work first: the dispatcher starts blocks in index order, and the tiles near
    // the central ray have the longest pixel strips, so items run in order of the tile's
    // distance from that ray (all views and row chunks of a tile together)
    const int bid = blockIdx.x;
    const int wv = (int)(threadIdx.x / kWave);
    const int item = SPLIT == 1 ? bid * kWavesPerBlock + wv : bid;
    const int share = SPLIT == 1 ? 0 : wv;
    if (item >= items) return;  // SPLIT > 1: the whole block leaves together
    const int vc = item % nvc;
    const int b = (item / nvc) % nviews;
    const int ut = tile_by_rank(item / (nvc * nviews), min(max((int)floor(-g.wu / T(TU)), 0), nut - 1), nut);
    const int lane = lane_id();
    const int loff = lane * (int)sizeof(Tin);  // byte offset of this lane's voxel in a prefetch slot
    const int iv_a = vc * rows;
    const int iu_a = ut * TU;
    const int iu_b = min(iu_a + TU, g.nu) - 1;
    const View<T> v = views[(size_t)b * vstride];
    const bool row_major = fabs_t(v.c) >= fabs_t(v.s);
    const int nmajor = row_major ? g.ny : g.nx;
    const int m_end = nmajor;
    const T kL = u_to_slope(g, T(iu_a) - T(0.5));
    const T kR = u_to_slope(g, T(iu_b) + T(0.5));
    const int iv_last = min(iv_a + rows, g.nv) - 1;
    T* zw = zbuf[wv];
    if (lane == 0) zw[0] = T(0);  // the word before the first staged voxel stays zero
    const T wzdz = v.wz * g.dz;                       // zc = z dz + wz dz
    const T fiv_lo = T(iv_a + lane) - T(0.5);          // lower edge of this lane's row
    // this lane's z offset in each prefetch slot: voxel za + i 64 + lane sits at
    // zc = (za dz + wz dz) + lzd[i]
    const T lzd0 = T(lane) * g.dz, lzd1 = T(kWave + lane) * g.dz, lzd2 = T(2 * kWave + lane) * g.dz;

    T* acc = &accs[wv][0][lane];
#pragma unroll
    for (int t = 0; t < TU; ++t) acc[t * kWave] = T(0);

    // the SPLIT waves of a block take every SPLIT-th image line of the strip (lane l of a
    // batch takes line m0 + l SPLIT + share), so that their shares are alike wherever the
    // strip is long: the block holds its LDS until its last wave ends
    for (int m0 = 0; m0 < m_end; m0 += SPLIT * kWave) {
        const int m = m0 + lane * SPLIT + share;
        int lo = 0, hi = -1;
        if (m < m_end) {
            if (!strip_range(g, v, row_major, m, kL, kR, exhaustive != 0, lo, hi)) hi = lo - 1;
        }
        const int cnt = hi - lo + 1 > 0 ? hi - lo + 1 : 0;
        const int incl = wave_incl_scan(cnt);
        const int excl = incl - cnt;
        const int total = rdl_i(incl, kWave - 1);
        for (int k0 = 0; k0 < total; k0 += kWave) {
            {
                const int k = k0 + lane;
                int j = 0;
#pragma unroll
                for (int step = 32; step >= 1; step >>= 1) {
                    const int e = __shfl(excl, j + step, 64);
                    if (e <= k) j += step;
                }
                const int lo_j = __shfl(lo, j, 64);
                const int ex_j = __shfl(excl, j, 64);
                const int major = m0 + j * SPLIT + share;
                const int minor = lo_j + (k - ex_j);
                const int ix = row_major ? minor : major;
                const int iy = row_major ? major : minor;

                // exact footprint of the candidate column and its weights on this tile
                bool ok = k < total;
                ColParams<T> cp{};
                cp.r2 = T(1);
                cp.zinc = T(1);
                if (ok && map != nullptr) ok = map[iy * g.nx + ix] != 0;
                if (ok) {
                    T p0, p1, p2, p3, amp, r, d;
                    int ilo = 0, ihi = -1, iz0 = 0, iz1 = 0;
                    ok = col_corners<T, FOOT>(g, v, ix, iy, p0, p1, p2, p3, ilo, ihi, amp, r, d);
                    ok = ok && ihi >= iu_a && ilo <= iu_b;
                    cp.cells = ilo | (ihi << 16);
                    T ampr = T(0);
                    if (ok) {
                        ok = v_window(g, d, v.wz, cp.zinc, cp.vstart, iz0, iz1) && iz0 < iz1 && cp.zinc > T(0);
                        ampr = amp * frcp(r);
                        cp.r2 = r * r;
                    }
                    if (ok) {
                        // voxel range needed by this wave's rows (exact ranges; a voxel left
                        // out at the rounding edge has a weight of the order of the rounding)
                        const T rz = frcp(cp.zinc);
                        const int za = max((int)ceil((T(iv_a) - T(0.5) - cp.vstart) * rz - T(1)), iz0);
                        const int zb = min((int)ceil((T(iv_last) + T(0.5) - cp.vstart) * rz) - 1, iz1 - 1);
                        cp.zab = za | (zb << kZBits) | (gather_class(cp.zinc) << (2 * kZBits));
                        ok = za <= zb;
                    }
                    if (ok) {
                        // tile weights of the first four footprint cells inside the tile,
                        // times amp / r
                        const int c0 = max(ilo, iu_a);
                        cp.w0 = u_weight<T, FOOT>(c0 + 0, ilo, ihi, p0, p1, p2, p3) * ampr;
                        cp.w1 = u_weight<T, FOOT>(c0 + 1, ilo, ihi, p0, p1, p2, p3) * ampr;
                        cp.w2 = u_weight<T, FOOT>(c0 + 2, ilo, ihi, p0, p1, p2, p3) * ampr;
                        cp.w3 = u_weight<T, FOOT>(c0 + 3, ilo, ihi, p0, p1, p2, p3) * ampr;
                        cp.col = iy * g.nx + ix;
                    }
                }
                // each lane keeps its candidate's parameters in registers; the column loop
                // reads those of the column it walks with readlane (scalar registers)
                unsigned long long mask = __ballot(ok);
                // software pipeline over the surviving columns: the loads of column i+1 are
                // in flight while column i is staged and gathered.  All three loads are
                // issued for every column (a buffer resource over voxels [za, zb] of the
                // column, wave-uniform base and range, so loads past zb read zero): loads
                // under a condition would make the compiler wait for them at once.
                int cur = mask ? __ffsll((long long)mask) - 1 : -1;
                if (cur >= 0) mask &= mask - 1;
                Tin pre0{}, pre1{}, pre2{};
                auto prefetch = [&](int c) {
                    const int zabn = rdl_i(cp.zab, c), zan = zabn & kZMask, zbn = (zabn >> kZBits) & kZMask;
                    const auto rs = buf_rsrc(vol + (size_t)rdl_i(cp.col, c) * g.nz + zan,
                                             (zbn - zan + 1) * (int)sizeof(Tin));
                    pre0 = buf_ld<Tin>(rs, loff);
                    pre1 = buf_ld<Tin>(rs, loff + kWave * (int)sizeof(Tin));
                    pre2 = buf_ld<Tin>(rs, loff + 2 * kWave * (int)sizeof(Tin));
                };
                if (cur >= 0) prefetch(cur);
                while (cur >= 0) {
                    const int me = cur;  // cur moves on to the next column inside the loop
                    const int zab = rdl_i(cp.zab, me), cells = rdl_i(cp.cells, me);
                    const int za = zab & kZMask, zb = (zab >> kZBits) & kZMask, cls = zab >> (2 * kZBits);
                    const int ilo = cells & 0xffff, ihi = cells >> 16;
                    const T r2 = rdl(cp.r2, me), zinc = rdl(cp.zinc, me), vstart = rdl(cp.vstart, me), rz = frcp(zinc);
                    const int iv = iv_a + lane;
                    // row iv against the column's voxels: voxel j covers detector rows
                    // [x_j, x_j + zinc] relative to the row's lower edge, with
                    // x_j = j zinc - (iv - 1/2 - vstart); at most ceil(1/zinc) + 1 of them
                    // overlap the row, so thick voxels take a fixed unrolled count (lanes
                    // past the last row compute a sum nobody stores; their reads stay in
                    // the window through the clamp of ja)
                    const T tt = fiv_lo - vstart;
                    const int ja = (int)ceil(tt * rz - T(1));
                    auto gather = [&](int zw0, int zw1) -> T {
                        if (cls == 3) {
                            // footprints that may reach the 16-cell cap, and thin voxels
                            T s = T(0);
                            if (iv <= iv_last) {
                                const int jb = min((int)ceil((T(iv) + T(0.5) - vstart) * rz) - 1, zw1);
                                for (int jj = max(ja, zw0); jj <= jb; ++jj) {
                                    const T vmin = vstart + T(jj) * zinc;
                                    s += (zinc > T(14) ? dd_overlap16(iv, vmin, vmin + zinc, g.nv)
                                                       : dd_overlap(iv, vmin, vmin + zinc)) * zw[jj - zw0 + 1];
                                }
                            }
                            return s;
                        }
                        if (cls == 0) return gather_row<T, 2>(zw, zw0, zw1, ja, tt, zinc);
                        if (cls == 1) return gather_row<T, 3>(zw, zw0, zw1, ja, tt, zinc);
                        return gather_row<T, 4>(zw, zw0, zw1, ja, tt, zinc);
                    };
                    auto next = [&]() {
                        cur = mask ? __ffsll((long long)mask) - 1 : -1;
                        if (cur >= 0) {
                            mask &= mask - 1;
                            prefetch(cur);
                        }
                    };
                    T q = T(0);
                    const int ns = (zb - za) >> 6;  // 64-voxel slots the range needs, minus one
                    const T zbase = T(za) * g.dz + wzdz;
                    if (zb - za < kFastSpan) {
                        // stage the slots the range needs (lanes past zb hold the zeros the
                        // range-checked loads returned) and four zero words after them
                        {
                            const T zc = zbase + lzd0;
                            zw[1 + lane] = fsqrt(fma_t(zc, zc, r2)) * T(to_f(pre0));
                        }
                        if (ns >= 1) {
                            const T zc = zbase + lzd1;
                            zw[1 + kWave + lane] = fsqrt(fma_t(zc, zc, r2)) * T(to_f(pre1));
                        }
                        if (ns >= 2) {
                            const T zc = zbase + lzd2;
                            zw[1 + 2 * kWave + lane] = fsqrt(fma_t(zc, zc, r2)) * T(to_f(pre2));
                        }
                        if (lane < kZPad) zw[1 + (ns + 1) * kWave + lane] = T(0);
                        __builtin_amdgcn_wave_barrier();
                        next();  // the next column's loads fly while this one is gathered
                        q = gather(za, zb);
                        __builtin_amdgcn_wave_barrier();
                    } else {
                        const Tin* colp = vol + (size_t)rdl_i(cp.col, me) * g.nz;
                        for (int zw0 = za; zw0 <= zb; zw0 += kZWin) {
                            const int zw1 = min(zb, zw0 + kZWin - 1);
                            // stage sqrt(r^2 + zc^2) * img of voxels [zw0, zw1] in LDS: the
                            // first window from the registers loaded one column earlier, the
                            // rest directly
                            int zfirst = zw0 + lane;
                            if (zw0 == za) {
                                const Tin pv[kPre] = {pre0, pre1, pre2};
#pragma unroll
                                for (int i = 0; i < kPre; ++i) {
                                    const int zz = za + i * kWave + lane;
                                    const T zc = T(zz) * g.dz + wzdz;
                                    zw[zz - zw0 + 1] = fsqrt(fma_t(zc, zc, r2)) * T(to_f(pv[i]));
                                }
                                zfirst = za + kPre * kWave + lane;
                            }
                            for (int zz = zfirst; zz <= zw1; zz += kWave) {
                                const T zc = T(zz) * g.dz + wzdz;
                                zw[zz - zw0 + 1] = fsqrt(fma_t(zc, zc, r2)) * ld<T>(colp + zz);
                            }
                            if (lane < kZPad) zw[zw1 - zw0 + 2 + lane] = T(0);  // zero pads
                            __builtin_amdgcn_wave_barrier();
                            if (zw0 == za) next();
                            q += gather(zw0, zw1);
                            __builtin_amdgcn_wave_barrier();
                        }
                    }
                    // first and last tile cell of the footprint; the same on every lane, so
                    // the weighted adds touch one LDS word per lane and bank.  The first four
                    // cells use the weights made with the column's parameters (scalar
                    // registers); footprints wider than four tile cells (rare) evaluate the
                    // further ones here.
                    const int c0 = max(ilo, iu_a) - iu_a;
                    const int c1 = min(ihi, iu_b) - iu_a;
                    {
                        T* pa = acc + c0 * kWave;
                        pa[0] += rdl(cp.w0, me) * q;
                        if (c0 + 1 <= c1) pa[kWave] += rdl(cp.w1, me) * q;
                        if (c0 + 2 <= c1) pa[2 * kWave] += rdl(cp.w2, me) * q;
                        if (c0 + 3 <= c1) pa[3 * kWave] += rdl(cp.w3, me) * q;
                    }
                    if (c0 + 4 <= c1) {
                        const int col = rdl_i(cp.col, me);
                        T p0, p1, p2, p3, amp, r, d;
                        int jlo, jhi;
                        col_corners<T, FOOT>(g, v, col % g.nx, col / g.nx, p0, p1, p2, p3, jlo, jhi, amp, r, d);
                        const T ampr = amp * frcp(r);
                        for (int cc = c0 + 4; cc <= c1; ++cc)
                            acc[cc * kWave] += u_weight<T, FOOT>(iu_a + cc, ilo, ihi, p0, p1, p2, p3) * ampr * q;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (SPLIT > 1) {
        // fixed-order combine of the shares by wave 0: ((s0 + s1) + s2) + s3
        __syncthreads();
        if (wv > 0) return;
    }
    Tout* out = proj + (size_t)b * g.nu * g.nv;
    const int iv = iv_a + lane;
    if (iv <= iv_last) {
#pragma unroll
        for (int t = 0; t < TU; ++t) {
            const int iu = iu_a + t;
            T sum = acc[t * kWave];
#pragma unroll
            for (int w = 1; w < SPLIT; ++w) sum += accs[w][t][lane];
            if (iu < g.nu) out[(size_t)iu * g.nv + iv] = cvt_store<Tout>(sum);
        }
    }
}

// LDS ring of the ring form: ring_len(NS) entries per wave, one per column.  An entry
// holds a zero word, the NS * 64 words a column's NS loads write (voxels za ..
// za + 64 NS - 1; MUBUF loads with the LDS bit put lane t's word at the entry address
// + 4 t, and lanes past the buffer range write zero) and kZPad zero words.  Only the loads
// write words 1 .. 64 NS, so word 0 and the pads stay zero for the whole kernel.  Columns
// of at most 128 voxels (NS = 2) take four entries (three columns in flight), columns of
// at most 192 (NS = 3) three: about the same LDS.  Columns of at most 256 (NS = 4) take
// two (one column in flight): with three, the ring's LDS and 91 registers held the kernel
// to five waves per SIMD, and the sixth wave measured 7-8% faster than the deeper ring
// (720-view 512^2 x 256 SF forward 154 -> 143 ms, DD 144 -> 132 ms).
constexpr int kRingWaves = 6;  // waves per SIMD the ring kernel is compiled for (the LDS of six
                               // blocks per CU; seven would spill registers)
__host__ __device__ constexpr int ring_len(int ns) { return ns == 2 ? 4 : ns == 4 ? 2 : 3; }
__host__ __device__ constexpr int ring_entry_words(int ns) { return 1 + ns * kWave + kZPad; }

// A buffer resource as four scalar words, for the inline assembly below.  The loads are
// issued from inline assembly because the compiler's wait insertion, which cannot tell the
// ring's slots apart, would otherwise wait for every outstanding LDS load before any LDS
// read of the ring; the kernel's explicit vmcnt waits order each slot's reads after its
// own loads instead.
typedef int rsrc_words __attribute__((ext_vector_type(4)));
__device__ __forceinline__ rsrc_words dma_rsrc(const void* base, int nbytes) {
    const unsigned long long a = (unsigned long long)base;
    rsrc_words r;
    r.x = (int)(unsigned)(a & 0xffffffffull);
    r.y = (int)(unsigned)((a >> 32) & 0xffffull);  // stride 0
    r.z = nbytes;                                  // range: loads at or past it read zero
    r.w = 0x00020000;
    return r;
}

// LDS byte address of a shared-memory pointer, in a scalar register
template <class P>
__device__ __forceinline__ unsigned lds_addr(P* p) {
    return (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(size_t)(__attribute__((address_space(3))) P*)p);
}

// wait until at most N of this wave's vector-memory operations are outstanding; the
// memory clobber keeps the compiler from moving LDS accesses across the wait
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// ---------------------------------------------------------------------------------
// Forward projection, ring form (FP32 volumes with nz <= kFastSpan: every column range
// fits three 64-voxel slots).  The same work items, strips and footprints as
// cbct_proj_kernel; what differs is how a wave walks the surviving columns of a batch:
//
//  * the survivors' parameters are moved to lanes 0 .. n-1 (ds_permute), so column k's
//    are read with readlane(k) and the walk needs no bit-mask bookkeeping;
//  * each candidate's lane precomputes what the walk would otherwise derive in scalar
//    registers: the byte address and length of its voxel range (the buffer resource),
//    the lower edge of its first voxel on the detector, zc of its first voxel, its tile
//    cells and gather class;
//  * a column's NS 64-voxel loads go straight into an LDS ring entry (MUBUF loads with
//    the LDS bit), ring_len(NS) - 1 columns ahead of the one being gathered.  Every column
//    issues exactly one group of NS loads (groups past the batch's last survivor have an
//    empty range and write zeros), so "this column's group has landed" is always the same
//    vmcnt: NS (ring_len(NS) - 1).
// ---------------------------------------------------------------------------------
struct RingCol {
    int alo;        // low word of the byte address of the column's first needed voxel
    // gather class | slots - 1 << 2 | cells past the first << 4 | first tile cell << 9 |
    // zb - za << 16 | (high word of that address - the volume's) << 24
    int meta;
    float zbase;    // zc of voxel za: za dz + wz dz
    float r2, zinc, rz, vst;  // r^2, magnified slice pitch, its reciprocal, lower edge of voxel za on the detector (rows)
    float w0, w1, w2, w3;  // tile weights of the first four footprint cells, times amp / r
    int col;        // image column (wide footprints only)
};

__device__ __forceinline__ int permute_i(int dst_lane_bytes, int x) {
    return __builtin_amdgcn_ds_permute(dst_lane_bytes, x);
}
__device__ __forceinline__ float permute_f(int dst_lane_bytes, float x) {
    return __int_as_float(__builtin_amdgcn_ds_permute(dst_lane_bytes, __float_as_int(x)));
}

// NS 64-lane dword loads of one column into the ring entry at LDS byte address la (the
// instruction offsets 0, 256, 512 apply to both the voxel and the LDS address)
template <int NS>
__device__ __forceinline__ void ring_load(rsrc_words rs, unsigned la, int voff) {
    static_assert(NS >= 1 && NS <= 4, "one to four slots per column");
    if constexpr (NS == 1)
        asm volatile(
            "s_mov_b32 m0, %1\n\t"
            "s_nop 0\n\t"
            "buffer_load_dword %0, %2, 0 offen lds"
            :
            : "v"(voff), "s"(la), "s"(rs)
            : "memory");
    else if constexpr (NS == 2)
        asm volatile(
            "s_mov_b32 m0, %1\n\t"
            "s_nop 0\n\t"
            "buffer_load_dword %0, %2, 0 offen lds\n\t"
            "buffer_load_dword %0, %2, 0 offen offset:256 lds"
            :
            : "v"(voff), "s"(la), "s"(rs)
            : "memory");
    else if constexpr (NS == 4)
        asm volatile(
            "s_mov_b32 m0, %1\n\t"
            "s_nop 0\n\t"
            "buffer_load_dword %0, %2, 0 offen lds\n\t"
            "buffer_load_dword %0, %2, 0 offen offset:256 lds\n\t"
            "buffer_load_dword %0, %2, 0 offen offset:512 lds\n\t"
            "buffer_load_dword %0, %2, 0 offen offset:768 lds"
            :
            : "v"(voff), "s"(la), "s"(rs)
            : "memory");
    else
        asm volatile(
            "s_mov_b32 m0, %1\n\t"
            "s_nop 0\n\t"
            "buffer_load_dword %0, %2, 0 offen lds\n\t"
            "buffer_load_dword %0, %2, 0 offen offset:256 lds\n\t"
            "buffer_load_dword %0, %2, 0 offen offset:512 lds"
            :
            : "v"(voff), "s"(la), "s"(rs)
            : "memory");
}

// sum over the K voxels from ja = ceil(tt / zinc - 1) on (clamped to [-1, span + 1]) of
// (overlap with the row) * staged value, voxel j covering [x_j, x_j + zinc] relative to the
// row's lower edge (x_j = j zinc - tt); zs[1 + j] holds voxel j of the range, zs[0] and the
// words past the range are zero.  jhi = span + 1.
//
// With c_k = clamp01(x_ja + k zinc) the sum is sum_{k < K} (c_{k+1} - c_k) v_k, which
// telescopes to c_K v_{K-1} - c_0 v_0 + sum_{0 < k < K} c_k (v_{k-1} - v_k).  ja puts x_ja
// in [-zinc, 0), so c_0 = 0, and the class's K has (K - 1) zinc >= 1, so c_K = 1: the sum is
// v_{K-1} + sum_{0 < k < K} c_k (v_{k-1} - v_k), 3 (K - 1) operations instead of 3 K + 1.
// Where ja is clamped the two hold as well or their values are zero words: below the range
// (ja = -1, x_ja >= 0) v_0 is the zero word and c_K = 1 still; above it every value read is
// zero.  (At the rounding edges of ja the first or last overlap moves by an ulp of the row.)
template <int K>
__device__ __forceinline__ float ring_gather(const float* zs, float jhi, float tt, float rz, float zinc) {
    const float jf = __builtin_amdgcn_fmed3f(ceilf(fmaf(tt, rz, -1.f)), -1.f, jhi);
    const float* pz = zs + ((int)jf + 1);
    const float x0 = fmaf(jf, zinc, -tt);
    float s = pz[K - 1];
#pragma unroll
    for (int k = 1; k < K; ++k) s = fmaf(clamp01(fmaf(float(k), zinc, x0)), pz[k - 1] - pz[k], s);
    return s;
}

// WPB waves per block; SPLIT = WPB shares each item's strip among them (eight waves: short
// view subsets make few items, and smaller shares even out the blocks' run times).
//
// HALF: an FP16 volume (the mixed PWLS form's projector input z - x0).  A slot of the ring
// is then 64 dwords holding 128 voxels (NS slots: up to 128 NS voxels per column), so the
// ring can be RING entries deep in the LDS of the FP32 form's four; the voxel range of a
// column is widened to whole dwords (an even first voxel: columns of an even nz start on a
// dword), and the scaling pass converts each lane's two voxels and writes them, scaled, as
// FP32 to a staging buffer that the row gather reads as the FP32 form reads its slot.  The
// one or two voxels the widening adds lie outside every row of the wave and weigh zero.
//
// RPL: detector rows per lane (RPL 64-row chunks per item): a column's voxel range, loads,
// scaling and weights serve all of the item's rows, so tall detectors pay the per-column
// set-up once per RPL chunks (with narrower tiles, so that the accumulators' LDS stays).
template <class Tout, int FOOT, int TU, int SPLIT, int NS, int WPB, bool HALF = false, int RING = ring_len(NS),
          int RPL = 1>
__global__ __launch_bounds__(WPB * kWave, kRingWaves) void cbct_proj_ring_kernel(Geom<float> g, const View<float>* __restrict__ views,
                                                                   int vstride, int nviews,
                                                                   const typename std::conditional<HALF, __half, float>::type* __restrict__ vol,
                                                                   const uint8_t* __restrict__ map,
                                                                   Tout* __restrict__ proj, int exhaustive) {
    using T = float;
    using Tin = typename std::conditional<HALF, __half, float>::type;
    static_assert(SPLIT == 1 || SPLIT == WPB, "SPLIT is 1 or the block's wave count");
    static_assert(HALF ? (NS == 1 || NS == 2) : (NS >= 2 && NS <= 4), "slots per column");
    static_assert((HALF ? 2 : 1) * NS * kWave <= 256 && TU <= 128, "a column's span and first cell fit its meta word");
    // one block of LDS: the waves' rings, (HALF) their staging buffers, then their tile
    // accumulators.  Accumulators: each lane's row of the tile, [cell][lane] (one word per
    // bank, one writer per word); a column's update reads and writes only its footprint's
    // cells inside the tile.  The rings come first so that their addresses fit the LDS
    // instructions' offset fields.
    constexpr int kRing = RING;
    constexpr int kSlotW = HALF ? NS * kWave : ring_entry_words(NS);
    // staging (HALF): an unused word, the zero word, 128 NS voxels, kZPad zero words; even
    // length, so that voxel pairs sit on 8-byte boundaries
    constexpr int kStageW = HALF ? 2 + 2 * NS * kWave + kZPad : 0;
    static_assert(!HALF || (kStageW % 2 == 0 && (kRing * kSlotW) % 2 == 0), "staging pairs on 8-byte boundaries");
    constexpr int kCellW = kWave * RPL;  // accumulator words per tile cell: [cell][row chunk][lane]
    __shared__ __align__(16) T lds_block[WPB * kRing * kSlotW + WPB * kStageW + WPB * TU * kCellW];
    T* const accs = lds_block + WPB * (kRing * kSlotW + kStageW);
    const int nut = (g.nu + TU - 1) / TU;
    const int nvc = (g.nv + kCellW - 1) / kCellW;
    const int items = nviews * nut * nvc;
    const int bid = blockIdx.x;
    const int wv = (int)(threadIdx.x / kWave);
    const int item = SPLIT == 1 ? bid * WPB + wv : bid;
    const int share = SPLIT == 1 ? 0 : wv;
    if (item >= items) return;  // SPLIT > 1: the whole block leaves together
    const int vc = item % nvc;
    const int b = (item / nvc) % nviews;
    // longest work first (tiles in order of their distance from the central ray's tile)
    const int ut = tile_by_rank(item / (nvc * nviews), min(max((int)floor(-g.wu / T(TU)), 0), nut - 1), nut);
    const int lane = lane_id();
    const int loff = lane * 4;
    const int iv_a = vc * kCellW;
    const int iu_a = ut * TU;
    const int iu_b = min(iu_a + TU, g.nu) - 1;
    const View<T> v = views[(size_t)b * vstride];
    const bool row_major = fabs_t(v.c) >= fabs_t(v.s);
    const int m_end = row_major ? g.ny : g.nx;
    const T kL = u_to_slope(g, T(iu_a) - T(0.5));
    const T kR = u_to_slope(g, T(iu_b) + T(0.5));
    const int iv = iv_a + lane;  // this lane's row of the first chunk (chunk r: iv + 64 r)
    const int iv_last = min(iv_a + kCellW, g.nv) - 1;
    T* zr = lds_block + wv * kRing * kSlotW;
    // FP32: word 0 and the pads of every slot stay zero; the loads write words 1 .. 192.
    // HALF: the staging buffer's zero word and pads stay zero; the scaling pass writes the rest
    T* const zstage = lds_block + WPB * kRing * kSlotW + wv * kStageW + 1;  // its zero word
    for (int i = lane; i < kRing * kSlotW; i += kWave) zr[i] = T(0);
    if constexpr (HALF)
        for (int i = lane; i < kStageW - 1; i += kWave) zstage[i] = T(0);
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // zeros in place before any load lands
    const unsigned ring_lds = lds_addr(zr);
    const unsigned slot_bytes = kSlotW * 4;
    const unsigned ring_end = ring_lds + kRing * slot_bytes;
    const T wzdz = v.wz * g.dz;
    const T fiv_lo = T(iv) - T(0.5);  // lower edge of this lane's row
    // zc of the voxels a lane scales, relative to the range's first: FP32 lane + 64 k,
    // HALF 2 lane (+ 1) + 128 k
    const T lzd0 = T(HALF ? 2 * lane : lane) * g.dz, lzd1 = T(kWave + lane) * g.dz,
            lzd2 = T(2 * kWave + lane) * g.dz, lzd3 = T(3 * kWave + lane) * g.dz;
    const T lzh1 = T(2 * kWave + 2 * lane) * g.dz;
    const rsrc_words empty_rs = dma_rsrc(vol, 0);  // a group of loads that only writes zeros

    T* acc = &accs[(wv * TU) * kCellW + lane];
#pragma unroll
    for (int t = 0; t < TU * RPL; ++t) acc[t * kWave] = T(0);

    // ring position of the next group to issue (byte address of its slot's word 1; HALF: of
    // the slot's first dword)
    unsigned isl = ring_lds + (HALF ? 0 : 4);
    auto advance = [&](unsigned a) { return a + slot_bytes >= ring_end ? a + slot_bytes - kRing * slot_bytes : a + slot_bytes; };

    // the SPLIT waves of a block take every SPLIT-th image line of the strip (lane l of a
    // batch takes line m0 + l SPLIT + share), so that their shares are alike wherever the
    // strip is long: the block holds its LDS until its last wave ends
    for (int m0 = 0; m0 < m_end; m0 += SPLIT * kWave) {
        const int m = m0 + lane * SPLIT + share;
        int lo = 0, hi = -1;
        if (m < m_end) {
            if (!strip_range(g, v, row_major, m, kL, kR, exhaustive != 0, lo, hi)) hi = lo - 1;
        }
        const int cnt = hi - lo + 1 > 0 ? hi - lo + 1 : 0;
        const int incl = wave_incl_scan(cnt);
        const int excl = incl - cnt;
        const int total = rdl_i(incl, kWave - 1);
        for (int k0 = 0; k0 < total; k0 += kWave) {
            const int k = k0 + lane;
            int j = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1) {
                const int e = __shfl(excl, j + step, 64);
                if (e <= k) j += step;
            }
            const int lo_j = __shfl(lo, j, 64);
            const int ex_j = __shfl(excl, j, 64);
            const int minor = lo_j + (k - ex_j);
            const int ix = row_major ? minor : m0 + j * SPLIT + share;
            const int iy = row_major ? m0 + j * SPLIT + share : minor;

            // exact footprint of the candidate column, its weights on this tile and the
            // quantities the walk needs
            bool ok = k < total;
            RingCol cp{};
            // the map byte is loaded first and tested after the footprint, so that its load
            // overlaps the footprint arithmetic (lanes it rejects would idle through that code)
            int in_map = 1;
            if (ok && map != nullptr) in_map = map[iy * g.nx + ix];
            if (ok) {
                T p0, p1, p2, p3, amp, r, d;
                int ilo = 0, ihi = -1, iz0 = 0, iz1 = 0;
                ok = col_corners<T, FOOT>(g, v, ix, iy, p0, p1, p2, p3, ilo, ihi, amp, r, d);
                ok = ok && ihi >= iu_a && ilo <= iu_b;
                T zinc = T(1), vstart = T(0), rz = T(1);
                if (ok) ok = v_window(g, d, v.wz, zinc, vstart, iz0, iz1) && iz0 < iz1 && zinc > T(0);
                int za = 0, zb = -1;
                if (ok) {
                    // voxel range needed by this wave's rows (exact ranges; a voxel left out at
                    // the rounding edge has a weight of the order of the rounding)
                    rz = frcp(zinc);
                    za = max((int)ceil((T(iv_a) - T(0.5) - vstart) * rz - T(1)), iz0);
                    zb = min((int)ceil((T(iv_last) + T(0.5) - vstart) * rz) - 1, iz1 - 1);
                    ok = za <= zb;
                }
                if (ok) {
                    if constexpr (HALF) {
                        // whole dwords: an even first voxel and an odd last one (nz is even)
                        za &= ~1;
                        zb |= 1;
                    }
                    const T ampr = amp * frcp(r);
                    const int c0 = max(ilo, iu_a);
                    const int c1 = min(ihi, iu_b);
                    cp.w0 = u_weight<T, FOOT>(c0 + 0, ilo, ihi, p0, p1, p2, p3) * ampr;
                    cp.w1 = u_weight<T, FOOT>(c0 + 1, ilo, ihi, p0, p1, p2, p3) * ampr;
                    cp.w2 = u_weight<T, FOOT>(c0 + 2, ilo, ihi, p0, p1, p2, p3) * ampr;
                    cp.w3 = u_weight<T, FOOT>(c0 + 3, ilo, ihi, p0, p1, p2, p3) * ampr;
                    const int span = zb - za;
                    const unsigned long long a =
                        (unsigned long long)(vol + ((size_t)(iy * g.nx + ix) * g.nz + za));
                    cp.alo = (int)(unsigned)(a & 0xffffffffull);
                    const unsigned ahi_off = (unsigned)(a >> 32) - (unsigned)(empty_rs.y);
                    cp.meta = gather_class(zinc) | ((span >> (HALF ? 7 : 6)) << 2) | ((c1 - c0) << 4) |
                              ((c0 - iu_a) << 9) | (span << 16) | (int)(ahi_off << 24);
                    cp.zbase = T(za) * g.dz + wzdz;
                    cp.r2 = r * r;
                    cp.zinc = zinc;
                    cp.rz = rz;
                    cp.vst = vstart + T(za) * zinc;
                    cp.col = iy * g.nx + ix;
                }
            }
            ok = ok && in_map != 0;
            // survivors to lanes 0 .. n-1 in lane order, the others after them
            const unsigned long long mask = __ballot(ok);
            const int n = __popcll(mask);
            if (n == 0) continue;
            {
                const unsigned long long below = mask & ((1ull << lane) - 1ull);
                const int rank = __popcll(below);
                const int dst = (ok ? rank : n + (lane - rank)) * 4;
                cp.alo = permute_i(dst, cp.alo);
                cp.meta = permute_i(dst, cp.meta);
                cp.zbase = permute_f(dst, cp.zbase);
                cp.r2 = permute_f(dst, cp.r2);
                cp.zinc = permute_f(dst, cp.zinc);
                cp.rz = permute_f(dst, cp.rz);
                cp.vst = permute_f(dst, cp.vst);
                cp.w0 = permute_f(dst, cp.w0);
                cp.w1 = permute_f(dst, cp.w1);
                cp.w2 = permute_f(dst, cp.w2);
                cp.w3 = permute_f(dst, cp.w3);
                cp.col = permute_i(dst, cp.col);
            }
            // group c's loads into the next ring slot; past the last survivor, a group that
            // writes zeros, so that every column has exactly two groups issued after it
            auto issue = [&](int c) {
                rsrc_words rs = empty_rs;
                if (c < n) {
                    const unsigned mt = (unsigned)rdl_i(cp.meta, c);
                    rs.x = rdl_i(cp.alo, c);
                    rs.y = empty_rs.y + (int)(mt >> 24);
                    rs.z = (int)(((mt >> 16) & 255u) + 1u) * (int)sizeof(Tin);
                }
                ring_load<NS>(rs, isl, loff);
                isl = advance(isl);
            };
#pragma unroll
            for (int i = 0; i < kRing - 1; ++i) issue(i);
            unsigned psl = isl;  // slot of column 0: kRing - 1 slots behind the issue position
            psl = advance(psl);
            for (int c = 0; c < n; ++c) {
                issue(c + kRing - 1);
                wait_vm<NS * (kRing - 1)>();  // column c's group has landed
                const unsigned slot_off = psl - (HALF ? 0 : 4) - ring_lds;  // the slot's first word
                T* zs = HALF ? zstage : (T*)((char*)zr + slot_off);         // word 0: zero
                psl = advance(psl);
                const int meta = rdl_i(cp.meta, c);
                const int cls = meta & 3, ns = (meta >> 2) & 3, nc = (meta >> 4) & 31, c0 = (meta >> 9) & 127;
                const T r2 = rdl(cp.r2, c), zinc = rdl(cp.zinc, c), vst = rdl(cp.vst, c), zbase = rdl(cp.zbase, c);
                const T rz = rdl(cp.rz, c);  // frcp(zinc), formed by the candidate's lane
                if constexpr (HALF) {
                    // each lane's dword is voxels 2 lane, 2 lane + 1 of a slot (zero past the
                    // range): converted, scaled by sqrt(r^2 + zc^2) and stored as an FP32 pair
                    const unsigned* slot = (const unsigned*)((const char*)zr + slot_off);
#pragma unroll
                    for (int k = 0; k < NS; ++k) {
                        // every slot, in range or not: the staging buffer must hold this
                        // column's zeros past its range, not the previous column's voxels
                        const unsigned h2 = slot[k * kWave + lane];
                        const T v0 = __half2float(__ushort_as_half((unsigned short)(h2 & 0xffffu)));
                        const T v1 = __half2float(__ushort_as_half((unsigned short)(h2 >> 16)));
                        const T zc0 = zbase + (k == 0 ? lzd0 : lzh1);
                        const T zc1 = zc0 + g.dz;
                        float2 o;
                        o.x = fsqrt(fmaf(zc0, zc0, r2)) * v0;
                        o.y = fsqrt(fmaf(zc1, zc1, r2)) * v1;
                        *(float2*)(zs + 1 + 2 * (k * kWave + lane)) = o;
                    }
                } else {
                    // scale the loaded voxels in place by sqrt(r^2 + zc^2) (the cone factor times
                    // r); words past the range are zero.  The first two slots in one straight
                    // pass (one LDS round trip; a slot past the range holds zeros and stays zero)
                    {
                        const T v0 = zs[1 + lane], v1 = zs[1 + kWave + lane];
                        const T zc0 = zbase + lzd0, zc1 = zbase + lzd1;
                        zs[1 + lane] = fsqrt(fmaf(zc0, zc0, r2)) * v0;
                        zs[1 + kWave + lane] = fsqrt(fmaf(zc1, zc1, r2)) * v1;
                    }
                    if (NS >= 3 && ns >= 2) {
                        const T zc = zbase + lzd2;
                        zs[1 + 2 * kWave + lane] = fsqrt(fmaf(zc, zc, r2)) * zs[1 + 2 * kWave + lane];
                    }
                    if (NS >= 4 && ns >= 3) {
                        const T zc = zbase + lzd3;
                        zs[1 + 3 * kWave + lane] = fsqrt(fmaf(zc, zc, r2)) * zs[1 + 3 * kWave + lane];
                    }
                }
                __builtin_amdgcn_wave_barrier();
                // row iv against the column's voxels (relative to voxel za): voxel j covers
                // rows [vst + j zinc, vst + (j + 1) zinc]; at most ceil(1/zinc) + 1 of them
                // overlap a row, so thick voxels take a fixed unrolled count
                const int span = (meta >> 16) & 255;
                const T w0 = rdl(cp.w0, c), w1 = rdl(cp.w1, c), w2 = rdl(cp.w2, c), w3 = rdl(cp.w3, c);
                const T jhi = T(span + 1);
#pragma unroll
                for (int rr = 0; rr < RPL; ++rr) {
                    const int ivr = iv + rr * kWave;
                    const T tt = (fiv_lo + T(rr * kWave)) - vst;
                    T q;
                    if (cls == 1) {
                        q = ring_gather<3>(zs, jhi, tt, rz, zinc);
                    } else if (cls == 2) {
                        q = ring_gather<4>(zs, jhi, tt, rz, zinc);
                    } else if (cls == 0) {
                        q = ring_gather<2>(zs, jhi, tt, rz, zinc);
                    } else {
                        // footprints that may reach the 16-cell cap, and thin voxels
                        q = T(0);
                        const int ja = (int)ceil(tt * rz - T(1));
                        if (ivr <= iv_last) {
                            const int jb = min((int)ceil((T(ivr) + T(0.5) - vst) * rz) - 1, span);
                            for (int jj = max(ja, 0); jj <= jb; ++jj) {
                                const T vmin = vst + T(jj) * zinc;
                                q += (zinc > T(14) ? dd_overlap16(ivr, vmin, vmin + zinc, g.nv)
                                                   : dd_overlap(ivr, vmin, vmin + zinc)) * zs[jj + 1];
                            }
                        }
                    }
                    // tile cells c0 .. c0 + nc; the same on every lane, so each weighted add
                    // touches one LDS word per lane and bank.  One branch per footprint width
                    // (uniform), each reading its cells together and writing them back: one LDS
                    // round trip, and no cell outside the footprint is read or written.
                    T* pa = acc + c0 * kCellW + rr * kWave;
                    if (nc == 0) {
                        pa[0] = pa[0] + w0 * q;
                    } else if (nc == 1) {
                        const T a0 = pa[0], a1 = pa[kCellW];
                        pa[0] = a0 + w0 * q;
                        pa[kCellW] = a1 + w1 * q;
                    } else if (nc == 2) {
                        const T a0 = pa[0], a1 = pa[kCellW], a2 = pa[2 * kCellW];
                        pa[0] = a0 + w0 * q;
                        pa[kCellW] = a1 + w1 * q;
                        pa[2 * kCellW] = a2 + w2 * q;
                    } else {
                        const T a0 = pa[0], a1 = pa[kCellW], a2 = pa[2 * kCellW], a3 = pa[3 * kCellW];
                        pa[0] = a0 + w0 * q;
                        pa[kCellW] = a1 + w1 * q;
                        pa[2 * kCellW] = a2 + w2 * q;
                        pa[3 * kCellW] = a3 + w3 * q;
                    }
                    if (nc >= 4) {
                        // wider than four tile cells (rare): the corners and cells again, from
                        // the column (the candidate's own evaluation, so the same values)
                        const int col = rdl_i(cp.col, c);
                        T p0, p1, p2, p3, amp, r, d;
                        int ilo, ihi;
                        col_corners<T, FOOT>(g, v, col % g.nx, col / g.nx, p0, p1, p2, p3, ilo, ihi, amp, r, d);
                        const T ampr = amp * frcp(r);
                        for (int cc = c0 + 4; cc <= c0 + nc; ++cc)
                            acc[cc * kCellW + rr * kWave] += u_weight<T, FOOT>(iu_a + cc, ilo, ihi, p0, p1, p2, p3) * ampr * q;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    wait_vm<0>();  // the trailing zero groups have landed before the ring is released
    if (SPLIT > 1) {
        // fixed-order combine of the shares by wave 0: ((s0 + s1) + s2) + ...
        __syncthreads();
        if (wv > 0) return;
    }
    Tout* out = proj + (size_t)b * g.nu * g.nv;
#pragma unroll
    for (int rr = 0; rr < RPL; ++rr) {
        const int ivr = iv + rr * kWave;
        if (ivr > iv_last) break;
#pragma unroll
        for (int t = 0; t < TU; ++t) {
            const int iu = iu_a + t;
            T sum = acc[t * kCellW + rr * kWave];
#pragma unroll
            for (int w = 1; w < SPLIT; ++w) sum += accs[(w * TU + t) * kCellW + rr * kWave + lane];
            if (iu < g.nu) out[(size_t)iu * g.nv + ivr] = cvt_store<Tout>(sum);
        }
    }
}

// ---------------------------------------------------------------------------------
// Matched backprojection: one wavefront per image column, NZL voxels per lane (z
// fastest).  The views' footprints of the column are evaluated 64 or 32 at a time (one view
// per lane) into LDS.  Per view the wave then combines the footprint's detector columns
// row by row, s(iv) = sum_u w_u proj(u, iv), into LDS for the rows the column covers
// (the next view's rows are already loading), and every voxel gathers its v footprint
// from LDS.  All views accumulate in registers; one store per voxel.
// ---------------------------------------------------------------------------------
// views whose footprint records a wave keeps in LDS at a time in the windowed form (tall
// detectors): 32 of them (lanes 32 .. 63 idle while the records are formed) halve the
// records' LDS, which otherwise limits its resident waves (measured: 720-view 512^2 x 256
// SF back 194 -> 173 ms).  Short detectors keep 64 (there the records' second half of the
// lanes pays: 1.13 vs 1.17 ms per PWLS subset).
constexpr int kBackViews = 32;
constexpr int kRowWin = 1019;  // detector rows staged per wave (nv larger: windowed); with the
                                // zero pads one wave's row buffer is 1024 words


template <class T>
// 8-byte aligned: 72 bytes in FP32 rather than 80, so that a short-detector block's 64-view
// records and row buffers fit eight blocks per CU (eight waves per SIMD)
struct __align__(8) ViewCol {  // one view's footprint on one image column
    int cells;                  // first footprint cell ilo | cell count nw << 16
    int rows;                   // rows covered r0 | r1 << 16
    int cls;                    // gather class: rows per voxel 2, 3, 4 (0, 1, 2) or the capped loop (3)
    unsigned voff;              // bytes from the batch's first view to row r0 of column ilo of this view
    T ampr, r2;                 // amp / r and r^2: the voxel factor is sqrt(r^2 + zc^2) amp / r
    T zinc, vs15, wzdz;         // v of voxel z spans [vs15 + z zinc, vs15 + (z + 1) zinc] - 1 (the +1/2 edge
                                // shift and a +1 that keeps the clamped row index non-negative folded in);
                                // zc = z dz + wzdz
    T vstart;                   // the unshifted start, for footprints that may reach the cap
    T w0, w1, w2, w3;           // weights of the first four footprint cells, times amp / r
    T p0, p1, p2, p3;           // footprint corners (cells past the fourth)
};

// Staged rows of a window [rw, rwe] sit at sw[1 + row - rw]; sw[0] and the kRowPad
// words after the last row are zero.  gather_rows sums, over the K rows from
// ra = floor(a) on, the overlap of the voxel's edge interval [a, b] with the row times
// the staged value.  ra is clamped to [rw - 1, rwe + 1]: rows outside the window then
// read the zero words (they belong to another window or lie off the detector), and the
// overlaps follow the DD rule: row ra takes min(b, ra + 1) - a, row f > ra takes
// max(min(b, f + 1) - f, 0), which equals clamp01(b - f) bit for bit (f is an integer:
// b - f is exact wherever the result lies inside (0, 1)).  A voxel meets at most
// K = floor(zinc) + 2 rows.
constexpr int kRowPad = 4;

// rows staged per wave for a detector of nv rows, the words of one wave's row buffer (at
// least one row per lane plus the zero words on either side) and the dynamic LDS of a block
inline int back_rowwin(int nv) { return nv < kRowWin ? nv : kRowWin; }
__host__ __device__ inline int back_rowbuf(int rowwin) { return (rowwin > kWave ? rowwin : kWave) + 1 + kRowPad; }
template <class T>
inline size_t back_lds_bytes(int rowwin) {
    return (size_t)kWavesPerBlock * (size_t)back_rowbuf(rowwin) * sizeof(T);
}

// The gathers take a voxel's edges a, b in row units plus one (so that the clamped row
// index is never negative), swr with swr[row + 1] = sw[1 + row - rw], and flo, fhi =
// rw, rwe + 2 as T: a shift of a, b, flo, fhi by one and of swr by minus one changes no
// weight and no word read.
//
// Voxels of at most one row (zinc <= 1) meet rows ra = floor(a) and ra + 1:
// s = (zinc - e) p[ra] + e p[ra + 1] with e = max(b - (ra + 1), 0) = max(frac(a) + zinc - 1, 0),
// the same two DD overlaps as below.  a is clamped to [flo, fhi] first: a clamped value is
// an integer, so e = 0 and the voxel reads only a zero word.
template <class T>
__device__ __forceinline__ T gather_rows2(const T* swr, T flo, T fhi, T a, T zinc, T zm1) {
    const T ac = fmin(fmax(a, flo), fhi);
    const T fa = floor(ac);
    const T* pr = swr + (int)fa;
    const T e = clamp01((ac - fa) + zm1);
    const T p0 = pr[0], p1 = pr[1];
    return fma_t(e, p1 - p0, zinc * p0);
}
// FP32: the clamp as one med3, the row as one truncating convert (flo >= 0, so ac >= 0) and
// frac(ac) as v_fract (exact: ac - floor(ac))
template <>
__device__ __forceinline__ float gather_rows2<float>(const float* swr, float flo, float fhi, float a, float zinc,
                                                      float zm1) {
    const float ac = __builtin_amdgcn_fmed3f(a, flo, fhi);
    const float* pr = swr + (int)ac;
    const float e = clamp01(__builtin_amdgcn_fractf(ac) + zm1);
    const float p0 = pr[0], p1 = pr[1];
    return fmaf(e, p1 - p0, zinc * p0);
}

// Thicker voxels: the clamp of floor(a) to [flo, fhi] is taken in T.
template <class T, int K>
__device__ __forceinline__ T gather_rows(const T* swr, T flo, T fhi, T a, T b) {
    static_assert(K - 1 <= kRowPad, "pad covers the rows past the window");
    const T fa = min(max(floor(a), flo), fhi);
    const T* pr = swr + (int)fa;
    T f = fa + T(1);
    T s = (fmin(b, f) - a) * pr[0];
#pragma unroll
    for (int k = 1; k < K; ++k) {
        s += clamp01(b - f) * pr[k];
        f += T(1);
    }
    return s;
}

// SHORT: detectors of at most 64 rows, where every view takes the one-pass row staging
// (the windowed form is compiled out, which keeps the loop's registers in place)
// FUSE: the column's finished gradient goes straight into the Nesterov-SQS update of its
// voxels (the update kernel's arithmetic with the Huber penalty, na's buffers) instead of
// to vol; columns outside the map are updated with a zero gradient.
template <class Tin, class Tout, class T, int FOOT, int NZL, bool SHORT, bool FUSE>
__global__ __launch_bounds__(kBlock) void cbct_back_kernel(Geom<T> g, const View<T>* __restrict__ views, int vstride,
                                                           int nviews, const Tin* __restrict__ proj,
                                                           const uint8_t* __restrict__ map,
                                                           Tout* __restrict__ vol, T scale, int accumulate,
                                                           int rowwin, NestArgs na, int order) {
    static_assert(!FUSE || (std::is_same<T, float>::value && SHORT), "the fused update is FP32, short detectors");
    // row buffers (rowwin + 1 + kRowPad words per wave) in dynamic LDS, sized at launch
    // for the detector, so short detectors leave room for more resident waves
    extern __shared__ __align__(16) unsigned char back_dyn_lds[];
    T* srow = reinterpret_cast<T*>(back_dyn_lds);
    constexpr int VB = SHORT ? kWave : kBackViews;  // records per batch
    __shared__ ViewCol<T> vbuf[kWavesPerBlock][VB];
    // blocks in dispatch order (consecutive blocks on different XCDs: each XCD's share of the
    // columns in flight stays within the current rows or super-tile; a contiguous range of
    // blocks per XCD measured 7% slower on the 720-view 512^2 x 256 SF back)
    const int bid = blockIdx.x;
    const int wv = (int)(threadIdx.x / kWave);
    // column order: 0 = four consecutive columns per block; 1 = 2 x 2 column patches in
    // super-tiles of kSuperTile x kSuperTile columns (the FDK backprojector's order)
    int ix, iy;
    if (order == 0) {
        const int colid = bid * kWavesPerBlock + wv;
        if (colid >= g.nx * g.ny) return;
        ix = colid % g.nx;
        iy = colid / g.nx;
    } else {
        constexpr int kHalfT = kSuperTile / 2;
        const int stl = bid / (kHalfT * kHalfT), r = bid % (kHalfT * kHalfT);
        const int stx = (g.nx + kSuperTile - 1) / kSuperTile;
        ix = (stl % stx) * kSuperTile + (r % kHalfT) * 2 + (wv & 1);
        iy = (stl / stx) * kSuperTile + (r / kHalfT) * 2 + (wv >> 1);
        if (ix >= g.nx || iy >= g.ny) return;
    }
    const int colid = iy * g.nx + ix;
    const int lane = lane_id();
    Tout* ocol = vol + (size_t)colid * g.zstride;
    const bool masked = map != nullptr && map[iy * g.nx + ix] == 0;
    if (masked && !FUSE) {
        if (!accumulate)
            for (int iz = lane; iz < g.nz; iz += kWave) ocol[iz] = cvt_store<Tout>(T(0));
        return;
    }
    T* sw = srow + (size_t)wv * back_rowbuf(rowwin);
    // the word before the first staged row and the kRowPad words after the 64 lane slots
    // start at zero; the one-pass form below never writes them
    if (lane <= kRowPad) sw[lane == 0 ? 0 : kWave + lane] = T(0);
    ViewCol<T>* vw = vbuf[wv];
    const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
    T img[NZL];
#pragma unroll
    for (int c = 0; c < NZL; ++c) img[c] = T(0);
    for (int b0 = 0; b0 < (masked ? 0 : nviews); b0 += VB) {
        // footprint of view b0 + lane on this column (lanes below VB)
        bool ok = lane < VB && b0 + lane < nviews;
        ViewCol<T> vc{};
        vc.zinc = T(1);
        int iz0 = 0, iz1 = 0, r0 = 0, r1 = 0;
        T vstart = T(0);
        if (ok) {
            const View<T> v = views[(size_t)(b0 + lane) * vstride];
            const Column<T> cg = column_geom(xc, yc, v, g.sad, g.dx);
            T clo, chi;
            int ilo = 0, ihi = -1;
            if (FOOT == FOOT_DD) {
                T umin, umax;
                dd_bounds(cg, v, g, umin, umax);
                ok = clamp_cells(umin, umax, g.nu, 16, clo, chi, ilo, ihi);
                vc.p0 = clo;
                vc.p1 = chi;
            } else {
                T tau[4];
                sf_taus(cg, v, g, tau);
                ok = clamp_cells(tau[0], tau[3], g.nu, 16, clo, chi, ilo, ihi);
                vc.p0 = tau[0]; vc.p1 = tau[1]; vc.p2 = tau[2]; vc.p3 = tau[3];
            }
            const T wz = v.wz + g.zoff;
            if (ok)
                ok = v_window(g, cg.d, wz, vc.zinc, vstart, iz0, iz1) && iz0 < iz1 && vc.zinc > T(0);
            if (ok) {
                // rows touched by voxels iz0 .. iz1-1
                T a0 = vstart + T(iz0) * vc.zinc + T(0.5);
                T a1 = vstart + T(iz1 - 1) * vc.zinc + vc.zinc + T(0.5);
                a0 = a0 > T(0) ? a0 : T(0);
                a1 = a1 < T(g.nv) ? a1 : T(g.nv);
                ok = a0 <= a1;
                r0 = min((int)floor(a0), g.nv - 1);
                r1 = min((int)floor(a1), g.nv - 1);
            }
            if (ok) {
                vc.cells = ilo | ((ihi - ilo + 1) << 16);
                vc.rows = r0 | (r1 << 16);
                vc.cls = vc.zinc <= T(1) ? 0 : (vc.zinc <= T(2) ? 1 : (vc.zinc <= T(3) ? 2 : 3));
                vc.ampr = cg.amp * frcp(cg.r);
                vc.w0 = u_weight<T, FOOT>(ilo + 0, ilo, ihi, vc.p0, vc.p1, vc.p2, vc.p3) * vc.ampr;
                vc.w1 = u_weight<T, FOOT>(ilo + 1, ilo, ihi, vc.p0, vc.p1, vc.p2, vc.p3) * vc.ampr;
                vc.w2 = u_weight<T, FOOT>(ilo + 2, ilo, ihi, vc.p0, vc.p1, vc.p2, vc.p3) * vc.ampr;
                vc.w3 = u_weight<T, FOOT>(ilo + 3, ilo, ihi, vc.p0, vc.p1, vc.p2, vc.p3) * vc.ampr;
                vc.r2 = cg.r * cg.r;
                vc.vs15 = vstart + T(1.5);
                vc.vstart = vstart;
                vc.wzdz = wz * g.dz;
                vc.voff = ((unsigned)(lane * g.nu + ilo) * (unsigned)g.nv + (unsigned)r0) * (unsigned)sizeof(Tin);
            }
        }
        if (lane < VB) vw[lane] = vc;
        __builtin_amdgcn_wave_barrier();
        unsigned long long mask = __ballot(ok);
        // the first row chunk of the next view is loaded while the current one is used
        int cur = mask ? __ffsll((long long)mask) - 1 : -1;
        if (cur >= 0) mask &= mask - 1;
        Tin pre0{}, pre1{}, pre2{}, pre3{};
        // the four footprint columns of view vb, rows r0 + lane (at most r1): loads through a
        // resource over the view's sinogram, so that columns past the detector's last read
        // zero; all four are loaded (loads under a condition would be waited for at once),
        // the columns past the footprint are not used
        // the batch's first view; a record's voff reaches any row of its view from here
        const char* bbase = reinterpret_cast<const char*>(proj + (size_t)b0 * g.nu * g.nv);
        auto load_rows = [&](unsigned voff, int cells, int rows) {
            // a resource from row r0 of the footprint's first column to the end of its last,
            // so that the loads of columns past the footprint (and past the detector) read
            // zero.  (One resource per column, with the column steps in the scalar bases and
            // one shared lane offset, saves three VALU adds but was measured 15% slower.)
            const int r0 = rows & 0xffff, r1 = rows >> 16, nw = cells >> 16;
            const auto rs = buf_rsrc(bbase + voff, (nw * g.nv - r0) * (int)sizeof(Tin));
            const int off = min(lane, r1 - r0) * (int)sizeof(Tin);
            const int step = g.nv * (int)sizeof(Tin);
            pre0 = buf_ld<Tin>(rs, off);
            pre1 = buf_ld<Tin>(rs, off + step);
            pre2 = buf_ld<Tin>(rs, off + 2 * step);
            pre3 = buf_ld<Tin>(rs, off + 3 * step);
        };
        // SHORT: the current view's integer words stay in scalar registers from the loading
        // of its rows on (the windowed form reads them again: its scalar registers are tighter)
        int cur_cells = 0, cur_rows = 0, cur_cls = 0;
        if (cur >= 0) {
            const ViewCol<T>& c = vw[cur];
            cur_cells = uni(c.cells);
            cur_rows = uni(c.rows);
            if constexpr (SHORT) cur_cls = uni(c.cls);
            load_rows((unsigned)uni((int)c.voff), cur_cells, cur_rows);
        }
        while (cur >= 0) {
            const ViewCol<T>& c = vw[cur];
            const int cells = SHORT ? cur_cells : uni(c.cells), rows = SHORT ? cur_rows : uni(c.rows);
            const int cls = SHORT ? cur_cls : uni(c.cls), vb = b0 + cur;
            const int ilo = cells & 0xffff, nw = cells >> 16, r0 = rows & 0xffff, r1 = rows >> 16;
            // SHORT: the record's reals stay in vector registers (every lane reads the same LDS
            // words): they are only operands of vector arithmetic, and moving them to scalar
            // registers would cost one instruction each per view.  The windowed form, with
            // more voxels per lane, moves them to scalar registers to keep its occupancy.
            auto rd = [&](const T& x) -> T { return SHORT ? x : uni(x); };
            const T w0 = rd(c.w0), w1 = rd(c.w1), w2 = rd(c.w2), w3 = rd(c.w3);  // amp / r folded in
            const T zinc = rd(c.zinc), vs15 = rd(c.vs15);
            const T r2 = rd(c.r2);      // ampv * r = sqrt(r^2 + zc^2)
            const T wzdz = rd(c.wzdz);  // zc = z dz + wz dz
            const Tin* pv = proj + ((size_t)vb * g.nu + ilo) * g.nv;
            // the next view's first row chunk starts loading once this one is staged
            auto next = [&]() {
                const int nxt = mask ? __ffsll((long long)mask) - 1 : -1;
                if (nxt >= 0) {
                    mask &= mask - 1;
                    const ViewCol<T>& cn = vw[nxt];
                    const int ncells = uni(cn.cells), nrows = uni(cn.rows);
                    if constexpr (SHORT) {
                        cur_cells = ncells;
                        cur_rows = nrows;
                        cur_cls = uni(cn.cls);
                    }
                    load_rows((unsigned)uni((int)cn.voff), ncells, nrows);
                }
                cur = nxt;
            };
            // the column's voxels gather the staged rows [rw, rwe].  Voxels outside the
            // column's range [iz0, iz1) need no test: their edges lie more than half a row
            // off the detector, so the clamped row index reads the zero words before or
            // after the staged rows (or overlaps of zero), as do voxels past nz, whose sums
            // are never stored.
            auto gather = [&](int rw, int rwe) {
                // edges and rows shifted by one: swr[row + 1] = sw[1 + row - rw], and the clamp
                // range [rw, rwe + 2] of floor(a + 1) is non-negative
                const T* swr = sw - rw;
                const T flo = T(rw), fhi = T(rwe + 2), zm1 = zinc - T(1);
                // the class is the same for all of the column's voxels: one branch per view,
                // each class with its own unrolled voxel loop
                auto voxels = [&](auto cls_k) {
                    constexpr int CLS = decltype(cls_k)::value;
#pragma unroll
                    for (int cz = 0; cz < NZL; ++cz) {
                        const int iz = cz * kWave + lane;
                        // voxel edges in row units plus one, +3/2 folded into vs15 (within an
                        // ulp of the reference's vstart + iz zinc + 1/2)
                        const T a = fma_t(T(iz), zinc, vs15);
                        T s;
                        if constexpr (CLS == 0) {
                            s = gather_rows2<T>(swr, flo, fhi, a, zinc, zm1);
                        } else if constexpr (CLS == 1) {
                            s = gather_rows<T, 3>(swr, flo, fhi, a, a + zinc);
                        } else if constexpr (CLS == 2) {
                            s = gather_rows<T, 4>(swr, flo, fhi, a, a + zinc);
                        } else {
                            // footprints that may reach the 16-cell cap: the reference's edges
                            const T vmin = uni(c.vstart) + T(iz) * zinc;
                            T clo, chi;
                            int vlo, vhi;
                            s = T(0);
                            if (clamp_cells(vmin, vmin + zinc, g.nv, 16, clo, chi, vlo, vhi)) {
                                const int ra = max(vlo, rw), rb = min(vhi, rwe);
                                for (int iv = ra; iv <= rb; ++iv)
                                    s += dd_overlap16(iv, vmin, vmin + zinc, g.nv) * sw[iv - rw + 1];
                            }
                        }
                        const T zcm = T(iz) * g.dz + wzdz;
                        img[cz] += fsqrt(fma_t(zcm, zcm, r2)) * s;
                    }
                };
                if (cls == 0) voxels(std::integral_constant<int, 0>{});
                else if (cls == 1) voxels(std::integral_constant<int, 1>{});
                else if (cls == 2) voxels(std::integral_constant<int, 2>{});
                else voxels(std::integral_constant<int, 3>{});
            };
            // the windowed form (tall detectors) instantiates the gather twice; there one copy of
            // the voxel loop with the class chosen per voxel keeps the code small enough for the
            // instruction cache
            auto gather_small = [&](int rw, int rwe) {
                const T* swr = sw - rw;
                const T flo = T(rw), fhi = T(rwe + 2), zm1 = zinc - T(1);
#pragma unroll
                for (int cz = 0; cz < NZL; ++cz) {
                    const int iz = cz * kWave + lane;
                    const T a = fma_t(T(iz), zinc, vs15);
                    T s;
                    if (cls == 0) {
                        s = gather_rows2<T>(swr, flo, fhi, a, zinc, zm1);
                    } else if (cls == 1) {
                        s = gather_rows<T, 3>(swr, flo, fhi, a, a + zinc);
                    } else if (cls == 2) {
                        s = gather_rows<T, 4>(swr, flo, fhi, a, a + zinc);
                    } else {
                        const T vmin = uni(c.vstart) + T(iz) * zinc;
                        T clo, chi;
                        int vlo, vhi;
                        s = T(0);
                        if (clamp_cells(vmin, vmin + zinc, g.nv, 16, clo, chi, vlo, vhi)) {
                            const int ra = max(vlo, rw), rb = min(vhi, rwe);
                            for (int iv = ra; iv <= rb; ++iv)
                                s += dd_overlap16(iv, vmin, vmin + zinc, g.nv) * sw[iv - rw + 1];
                        }
                    }
                    const T zcm = T(iz) * g.dz + wzdz;
                    img[cz] += fsqrt(fma_t(zcm, zcm, r2)) * s;
                }
            };
            // (short columns only: with more voxels per lane the second copy of the gather
            // costs the longer columns a wave per SIMD)
            if (SHORT || (NZL <= 2 && r1 - r0 + 1 + kRowPad <= kWave)) {
                // one row per lane: lanes past r1 write the zero words the gather reads past
                // the rows (past the 64 slots they were zeroed at the start)
                // columns past the footprint were loaded as zeros (and weigh zero)
                T s = w0 * T(to_f(pre0)) + w1 * T(to_f(pre1)) + w2 * T(to_f(pre2)) + w3 * T(to_f(pre3));
                for (int k = 4; k < nw; ++k)
                    s += u_weight<T, FOOT>(ilo + k, ilo, ilo + nw - 1, c.p0, c.p1, c.p2, c.p3) * c.ampr *
                         ld<T>(pv + (size_t)k * g.nv + min(r0 + lane, r1));
                sw[1 + lane] = r0 + lane <= r1 ? s : T(0);
                __builtin_amdgcn_wave_barrier();
                next();
                if constexpr (SHORT) gather(r0, r1);
                else gather_small(r0, r1);
                __builtin_amdgcn_wave_barrier();
                continue;
            }
            for (int rw = r0; rw <= r1; rw += rowwin) {
                const int rwe = min(r1, rw + rowwin - 1);
                // u combination of the footprint's detector columns, row by row
                int rfirst = rw + lane;
                if (rw == r0) {
                    if (rfirst <= rwe) {
                        T s = w0 * T(to_f(pre0)) + w1 * T(to_f(pre1)) + w2 * T(to_f(pre2)) + w3 * T(to_f(pre3));
                        for (int k = 4; k < nw; ++k)
                            s += u_weight<T, FOOT>(ilo + k, ilo, ilo + nw - 1, c.p0, c.p1, c.p2, c.p3) * c.ampr *
                                 ld<T>(pv + (size_t)k * g.nv + rfirst);
                        sw[rfirst - rw + 1] = s;
                    }
                    rfirst += kWave;
                }
                if (lane <= kRowPad) sw[lane == 0 ? 0 : rwe - rw + 1 + lane] = T(0);  // zero pads
                // the other row chunks, each chunk's loads issued while the one before is
                // combined: the four footprint columns through one resource from row r0 of
                // column ilo to the end of the footprint's last column (columns past it read
                // zero and weigh zero, as in the first chunk; lanes past rwe read row r1)
                const auto rs = buf_rsrc(bbase + uni((int)c.voff), (nw * g.nv - r0) * (int)sizeof(Tin));
                const int step = g.nv * (int)sizeof(Tin);
                Tin q0{}, q1{}, q2{}, q3{};
                auto ld4 = [&](int row, Tin& a0, Tin& a1, Tin& a2, Tin& a3) {
                    const int off = (min(row, r1) - r0) * (int)sizeof(Tin);
                    a0 = buf_ld<Tin>(rs, off);
                    a1 = buf_ld<Tin>(rs, off + step);
                    a2 = buf_ld<Tin>(rs, off + 2 * step);
                    a3 = buf_ld<Tin>(rs, off + 3 * step);
                };
                if (rfirst - lane <= rwe) ld4(rfirst, q0, q1, q2, q3);
                for (int row = rfirst; row - lane <= rwe; row += kWave) {
                    Tin n0{}, n1{}, n2{}, n3{};
                    if (row - lane + kWave <= rwe) ld4(row + kWave, n0, n1, n2, n3);
                    if (row <= rwe) {
                        T s = w0 * T(to_f(q0)) + w1 * T(to_f(q1)) + w2 * T(to_f(q2)) + w3 * T(to_f(q3));
                        for (int k = 4; k < nw; ++k)
                            s += u_weight<T, FOOT>(ilo + k, ilo, ilo + nw - 1, c.p0, c.p1, c.p2, c.p3) * c.ampr *
                                 ld<T>(pv + (size_t)k * g.nv + row);
                        sw[row - rw + 1] = s;
                    }
                    q0 = n0;
                    q1 = n1;
                    q2 = n2;
                    q3 = n3;
                }
                __builtin_amdgcn_wave_barrier();
                if (rw == r0) next();
                gather_small(rw, rwe);
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if constexpr (FUSE) {
        // reconPwlsSeNesterovSqs.m:67-87 for the column's voxels, as the update kernel does it
        double t = na.t, t_next = na.t_next;
        if (na.tseq) {
            const int i0 = *na.base + na.k;
            t = na.tseq[i0];
            t_next = na.tseq[i0 + 1];
        }
        const double a = 1.0 - 1.0 / t_next, b = 1.0 / t_next;
#pragma unroll
        for (int c = 0; c < NZL; ++c) {
            const int iz = c * kWave + lane;
            if (iz < g.nz) {
                const size_t i = (size_t)colid * g.nz + iz;
                const float sz = pen_point<1, POT_HUBER>(na.z, g.nx, g.ny, g.nz, ix, iy, iz, na.delta);
                const float tx = pen_point<2, POT_HUBER>(na.x, g.nx, g.ny, g.nz, ix, iy, iz, na.delta);
                // the gradient rounded as a stored one is (no fused multiply-add across it)
                const float grad = __fmul_rn(scale, img[c]) + (float)(na.beta * (double)sz);
                const double den = (double)na.precom[i] + na.beta * (double)tx;
                const float du = (float)((double)grad / den);
                const float xn = na.z[i] - du;
                const float sn = na.s[i] - (float)(t * (double)du);
                const float zn = (float)(a * (double)xn + b * (double)sn);
                na.x_out[i] = scrub(xn);
                na.s[i] = scrub(sn);
                const float zs = scrub(zn);
                na.z_out[i] = zs;
                if (na.zcopy) na.zcopy[i] = na.zref ? zs - na.zref[i] : zs;
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < NZL; ++c) {
            const int iz = c * kWave + lane;
            if (iz < g.nz) {
                const T val = scale * img[c];
                ocol[iz] = cvt_store<Tout>(accumulate ? ld<T>(ocol + iz) + val : val);
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// FDK voxel-driven bilinear backprojection (cbct_back1_pd).  off_boundary = 1 is
// 'back,pd' (skip below the detector, stop above it), 0 is 'back,hpd' (clamp v).
// One wavefront per image column, NZL voxels per lane.
//
// Every wave walks all views, so what stays in L2 is the detector footprint of the
// columns resident over the few views they are apart.  Blocks therefore take 2 x 2 column
// patches in super-tiles of 32 x 32 columns (256 blocks) in dispatch order: the columns in
// flight form a compact patch whose per-view footprint is about a tenth of the detector,
// not two full image rows.  (Giving each XCD a contiguous range of super-tiles instead was
// measured 10% slower at 512^3 from 720 views.)
// ---------------------------------------------------------------------------------
constexpr int kPdTile = kSuperTile;  // super-tile edge in columns

// The reference tests u and v in double against 1e-6 and n - 1 - 1e-6.  In T these are
// the same tests against lo = the smallest T whose double value is >= 1e-6 and hi = the
// largest T whose double value is <= n - 1 - 1e-6 (computed on the host).
template <class T>
struct PdLimits {
    T lo, u_hi, v_hi;
};
static_assert(kWavesPerBlock == 4, "a block's waves form a 2 x 2 column patch");

__host__ __device__ inline int pd_blocks(int nx, int ny) {
    return ((nx + kPdTile - 1) / kPdTile) * ((ny + kPdTile - 1) / kPdTile) * (kPdTile * kPdTile / kWavesPerBlock);
}
// One view's u, w2 and w2s at the column (xc, yc) in the reference's float operations,
// one rounding each and no contraction (cbct_back.c:195-212): u, its cell and weight, and
// w2s are then the reference's own.
__device__ __forceinline__ void pd_view_col_f32(const Geom<float>& g, const View<float>& v, float xc, float yc,
                                                float& u, float& w2, float& w2s) {
#pragma clang fp contract(off)
    const float xr = xc * v.c - yc * v.s;
    const float yr = xc * v.s + yc * v.c;
    if (g.dettype == 1) {
        w2 = g.sdd / (g.sad + yr);
        w2s = w2 * w2;
        u = xr * w2 / g.du - g.wu;
    } else {
        const float roff = g.ou * g.sad / g.sdd;
        u = atanf((xr - roff) / (g.sad + yr)) * g.sdd / g.du - g.wu;
        w2s = g.sdd * g.sdd / ((g.sad + yr) * (g.sad + yr) + (xr - roff) * (xr - roff));
        w2 = g.sdd / (g.sad + yr);
    }
}

template <class Tin, class Tout, class T, int NZL, int OFF_BOUNDARY>
__global__ __launch_bounds__(kBlock) void cbct_back_pd_kernel(Geom<T> g, const View<T>* __restrict__ views,
                                                              int vstride, int nviews, const Tin* __restrict__ proj,
                                                              const uint8_t* __restrict__ map,
                                                              Tout* __restrict__ vol, T scale, int accumulate,
                                                              PdLimits<T> lim) {
    constexpr bool off_boundary = OFF_BOUNDARY != 0;
    const int bid = blockIdx.x;
    const int wv = (int)(threadIdx.x / kWave);
    constexpr int kHalf = kPdTile / 2;  // 2 x 2 patches per super-tile edge
    const int st = bid / (kHalf * kHalf), r = bid % (kHalf * kHalf);
    const int stx = (g.nx + kPdTile - 1) / kPdTile;
    const int ix = (st % stx) * kPdTile + (r % kHalf) * 2 + (wv & 1);
    const int iy = (st / stx) * kPdTile + (r / kHalf) * 2 + (wv >> 1);
    if (ix >= g.nx || iy >= g.ny) return;
    const int colid = iy * g.nx + ix;
    const int lane = lane_id();
    Tout* ocol = vol + (size_t)colid * g.zstride;
    if (map != nullptr && map[iy * g.nx + ix] == 0) {
        if (!accumulate)
            for (int iz = lane; iz < g.nz; iz += kWave) ocol[iz] = cvt_store<Tout>(T(0));
        return;
    }
    const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
    T izf[NZL];  // voxel indices as T, once
#pragma unroll
    for (int c = 0; c < NZL; ++c) izf[c] = T(c * kWave + lane);
    T img[NZL];
#pragma unroll
    for (int c = 0; c < NZL; ++c) img[c] = T(0);
    for (int b = 0; b < nviews; ++b) {
        const View<T> v = views[(size_t)b * vstride];
        T u, w2, w2s;
        if constexpr (std::is_same<T, float>::value) {
            pd_view_col_f32(g, v, xc, yc, u, w2, w2s);
        } else {
            const T xr = xc * v.c - yc * v.s;
            const T yr = xc * v.s + yc * v.c;
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
        }
        if (!(u >= lim.lo && u <= lim.u_hi)) continue;  // also skips NaN
        const int iu = (int)floor(u);
        const T wul = u - T(iu);
        // v of voxel iz: (iz + wz) dz w2 / dv - wv, evaluated as iz A + B with the view's
        // A and B (one FMA per voxel; within an ulp or two of the reference's order)
        const T A = g.dz * w2 / g.dv;
        const T B = (v.wz + g.zoff) * A - g.wv;
        // 'back,pd' stops at the first voxel above the detector; v is monotone in iz, so
        // that is a per-voxel test when v grows and "nothing if voxel 0 of the volume is
        // already above" when it shrinks
        if (off_boundary && !(A > T(0)) && v.wz * A - g.wv > lim.v_hi) continue;
        const Tin* p = proj + ((size_t)b * g.nu + iu) * g.nv;
#pragma unroll
        for (int c = 0; c < NZL; ++c) {
            const int iz = c * kWave + lane;
            if (iz >= g.nz) continue;
            T vv = fma_t(izf[c], A, B);
            if (off_boundary) {
                if (!(vv >= lim.lo && vv <= lim.v_hi)) continue;
            } else {
                if (vv < lim.lo) vv = T(0.5);
                if (vv > lim.v_hi) vv = T(g.nv) - T(1.5);
                if (vv != vv) continue;  // NaN
            }
            // the tests above leave v in [1e-6, nv - 1 - 1e-6] or at 0.5 / nv - 1.5, so the
            // clamp is a no-op that keeps the reads in range whatever the arithmetic
            const int iv = max(0, min((int)floor(vv), g.nv - 2));
            const T wvl = vv - T(iv);
            const T a0 = ld<T>(p + iv), a1 = ld<T>(p + iv + 1), a2 = ld<T>(p + g.nv + iv),
                    a3 = ld<T>(p + g.nv + iv + 1);
            img[c] += w2s * ((T(1) - wul) * ((T(1) - wvl) * a0 + wvl * a1) + wul * ((T(1) - wvl) * a2 + wvl * a3));
        }
    }
#pragma unroll
    for (int c = 0; c < NZL; ++c) {
        const int iz = c * kWave + lane;
        if (iz < g.nz) {
            const T val = scale * img[c];
            ocol[iz] = cvt_store<Tout>(accumulate ? ld<T>(ocol + iz) + val : val);
        }
    }
}

// ---------------------------------------------------------------------------------
// FDK backprojection, staged form (FP32 arithmetic, detectors of at most kPdRows rows).
// The same columns, views, tests and per-voxel v as cbct_back_pd_kernel; what differs:
//
//  * the per-view quantities of the column (u, its cell and weight, w2s, the v line
//    A iz + B and the detector rows the column's voxels reach) are evaluated one view per
//    lane, 64 views at a time, and read with readlane in the view loop;
//  * per view the wave interpolates the two detector columns iu, iu + 1 in u once, row by
//    row, c(iv) = (1 - wul) p(iu, iv) + wul p(iu + 1, iv), over the rows its voxels reach
//    (coalesced 8-byte loads), into its LDS row buffer; each voxel then interpolates c in v:
//    w2s ((1 - wvl) c(iv) + wvl c(iv + 1)).  This is the reference's bilinear sum with the
//    u and v interpolations taken in the other order (the same value up to FP32 rounding).
//
// The rows a column reaches: iv(v) = clamp(floor(m(v)), 0, nv - 2), m being the
// boundary rule of the mode (hpd: 0.5 below, nv - 1.5 above; pd: voxels outside are
// skipped, so the clamp to [lo, v_hi] bounds the ones kept), is non-decreasing in v, and
// v is monotone in iz, so the rows lie between those of the first and last voxel.
// ---------------------------------------------------------------------------------
constexpr int kPdRows = 1024;             // detector rows a wave stages (taller: the direct kernel)
constexpr int kPdPad = 4;                 // words before the first staged row
constexpr int kPdStageW = kPdRows + 16;   // words per wave: pad, rows, the last load's overshoot

// the E = 8 / sizeof(Tin) elements from p on as one 8-byte word; when only avail < E of
// them lie inside the column, those, and zero bits for the rest
template <class Tin>
__device__ __forceinline__ uint2 pd_load8(const Tin* p, int avail) {
    uint2 x = make_uint2(0u, 0u);
    if (avail >= 8 / (int)sizeof(Tin)) {
        __builtin_memcpy(&x, p, 8);
    } else if constexpr (sizeof(Tin) == 4) {
        if (avail > 0) x.x = *(const unsigned*)p;
    } else {
        const unsigned short* q = (const unsigned short*)p;
        if (avail > 0) x.x = q[0];
        if (avail > 1) x.x |= (unsigned)q[1] << 16;
        if (avail > 2) x.y = q[2];
    }
    return x;
}

template <class Tin>
__device__ __forceinline__ void pd_unpack8(uint2 x, float* out);
template <>
__device__ __forceinline__ void pd_unpack8<float>(uint2 x, float* out) {
    out[0] = __uint_as_float(x.x);
    out[1] = __uint_as_float(x.y);
}
template <>
__device__ __forceinline__ void pd_unpack8<__half>(uint2 x, float* out) {
    out[0] = __half2float(__ushort_as_half((unsigned short)(x.x & 0xffffu)));
    out[1] = __half2float(__ushort_as_half((unsigned short)(x.x >> 16)));
    out[2] = __half2float(__ushort_as_half((unsigned short)(x.y & 0xffffu)));
    out[3] = __half2float(__ushort_as_half((unsigned short)(x.y >> 16)));
}
template <>
__device__ __forceinline__ void pd_unpack8<__hip_bfloat16>(uint2 x, float* out) {
    out[0] = __uint_as_float(x.x << 16);
    out[1] = __uint_as_float(x.x & 0xffff0000u);
    out[2] = __uint_as_float(x.y << 16);
    out[3] = __uint_as_float(x.y & 0xffff0000u);
}

// iv of a voxel at v (the row pair it reads), for the row range of a column
template <int OFF_BOUNDARY>
__device__ __forceinline__ int pd_row_of(float v, const PdLimits<float>& lim, int nv) {
    float m;
    if (OFF_BOUNDARY) {
        m = fminf(fmaxf(v, lim.lo), lim.v_hi);
    } else {
        m = v;
        if (m < lim.lo) m = 0.5f;
        if (m > lim.v_hi) m = float(nv) - 1.5f;
    }
    return max(0, min((int)floorf(m), nv - 2));
}

template <class Tin, class Tout, int NZL, int OFF_BOUNDARY>
__global__ __launch_bounds__(kBlock) void cbct_back_pd_stage_kernel(Geom<float> g, const View<float>* __restrict__ views,
                                                                    int vstride, int nviews,
                                                                    const Tin* __restrict__ proj,
                                                                    const uint8_t* __restrict__ map,
                                                                    Tout* __restrict__ vol, float scale,
                                                                    int accumulate, PdLimits<float> lim) {
    using T = float;
    constexpr bool off_boundary = OFF_BOUNDARY != 0;
    constexpr int E = 8 / (int)sizeof(Tin);  // rows per 8-byte load
    __shared__ __align__(16) float stage[kWavesPerBlock * kPdStageW];
    const int bid = blockIdx.x;
    const int wv = (int)(threadIdx.x / kWave);
    constexpr int kHalf = kPdTile / 2;
    const int st = bid / (kHalf * kHalf), r = bid % (kHalf * kHalf);
    const int stx = (g.nx + kPdTile - 1) / kPdTile;
    const int ix = (st % stx) * kPdTile + (r % kHalf) * 2 + (wv & 1);
    const int iy = (st / stx) * kPdTile + (r / kHalf) * 2 + (wv >> 1);
    if (ix >= g.nx || iy >= g.ny) return;
    const int colid = iy * g.nx + ix;
    const int lane = lane_id();
    Tout* ocol = vol + (size_t)colid * g.zstride;
    if (map != nullptr && map[iy * g.nx + ix] == 0) {
        if (!accumulate)
            for (int iz = lane; iz < g.nz; iz += kWave) ocol[iz] = cvt_store<Tout>(T(0));
        return;
    }
    float* cs = stage + wv * kPdStageW;
    const T xc = (T(ix) + g.wx) * g.dx, yc = (T(iy) + g.wy) * g.dy;
    T izf[NZL];
#pragma unroll
    for (int c = 0; c < NZL; ++c) izf[c] = T(c * kWave + lane);
    T img[NZL];
#pragma unroll
    for (int c = 0; c < NZL; ++c) img[c] = T(0);
    for (int b0 = 0; b0 < nviews; b0 += kWave) {
        // view b0 + lane: cell iu (-1: the column is skipped in this view), weight wul, w2s,
        // the v line (A, B) and the rows r0 .. r1 its voxels read
        int h_iu = -1, h_rows = 0;
        T h_wul = T(0), h_w2s = T(0), h_A = T(0), h_B = T(0);
        if (b0 + lane < nviews) {
            const View<T> v = views[(size_t)(b0 + lane) * vstride];
            T u, w2, w2s;
            pd_view_col_f32(g, v, xc, yc, u, w2, w2s);
            if (u >= lim.lo && u <= lim.u_hi) {
                const T A = g.dz * w2 / g.dv;
                const T B = (v.wz + g.zoff) * A - g.wv;
                // (A and B are finite wherever u passes its test; a view where they are not
                // is skipped, so that every voxel's v below is a number)
                if (!(off_boundary && !(A > T(0)) && v.wz * A - g.wv > lim.v_hi) && __builtin_isfinite(A) &&
                    __builtin_isfinite(B)) {
                    h_iu = (int)floor(u);
                    h_wul = u - T(h_iu);
                    h_w2s = w2s;
                    h_A = A;
                    h_B = B;
                    const T va = B, vb = fma_t(T(g.nz - 1), A, B);
                    int ra = 0, rb = g.nv - 2;
                    if (va == va && vb == vb) {
                        ra = pd_row_of<OFF_BOUNDARY>(va, lim, g.nv);
                        rb = pd_row_of<OFF_BOUNDARY>(vb, lim, g.nv);
                    }
                    h_rows = min(ra, rb) | ((max(ra, rb) + 1) << 16);
                }
            }
        }
        const int nb = min(kWave, nviews - b0);
        for (int k = 0; k < nb; ++k) {
            const int iu = rdl_i(h_iu, k);
            if (iu < 0) continue;
            const T wul = rdl(h_wul, k), w2s = rdl(h_w2s, k), A = rdl(h_A, k), B = rdl(h_B, k);
            const int rows = rdl_i(h_rows, k);
            const int r0 = rows & 0xffff, r1 = rows >> 16;
            const int rbase = r0 & ~(E - 1);
            // the previous view's reads of the row buffer come before these writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const Tin* p0 = proj + ((size_t)(b0 + k) * g.nu + iu) * g.nv;
            const Tin* p1 = p0 + g.nv;
            for (int ra = rbase + lane * E; ra <= r1; ra += kWave * E) {
                const int avail = g.nv - ra;
                const uint2 x0 = pd_load8<Tin>(p0 + (unsigned)ra, avail), x1 = pd_load8<Tin>(p1 + (unsigned)ra, avail);
                T a[E], bb[E], c[E];
                pd_unpack8<Tin>(x0, a);
                pd_unpack8<Tin>(x1, bb);
#pragma unroll
                for (int e = 0; e < E; ++e) c[e] = fmaf(wul, bb[e], fmaf(a[e], -wul, a[e]));
                float* dst = cs + (ra - rbase + kPdPad);
                if constexpr (E == 4)
                    *(float4*)dst = make_float4(c[0], c[1], c[2], c[3]);
                else
                    *(float2*)dst = make_float2(c[0], c[1]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const float* crow = cs + (kPdPad - rbase);
            // B in a vector register once per view (the voxel FMA takes A from a scalar one)
            T Bv;
            asm volatile("v_mov_b32 %0, %1" : "=v"(Bv) : "s"(B));
            auto voxel = [&](int c) {
                T vv = fma_t(izf[c], A, Bv);
                if (off_boundary) {
                    if (!(vv >= lim.lo && vv <= lim.v_hi)) return;
                } else {
                    if (vv < lim.lo) vv = T(0.5);
                    if (vv > lim.v_hi) vv = T(g.nv) - T(1.5);
                }
                // v >= 0 here, so truncation is floor; floor(v) lies in [r0, r1 - 1], the rows
                // staged: the voxel FMA is monotone in iz, so v lies between the first and the
                // last voxel's, and the row of v is non-decreasing in v
                const int iv = (int)vv;
                const T wvl = __builtin_amdgcn_fractf(vv);
                const T c0 = crow[iv], c1 = crow[iv + 1];
                img[c] = fmaf(w2s, fmaf(wvl, c1 - c0, c0), img[c]);
            };
            if (g.nz >= NZL * kWave) {
                // every lane's NZL voxels exist: one straight pass (the LDS reads overlap)
#pragma unroll
                for (int c = 0; c < NZL; ++c) voxel(c);
            } else {
#pragma unroll
                for (int c = 0; c < NZL; ++c)
                    if (c * kWave + lane < g.nz) voxel(c);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NZL; ++c) {
        const int iz = c * kWave + lane;
        if (iz < g.nz) {
            const T val = scale * img[c];
            ocol[iz] = cvt_store<Tout>(accumulate ? ld<T>(ocol + iz) + val : val);
        }
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
constexpr int kTU2 = 8;  // the same for the two-row-chunk ring form
// ring depth of the FP16 ring form: one 64-dword slot per column (at most 128 slices), two
// (at most 256); six entries keep five columns in flight in the LDS the FP32 form's four take
constexpr int kHalfRing1 = 6, kHalfRing2 = 4;

template <class Tin, class Tout, class T, int FOOT, int RPL, int SPLIT>
static void launch_proj(unsigned blocks, hipStream_t st, const Geom<T>& g, const View<T>* views, int vstride,
                        int nviews, const void* vol, const uint8_t* map, void* proj, int exhaustive) {
    const int form = g.sched_proj & 3;  // 0 auto, 1 ring, 2 register pipeline, 3 ring with two rows per lane
    // FP16 volumes of an even number of at most 256 slices: the ring form with half slots
    // (auto: the register pipeline, which measured faster at 128 slices)
    if constexpr (std::is_same<Tin, __half>::value && std::is_same<T, float>::value) if (form == 1) {
        constexpr int kRW = SPLIT > 1 ? 8 : kWavesPerBlock, kRS = SPLIT > 1 ? 8 : 1;
        if (g.nz % 2 == 0 && g.nz <= 2 * kWave) {
            hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU, kRS, 1, kRW, true, kHalfRing1>), dim3(blocks),
                               dim3(kRW * kWave), 0, st, g, views, vstride, nviews, (const __half*)vol, map,
                               (Tout*)proj, exhaustive);
            return;
        }
        if (g.nz % 2 == 0 && g.nz <= 4 * kWave) {
            hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU, kRS, 2, kRW, true, kHalfRing2>), dim3(blocks),
                               dim3(kRW * kWave), 0, st, g, views, vstride, nviews, (const __half*)vol, map,
                               (Tout*)proj, exhaustive);
            return;
        }
    }
    // form 3 (FP32 volumes of at most 256 slices, one wave per item): the ring form with two
    // 64-row chunks per item and eight-cell tiles (proj3_foot sizes the grid for it)
    if constexpr (std::is_same<Tin, float>::value && std::is_same<T, float>::value) if (form == 3 && SPLIT == 1) {
        constexpr int kRW = kWavesPerBlock;
        if (g.nz <= 2 * kWave) {
            hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU2, 1, 2, kRW, false, ring_len(2), 2>), dim3(blocks),
                               dim3(kRW * kWave), 0, st, g, views, vstride, nviews, (const float*)vol, map, (Tout*)proj,
                               exhaustive);
            return;
        }
        if (g.nz <= kFastSpan) {
            hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU2, 1, 3, kRW, false, ring_len(3), 2>), dim3(blocks),
                               dim3(kRW * kWave), 0, st, g, views, vstride, nviews, (const float*)vol, map, (Tout*)proj,
                               exhaustive);
            return;
        }
        hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU2, 1, 4, kRW, false, ring_len(4), 2>), dim3(blocks),
                           dim3(kRW * kWave), 0, st, g, views, vstride, nviews, (const float*)vol, map, (Tout*)proj,
                           exhaustive);
        return;
    }
    // FP32 volumes whose columns fit the three 64-voxel slots take the ring form (auto);
    // four slots (at most 256 slices) for whole work items (auto: 5% faster than the
    // register pipeline on the 720-view 512^2 x 256 operator) or on request
    if constexpr (std::is_same<Tin, float>::value && std::is_same<T, float>::value) if (form != 2) {
        // split items (SPLIT > 1: one item per block) go to blocks of eight waves
        constexpr int kRW = SPLIT > 1 ? 8 : kWavesPerBlock, kRS = SPLIT > 1 ? 8 : 1;
        if (g.nz <= 2 * kWave) {
            hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU, kRS, 2, kRW>), dim3(blocks), dim3(kRW * kWave), 0,
                               st, g, views, vstride, nviews, (const float*)vol, map, (Tout*)proj, exhaustive);
            return;
        }
        if (g.nz <= kFastSpan) {
            hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU, kRS, 3, kRW>), dim3(blocks), dim3(kRW * kWave), 0,
                               st, g, views, vstride, nviews, (const float*)vol, map, (Tout*)proj, exhaustive);
            return;
        }
        if ((form == 1 || (form == 0 && SPLIT == 1)) && g.nz <= 4 * kWave) {
            hipLaunchKernelGGL((cbct_proj_ring_kernel<Tout, FOOT, kTU, kRS, 4, kRW>), dim3(blocks), dim3(kRW * kWave), 0,
                               st, g, views, vstride, nviews, (const float*)vol, map, (Tout*)proj, exhaustive);
            return;
        }
    }
    hipLaunchKernelGGL((cbct_proj_kernel<Tin, Tout, T, FOOT, kTU, RPL, SPLIT>), dim3(blocks), dim3(kBlock), 0, st, g,
                       views, vstride, nviews, (const Tin*)vol, map, (Tout*)proj, exhaustive);
}

template <class Tin, class Tout, class T, int FOOT>
static void proj3_foot(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* vol,
                       const uint8_t* map, void* proj, int exhaustive, hipStream_t st) {
    // form 3: eight-cell tiles of two 64-row chunks, one wave per item (FP32 volumes of at
    // most 256 slices; other volumes take the automatic choice).  Automatic: taken for
    // detectors over 64 rows whose items need no split (measured on the 720-view
    // 512^2 x 256 pair at 1024 x 384: SF forward 146 -> 115 ms, DD 134 -> 102 ms)
    constexpr bool fp32 = std::is_same<Tin, float>::value && std::is_same<T, float>::value;
    const int form = g.sched_proj & 3, ssplit0 = (g.sched_proj >> 2) & 3;
    const long long items1 = (long long)nviews * ((g.nu + kTU - 1) / kTU) * ((g.nv + kWave - 1) / kWave);
    const bool two = fp32 && g.nz <= 4 * kWave &&
                     (form == 3 || (form == 0 && ssplit0 != 2 && g.nv > kWave && items1 >= 16384));
    const int tu = two ? kTU2 : kTU, rows = two ? 2 * kWave : kWave;
    const long long nut = (g.nu + tu - 1) / tu, nvc = (g.nv + rows - 1) / rows;
    const long long items = (long long)nviews * nut * nvc;
    if (two) {
        // the launch reads the form from the schedule: make it form 3 (the grid is sized for it)
        Geom<T> g3 = g;
        g3.sched_proj = (g.sched_proj & ~3) | 3;
        const unsigned blocks = (unsigned)((items + kWavesPerBlock - 1) / kWavesPerBlock);
        launch_proj<Tin, Tout, T, FOOT, 1, 1>(blocks, st, g3, views, vstride, nviews, vol, map, proj, exhaustive);
        return;
    }
    Geom<T> ga = g;
    if ((g.sched_proj & 3) == 3) ga.sched_proj = g.sched_proj & ~3;  // form 3 not applicable: automatic
    // few work items (short view subsets, thin detectors): split each item's strip over
    // the waves of a block so that the chip still holds several waves per SIMD; the plan's
    // schedule can force either form
    const int ssplit = (g.sched_proj >> 2) & 3;
    const bool split = ssplit == 1 ? false : (ssplit == 2 ? true : items < 16384);
    const unsigned blocks = (unsigned)(split ? items : (items + kWavesPerBlock - 1) / kWavesPerBlock);
    if (split) launch_proj<Tin, Tout, T, FOOT, 1, kWavesPerBlock>(blocks, st, ga, views, vstride, nviews, vol, map, proj, exhaustive);
    else launch_proj<Tin, Tout, T, FOOT, 1, 1>(blocks, st, ga, views, vstride, nviews, vol, map, proj, exhaustive);
}

template <class Tin, class Tout, class T>
static hipError_t proj3(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* vol,
                        const uint8_t* map, void* proj, int foot, int exhaustive, hipStream_t st) {
    if (nviews <= 0) return hipSuccess;
    if (g.nu >= 32768) return hipErrorNotSupported;  // footprint cells are packed in 16 bits
    if (g.nz > kMaxNz) return hipErrorNotSupported;   // voxel ranges are packed in kZBits bits
    if (foot == FOOT_DD)
        proj3_foot<Tin, Tout, T, FOOT_DD>(g, views, vstride, nviews, vol, map, proj, exhaustive, st);
    else
        proj3_foot<Tin, Tout, T, FOOT_TF>(g, views, vstride, nviews, vol, map, proj, exhaustive, st);
    return hipGetLastError();
}

// PdLimits for the T arithmetic of the device (float: the nearest floats on the safe side)
static PdLimits<double> pd_limits_d(int nu, int nv) {
    return PdLimits<double>{1e-6, (double)(nu - 1) - 1e-6, (double)(nv - 1) - 1e-6};
}
static float float_at_least(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}
static float float_at_most(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}
template <class T>
static PdLimits<T> pd_limits(int nu, int nv);
template <>
PdLimits<double> pd_limits<double>(int nu, int nv) { return pd_limits_d(nu, nv); }
template <>
PdLimits<float> pd_limits<float>(int nu, int nv) {
    const PdLimits<double> d = pd_limits_d(nu, nv);
    return PdLimits<float>{float_at_least(d.lo), float_at_most(d.u_hi), float_at_most(d.v_hi)};
}

template <class Tin, class Tout, class T, int NZL, bool SHORT>
static void launch_back_dd_tf(unsigned blocks, hipStream_t st, const Geom<T>& g, const View<T>* views, int vstride,
                              int nviews, const void* proj, const uint8_t* map, void* vol, int foot, T scale,
                              int accumulate, int rowwin) {
    if (foot == FOOT_DD)
        hipLaunchKernelGGL((cbct_back_kernel<Tin, Tout, T, FOOT_DD, NZL, SHORT, false>), dim3(blocks), dim3(kBlock),
                           back_lds_bytes<T>(rowwin), st, g, views, vstride, nviews, (const Tin*)proj, map,
                           (Tout*)vol, scale, accumulate, rowwin, NestArgs{}, g.sched_back & 1);
    else
        hipLaunchKernelGGL((cbct_back_kernel<Tin, Tout, T, FOOT_TF, NZL, SHORT, false>), dim3(blocks), dim3(kBlock),
                           back_lds_bytes<T>(rowwin), st, g, views, vstride, nviews, (const Tin*)proj, map,
                           (Tout*)vol, scale, accumulate, rowwin, NestArgs{}, g.sched_back & 1);
}

template <class Tin, class Tout, class T, int NZL>
static void launch_back(unsigned blocks, hipStream_t st, const Geom<T>& g, const View<T>* views, int vstride,
                        int nviews, const void* proj, const uint8_t* map, void* vol, int foot, int pd_mode, T scale,
                        int accumulate) {
    const int rowwin = back_rowwin(g.nv);
    if constexpr (std::is_same<T, float>::value) {
        // the staged FDK form unless the schedule asks for the direct one (sched_back bit 1)
        if (foot == 2 && (g.sched_back & 2) == 0 && g.nv <= kPdRows) {
            if (pd_mode)
                hipLaunchKernelGGL((cbct_back_pd_stage_kernel<Tin, Tout, NZL, 1>), dim3(blocks), dim3(kBlock), 0, st,
                                   g, views, vstride, nviews, (const Tin*)proj, map, (Tout*)vol, scale, accumulate,
                                   pd_limits<T>(g.nu, g.nv));
            else
                hipLaunchKernelGGL((cbct_back_pd_stage_kernel<Tin, Tout, NZL, 0>), dim3(blocks), dim3(kBlock), 0, st,
                                   g, views, vstride, nviews, (const Tin*)proj, map, (Tout*)vol, scale, accumulate,
                                   pd_limits<T>(g.nu, g.nv));
            return;
        }
    }
    if (foot == 2 && pd_mode)
        hipLaunchKernelGGL((cbct_back_pd_kernel<Tin, Tout, T, NZL, 1>), dim3(blocks), dim3(kBlock), 0, st, g, views,
                           vstride, nviews, (const Tin*)proj, map, (Tout*)vol, scale, accumulate,
                           pd_limits<T>(g.nu, g.nv));
    else if (foot == 2)
        hipLaunchKernelGGL((cbct_back_pd_kernel<Tin, Tout, T, NZL, 0>), dim3(blocks), dim3(kBlock), 0, st, g, views,
                           vstride, nviews, (const Tin*)proj, map, (Tout*)vol, scale, accumulate,
                           pd_limits<T>(g.nu, g.nv));
    else if (g.nv <= kWave)
        launch_back_dd_tf<Tin, Tout, T, NZL, true>(blocks, st, g, views, vstride, nviews, proj, map, vol, foot, scale,
                                                   accumulate, rowwin);
    else
        launch_back_dd_tf<Tin, Tout, T, NZL, false>(blocks, st, g, views, vstride, nviews, proj, map, vol, foot, scale,
                                                    accumulate, rowwin);
}

template <class Tin, class Tout, class T>
static hipError_t back3(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* proj,
                        const uint8_t* map, void* vol, int foot, int pd_mode, T scale, int accumulate,
                        hipStream_t st) {
    const long long cols = (long long)g.nx * g.ny;
    const unsigned blocks = (foot == 2 || (g.sched_back & 1)) ? (unsigned)pd_blocks(g.nx, g.ny)
                                                             : (unsigned)((cols + kWavesPerBlock - 1) / kWavesPerBlock);
    if (cols == 0 || blocks == 0) return hipSuccess;
    if (foot == 2 && (g.nu < 2 || g.nv < 2)) return hipErrorInvalidValue;  // bilinear needs 2x2 cells
    if (foot != 2 && (long long)kWave * g.nu * g.nv * (long long)sizeof(Tin) > 0xffffffffLL)
        return hipErrorNotSupported;  // 64 views' sinograms are addressed with 32-bit byte offsets
    if (g.nz > kBackSlab) {
        // a lane holds at most 8 voxels of a column: taller volumes are backprojected in
        // z-slabs of kBackSlab slices, each a launch over its slices of every column
        for (int z0 = 0; z0 < g.nz; z0 += kBackSlab) {
            Geom<T> gs = g;
            gs.nz = std::min(kBackSlab, g.nz - z0);
            gs.zstride = g.zstride;
            gs.zoff = g.zoff + T(z0);
            hipError_t e = back3<Tin, Tout, T>(gs, views, vstride, nviews, proj, map,
                                               static_cast<Tout*>(vol) + z0, foot, pd_mode, scale, accumulate, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const int nzl = (g.nz + kWave - 1) / kWave;
    if (nzl <= 1) launch_back<Tin, Tout, T, 1>(blocks, st, g, views, vstride, nviews, proj, map, vol, foot, pd_mode, scale, accumulate);
    else if (nzl <= 2) launch_back<Tin, Tout, T, 2>(blocks, st, g, views, vstride, nviews, proj, map, vol, foot, pd_mode, scale, accumulate);
    else if (nzl <= 4) launch_back<Tin, Tout, T, 4>(blocks, st, g, views, vstride, nviews, proj, map, vol, foot, pd_mode, scale, accumulate);
    else launch_back<Tin, Tout, T, 8>(blocks, st, g, views, vstride, nviews, proj, map, vol, foot, pd_mode, scale, accumulate);
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

// backprojection of a view selection into the gradient, fused with the Nesterov-SQS update
// of every voxel (Huber penalty); FP32, detectors of at most 64 rows
template <class Tin, int NZL>
static void launch_back_nesterov(unsigned blocks, hipStream_t st, const Geom<float>& g, const View<float>* views,
                                 int vstride, int nviews, const void* proj, const uint8_t* map, int foot, float scale,
                                 const NestArgs& na) {
    const int rowwin = back_rowwin(g.nv);
    if (foot == FOOT_DD)
        hipLaunchKernelGGL((cbct_back_kernel<Tin, float, float, FOOT_DD, NZL, true, true>), dim3(blocks), dim3(kBlock),
                           back_lds_bytes<float>(rowwin), st, g, views, vstride, nviews, (const Tin*)proj, map,
                           (float*)nullptr, scale, 0, rowwin, na, 0);
    else
        hipLaunchKernelGGL((cbct_back_kernel<Tin, float, float, FOOT_TF, NZL, true, true>), dim3(blocks), dim3(kBlock),
                           back_lds_bytes<float>(rowwin), st, g, views, vstride, nviews, (const Tin*)proj, map,
                           (float*)nullptr, scale, 0, rowwin, na, 0);
}

template <class Tin>
static hipError_t back_nesterov3(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                 const void* proj, const uint8_t* map, int foot, float scale, const NestArgs& na,
                                 hipStream_t st) {
    if (g.nv > kWave) return hipErrorNotSupported;
    if (g.nz > 8 * kWave) return hipErrorNotSupported;
    if ((long long)kWave * g.nu * g.nv * (long long)sizeof(Tin) > 0xffffffffLL) return hipErrorNotSupported;
    const long long cols = (long long)g.nx * g.ny;
    const unsigned blocks = (unsigned)((cols + kWavesPerBlock - 1) / kWavesPerBlock);
    if (cols == 0) return hipSuccess;
    const int nzl = (g.nz + kWave - 1) / kWave;
    if (nzl <= 1) launch_back_nesterov<Tin, 1>(blocks, st, g, views, vstride, nviews, proj, map, foot, scale, na);
    else if (nzl <= 2) launch_back_nesterov<Tin, 2>(blocks, st, g, views, vstride, nviews, proj, map, foot, scale, na);
    else if (nzl <= 4) launch_back_nesterov<Tin, 4>(blocks, st, g, views, vstride, nviews, proj, map, foot, scale, na);
    else launch_back_nesterov<Tin, 8>(blocks, st, g, views, vstride, nviews, proj, map, foot, scale, na);
    return hipGetLastError();
}

hipError_t cbct_back_nesterov_launch(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                     const void* proj, int proj_dt, const uint8_t* map, int foot, float scale,
                                     const NestArgs& na, hipStream_t st) {
    if (foot != FOOT_DD && foot != FOOT_TF) return hipErrorInvalidValue;
    if (nviews <= 0) return hipErrorInvalidValue;
    switch (proj_dt) {
        case DT_F32: return back_nesterov3<float>(g, views, vstride, nviews, proj, map, foot, scale, na, st);
        case DT_F16: return back_nesterov3<__half>(g, views, vstride, nviews, proj, map, foot, scale, na, st);
        case DT_BF16: return back_nesterov3<__hip_bfloat16>(g, views, vstride, nviews, proj, map, foot, scale, na, st);
        default: return hipErrorInvalidValue;
    }
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
