// Projection-matrix (PMat / C-arm) cone-beam operators: the reference's geom/pcbct
// gateway (cbct_geom_pmat_mex.c), whose views are given by a 3x4 projection matrix, a
// camera position, a source-to-detector distance and a detector offset each.
//
//   proj,dd  (proj,tf runs it too)   geom/pcbct/cbct_proj_pmat.c:8-215
//   back,dd  (back,tf runs it too)   geom/pcbct/cbct_back_pmat.c:137-321
//   back,pd / back,hpd               geom/pcbct/cbct_back_pmat.c:9-134
//
// Volumes are device ZXY (z fastest), sinograms [view][u][v].  FP32 throughout; the
// footprint arithmetic follows the reference's float / double mix with no contraction,
// so footprints, weights and amplitudes are the reference's own.
//
//  * back,pd / back,hpd and back,dd gather: one wavefront per image column, lanes are the
//    column's voxels, every view in order; one writer per voxel, no atomics.
//  * proj,dd: voxel-driven like the reference (its per-cell amplitude applied afterwards),
//    one block per (view, 16 x 16 image columns, 64 slices), the voxels' footprint cells
//    summed in an LDS patch over the block's detector box and the patch added to the view
//    with one FP32 global atomic per cell.  Results agree with the reference to FP32
//    rounding; the order of the additions is not fixed, so the last bits can vary between
//    runs.  The footprint of a general projection matrix has no cheap inverse (the circular
//    kernels' strip walk assumes the orbit), so the form stays a scatter.
#include <algorithm>

#include "internal.h"

namespace ctsim {

constexpr int kWave = 64;      // wavefront width
constexpr int kPmFootU = 16;  // u cells per footprint (iumin + 15, cbct_proj_pmat.c:118)
constexpr int kPmFootV = 10;  // v rows per voxel (FOOTPRINT_SIZE_V, ct_def.h:82)

// u footprint of column (ix, iy) in view pv (cbct_proj_pmat.c:69-147): cells
// [iumin, iumax] covering [umin, umax], ampu, and v0 / f0 and dv/dz of the column; false
// when the column misses the detector.
__device__ __forceinline__ bool pmat_ucell(const PmatConst& c, const PmatView& pv, int ix, int iy, float& ampu,
                                           float& umin, float& umax, int& iumin, int& iumax, float& v0f,
                                           float& dvdz) {
#pragma clang fp contract(off)
    const float* pm = pv.pm;
    const float xc = ((float)ix + c.wx) * c.dx;
    const float yc = ((float)iy + c.wy) * c.dy;
    const float u0 = xc * pm[0] + yc * pm[3] + c.wz * c.dz * pm[6] + pm[9];
    const float v0 = xc * pm[1] + yc * pm[4] + c.wz * c.dz * pm[7] + pm[10];
    const float f0 = xc * pm[2] + yc * pm[5] + pm[11];
    const float dsp_x = (float)fabs((double)(pv.cx - xc));
    const float dsp_y = (float)fabs((double)(pv.cy - yc));
    const double dist = sqrt((double)(dsp_x * dsp_x + dsp_y * dsp_y));
    if (fabs((double)dsp_x) > fabs((double)dsp_y)) {
        umin = (u0 - c.dy * pm[3] / 2.0f) / (f0 - c.dy * pm[5] / 2.0f);
        umax = (u0 + c.dy * pm[3] / 2.0f) / (f0 + c.dy * pm[5] / 2.0f);
        ampu = (float)((double)c.dx * dist / (double)dsp_x / 10.0);
    } else {
        umin = (u0 - c.dx * pm[0] / 2.0f) / (f0 - c.dx * pm[2] / 2.0f);
        umax = (u0 + c.dx * pm[0] / 2.0f) / (f0 + c.dx * pm[2] / 2.0f);
        ampu = (float)((double)c.dx * dist / (double)dsp_y / 10.0);
    }
    if (umin >= umax) {
        const float t = umax;
        umax = umin;
        umin = t;
    }
    const double a = (double)umin + 0.5, b = (double)umax + 0.5;
    umin = (float)(a > 0.0 ? a : 0.0);
    umax = (float)(b < (double)c.nu ? b : (double)c.nu);
    if (umin > umax) return false;
    iumin = (int)floorf(umin);
    iumax = min((int)floorf(umax), c.nu - 1);
    iumax = min(iumax, iumin + kPmFootU - 1);
    v0f = v0 / f0;
    dvdz = pm[7] / f0;
    return true;
}

// weight of cell i in [imin, imax] of a footprint covering [lo, hi] (cbct_proj_pmat.c:
// 126-145 for u, 170-188 for v): the covered fraction of the first and last cells, 1
// between them, hi - lo for a one-cell footprint
__device__ __forceinline__ float pmat_weight(float lo, float hi, int imin, int imax, int i) {
#pragma clang fp contract(off)
    if (imin == imax) return hi - lo;
    return i == imin ? (float)((double)(float)imin + 1.0 - (double)lo) : (i == imax ? hi - (float)imax : 1.0f);
}

// v footprint of voxel iz (cbct_proj_pmat.c:155-190): rows [ivmin, ivmax] covering
// [vmin, vmax]
__device__ __forceinline__ bool pmat_vcell(const PmatConst& c, int iz, float v0f, float dvdz, float& vmin,
                                           float& vmax, int& ivmin, int& ivmax) {
#pragma clang fp contract(off)
    const double fz = (double)(float)iz;
    double lo, hi;
    if (dvdz > 0.0f) {
        lo = (double)v0f + (fz - 0.5) * (double)c.dz * (double)dvdz;
        hi = (double)v0f + (fz + 0.5) * (double)c.dz * (double)dvdz;
    } else {
        lo = (double)v0f + (fz + 0.5) * (double)c.dz * (double)dvdz;
        hi = (double)v0f + (fz - 0.5) * (double)c.dz * (double)dvdz;
    }
    vmin = (float)lo;
    vmax = (float)hi;
    const double a = (double)vmin + 0.5, b = (double)vmax + 0.5;
    vmin = (float)(a > 0.0 ? a : 0.0);
    vmax = (float)(b < (double)c.nv ? b : (double)c.nv);
    if (vmin > vmax) return false;
    ivmin = (int)floorf(vmin);
    ivmax = min((int)floorf(vmax), c.nv - 1);
    ivmax = min(ivmax, ivmin + kPmFootV - 1);
    return true;
}

// amplitude of detector cell (iu, iv) of a view (cbct_proj_pmat.c:197-210); the v
// coordinate takes the u offset, as the reference writes it
__device__ __forceinline__ float pmat_cell_amp(const PmatConst& c, const PmatView& pv, int iu, int iv) {
#pragma clang fp contract(off)
    const float u = ((float)iu + c.wu + pv.ou) * c.du;
    const float uu = pv.sdd * pv.sdd + u * u;
    const float v = ((float)iv + c.wv + pv.ou) * c.dv;
    return (float)sqrt((double)((uu + v * v) / uu));
}

__global__ __launch_bounds__(256) void pmat_amp_kernel(PmatConst c, const PmatView* __restrict__ views,
                                                       const float* __restrict__ in, float* __restrict__ out,
                                                       size_t n) {
    const size_t cells = (size_t)c.nu * c.nv;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(i / cells);
        const int r = (int)(i - (size_t)b * cells);
        const int iu = r / c.nv, iv = r - iu * c.nv;
        out[i] = in[i] * pmat_cell_amp(c, views[b], iu, iv);
    }
}

// forward: one block per (view, tile of kPtX x kPtY image columns, chunk of kWave
// slices).  Its threads first take one column each: the u footprint goes to LDS and the
// block's detector box (the tile's u cells by the chunk's v rows, one row of margin) is
// reduced there.  The columns are then ranked by the centre of their u footprint and cut
// into kSeg runs of neighbouring u; in two phases each wave walks one run (runs 0, 2, 4, 6,
// then 1, 3, 5, 7), its lanes on the chunk's voxels, adding the footprint cells to an LDS
// patch over the box.  Within one (column, u cell) the lanes hold distinct v rows whenever
// the v footprint advances at least one row per slice, so a cell that no other wave's
// current run reaches takes a plain read-add-write, in order; a cell shared with another
// run, or a column whose slices share a first row, takes an LDS atomic.  (A column with a
// cell outside the box, or a voxel over more than kFastRows rows, takes atomics for all of
// its cells.)  After a barrier
// every patch cell that received anything is added to the view with one FP32 global
// atomic.  A cell outside the box, or every cell of a block whose box exceeds kPatch,
// takes a global atomic directly.
constexpr int kPtX = 16, kPtY = 16;  // image columns per tile
constexpr int kSeg = 8;              // runs of columns, two per wave
constexpr int kPatch = 7168;         // LDS patch cells (28 KB; four blocks per CU)
constexpr int kFastRows = 4;         // v rows per voxel on the in-patch path

struct PmatCol {
    float umin, umax, ampu, v0f, dvdz;
    int iumin, iumax, ok;
};

__global__ __launch_bounds__(256) void pmat_proj_kernel(PmatConst c, const PmatView* __restrict__ views,
                                                        const float* __restrict__ vol, const uint8_t* __restrict__ map,
                                                        float* __restrict__ acc, int ntx, int nty) {
#pragma clang fp contract(off)
    static_assert(kPtX * kPtY == 256 && kSeg == 2 * (256 / kWave), "one column per thread, two runs per wave");
    __shared__ PmatCol cols[kPtX * kPtY];
    __shared__ float ukey[kPtX * kPtY];
    __shared__ int order[kPtX * kPtY];
    __shared__ float patch[kPatch];
    __shared__ int box[4];          // u lo, u hi, v lo, v hi
    __shared__ int run[2 * kSeg];  // u lo, u hi of each run
    __shared__ int nok_s;
    const int t = threadIdx.x;
    int r = blockIdx.x;
    const int b = r % c.nviews;
    r /= c.nviews;
    const int tx = r % ntx;
    r /= ntx;
    const int ty = r % nty;
    const int z0 = (r / nty) * kWave;
    const int z1 = min(z0 + kWave, c.nz) - 1;
    if (t < 2 * kSeg) run[t] = (t & 1) ? INT_MIN : INT_MAX;
    if (t == 0) {
        box[0] = INT_MAX;
        box[1] = INT_MIN;
        box[2] = INT_MAX;
        box[3] = INT_MIN;
        nok_s = 0;
    }
    __syncthreads();
    {
        const int ix = tx * kPtX + t % kPtX, iy = ty * kPtY + t / kPtX;
        PmatCol pc;
        pc.ok = 0;
        if (ix < c.nx && iy < c.ny && map[iy * c.nx + ix])
            pc.ok = pmat_ucell(c, views[b], ix, iy, pc.ampu, pc.umin, pc.umax, pc.iumin, pc.iumax, pc.v0f, pc.dvdz);
        if (pc.ok) {
            atomicAdd(&nok_s, 1);
            // v rows of the chunk's end faces; v is linear in z, so they bound every voxel's
            const float e0 = pc.v0f + ((float)z0 - 0.5f) * c.dz * pc.dvdz;
            const float e1 = pc.v0f + ((float)z1 + 0.5f) * c.dz * pc.dvdz;
            const int vlo = max((int)floorf(fminf(e0, e1) + 0.5f) - 1, 0);
            const int vhi = min((int)floorf(fmaxf(e0, e1) + 0.5f) + 1, c.nv - 1);
            if (vlo <= vhi) {
                atomicMin(&box[0], pc.iumin);
                atomicMax(&box[1], pc.iumax);
                atomicMin(&box[2], vlo);
                atomicMax(&box[3], vhi);
            }
        }
        cols[t] = pc;
        ukey[t] = pc.ok ? pc.umin + pc.umax : HUGE_VALF;
    }
    __syncthreads();
    const bool boxed = box[0] <= box[1] && box[2] <= box[3];
    const int ulo = box[0], vlo = box[2];
    const int pu_n = boxed ? box[1] - ulo + 1 : 0, pv_n = boxed ? box[3] - vlo + 1 : 0;
    const bool use_patch = boxed && (long long)pu_n * pv_n <= kPatch;
    const int pcells = use_patch ? pu_n * pv_n : 0;
    const int nok = uni(nok_s);
    for (int i = t; i < pcells; i += 256) patch[i] = 0.0f;
    {
        // rank by footprint centre, ties by thread: the order is a permutation
        const float kt = ukey[t];
        int rank = 0;
        for (int s2 = 0; s2 < kPtX * kPtY; ++s2) {
            const float ks = ukey[s2];
            rank += (ks < kt || (ks == kt && s2 < t)) ? 1 : 0;
        }
        order[rank] = t;
    }
    __syncthreads();
    // run g holds ranks [ceil(g nok / kSeg), ceil((g + 1) nok / kSeg)); its u cells
    if (t < nok) {
        const int g = t * kSeg / nok;
        const PmatCol& pc = cols[order[t] & 255];
        atomicMin(&run[2 * g], pc.iumin);
        atomicMax(&run[2 * g + 1], pc.iumax);
    }
    __syncthreads();
    float* pview = acc + (size_t)b * c.nu * c.nv;
    const int lane = lane_id();
    const int w = uni(t / kWave);
    const int iz = z0 + lane;
    for (int h = 0; h < 2; ++h) {
        const int g = 2 * w + h;
        // u cells of the runs the other waves walk in this phase
        int olo[256 / kWave - 1], ohi[256 / kWave - 1];
#pragma unroll
        for (int w2 = 0, o = 0; w2 < 256 / kWave; ++w2) {
            if (w2 == w) continue;
            olo[o] = uni(run[2 * (2 * w2 + h)]);
            ohi[o] = uni(run[2 * (2 * w2 + h) + 1]);
            ++o;
        }
        const int p0 = (g * nok + kSeg - 1) / kSeg, p1 = ((g + 1) * nok + kSeg - 1) / kSeg;
        for (int p = p0; p < p1; ++p) {
            const int k = uni(order[p]) & 255;
            PmatCol pc = cols[k];
            pc.umin = uni(pc.umin);
            pc.umax = uni(pc.umax);
            pc.ampu = uni(pc.ampu);
            pc.v0f = uni(pc.v0f);
            pc.dvdz = uni(pc.dvdz);
            const int iumin = uni(pc.iumin), iumax = uni(pc.iumax);
            const int ix = tx * kPtX + k % kPtX, iy = ty * kPtY + k / kPtX;
            const float val = iz < c.nz ? vol[((size_t)iy * c.nx + ix) * c.nz + iz] : 0.0f;
            float vmin = 0.0f, vmax = 0.0f;
            int ivmin = 0, ivmax = -1;
            const bool vok = iz < c.nz && pmat_vcell(c, iz, pc.v0f, pc.dvdz, vmin, vmax, ivmin, ivmax);
            const bool wr = vok && val != 0.0f;
            // plain read-add-write needs distinct first rows on the lanes that write
            const int prev_iv = __shfl_up(ivmin, 1);
            const int prev_ok = __shfl_up(vok ? 1 : 0, 1);
            const bool plain = !__any(lane > 0 && vok && prev_ok && prev_iv == ivmin);
            const int nrows = wr ? ivmax - ivmin + 1 : 0;
            const bool inbox = use_patch && iumin >= ulo && iumax < ulo + pu_n &&
                               !__any(wr && (ivmin < vlo || ivmax >= vlo + pv_n || nrows > kFastRows));
            const float voxel = pc.ampu * val;
            const float wv0 = pmat_weight(vmin, vmax, ivmin, ivmax, ivmin);
            const float wv1 = pmat_weight(vmin, vmax, ivmin, ivmax, ivmax);
            if (inbox) {
                // every cell of the column lies in the patch and every voxel spans at most
                // kFastRows rows: rows walked in step, lanes past their last row idle
                const float wu0 = pmat_weight(pc.umin, pc.umax, iumin, iumax, iumin);
                const float wu1 = pmat_weight(pc.umin, pc.umax, iumin, iumax, iumax);
                for (int iu = iumin; iu <= iumax; ++iu) {
                    const float wu = iu == iumin ? wu0 : (iu == iumax ? wu1 : 1.0f);
                    bool own = plain;
#pragma unroll
                    for (int o = 0; o < 256 / kWave - 1; ++o)
                        if (iu >= olo[o] && iu <= ohi[o]) own = false;
                    float* row = patch + (iu - ulo) * pv_n + (ivmin - vlo);
                    if (own) {
#pragma unroll
                        for (int j = 0; j < kFastRows; ++j) {
                            if (!__any(j < nrows)) break;
                            if (j < nrows) {
                                const float add = wu * (j == 0 ? wv0 : (ivmin + j == ivmax ? wv1 : 1.0f)) * voxel;
                                row[j] = row[j] + add;
                            }
                            // the next row's read must follow this write: the LDS runs
                            // one wave's operations in order, and the empty asm keeps the
                            // compiler from moving memory operations across it
                            asm volatile("" ::: "memory");
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < kFastRows; ++j) {
                            if (!__any(j < nrows)) break;
                            if (j < nrows)
                                atomicAdd(&row[j], wu * (j == 0 ? wv0 : (ivmin + j == ivmax ? wv1 : 1.0f)) * voxel);
                        }
                    }
                }
                continue;
            }
            if (!wr) continue;
            for (int iu = iumin; iu <= iumax; ++iu) {
                const float wu = pmat_weight(pc.umin, pc.umax, iumin, iumax, iu);
                const int pu = iu - ulo;
                for (int iv = ivmin; iv <= ivmax; ++iv) {
                    const float add = wu * (iv == ivmin ? wv0 : (iv == ivmax ? wv1 : 1.0f)) * voxel;
                    const int pv = iv - vlo;
                    if (use_patch && (unsigned)pu < (unsigned)pu_n && (unsigned)pv < (unsigned)pv_n)
                        atomicAdd(&patch[pu * pv_n + pv], add);
                    else
                        atomicAdd(&pview[(size_t)iu * c.nv + iv], add);
                }
            }
        }
        __syncthreads();
    }
    for (int i = t; i < pcells; i += 256) {
        const float v = patch[i];
        if (v != 0.0f) {
            const int pu = i / pv_n;
            atomicAdd(&pview[(size_t)(ulo + pu) * c.nv + vlo + (i - pu * pv_n)], v);
        }
    }
}

// back,dd: one wave per image column, lanes are voxels; sino is already scaled by the
// cell amplitudes
template <int NZL>
__global__ __launch_bounds__(256) void pmat_back_dd_kernel(PmatConst c, const PmatView* __restrict__ views,
                                                           const float* __restrict__ sino,
                                                           const uint8_t* __restrict__ map, float* __restrict__ vol) {
#pragma clang fp contract(off)
    const int col = xcd_remap(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (col >= c.nx * c.ny) return;
    const int lane = lane_id();
    float* ocol = vol + (size_t)col * c.nz;
    float img[NZL];
#pragma unroll
    for (int k = 0; k < NZL; ++k) img[k] = 0.0f;
    if (map[col]) {
        const int ix = col % c.nx, iy = col / c.nx;
        for (int b = 0; b < c.nviews; ++b) {
            // the column's u footprint is the wave's: kept in scalar registers, and the
            // weights of its first and last cells (1 between) formed once per view
            const PmatView pv = views[b];
            float ampu, umin, umax, v0f, dvdz;
            int iumin, iumax;
            if (!pmat_ucell(c, pv, ix, iy, ampu, umin, umax, iumin, iumax, v0f, dvdz)) continue;
            iumin = uni(iumin);
            iumax = uni(iumax);
            ampu = uni(ampu);
            v0f = uni(v0f);
            dvdz = uni(dvdz);
            const float wu0 = uni(pmat_weight(umin, umax, iumin, iumax, iumin));
            const float wu1 = uni(pmat_weight(umin, umax, iumin, iumax, iumax));
            const float* pview = sino + (size_t)b * c.nu * c.nv;
#pragma unroll
            for (int k = 0; k < NZL; ++k) {
                const int iz = k * kWave + lane;
                if (iz >= c.nz) continue;
                float vmin, vmax;
                int ivmin, ivmax;
                if (!pmat_vcell(c, iz, v0f, dvdz, vmin, vmax, ivmin, ivmax)) continue;
                const float wv0 = pmat_weight(vmin, vmax, ivmin, ivmax, ivmin);
                const float wv1 = pmat_weight(vmin, vmax, ivmin, ivmax, ivmax);
                float voxel = 0.0f;
                for (int iu = iumin; iu <= iumax; ++iu) {
                    const float wu = iu == iumin ? wu0 : (iu == iumax ? wu1 : 1.0f);
                    const float* prow = pview + (size_t)iu * c.nv;
                    for (int iv = ivmin; iv <= ivmax; ++iv)
                        voxel += prow[iv] * wu * (iv == ivmin ? wv0 : (iv == ivmax ? wv1 : 1.0f));
                }
                img[k] += ampu * voxel;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NZL; ++k) {
        const int iz = k * kWave + lane;
        if (iz < c.nz) ocol[iz] = img[k];
    }
}

// back,pd (OFF = 1: skip below the detector, stop the column above it) and back,hpd
// (OFF = 0: clamp v to [0.5, nv - 1.5]); one wave per image column, lanes are voxels
template <int NZL, int OFF>
__global__ __launch_bounds__(256) void pmat_back_pd_kernel(PmatConst c, const PmatView* __restrict__ views,
                                                           const float* __restrict__ sino,
                                                           const uint8_t* __restrict__ map, float* __restrict__ vol) {
#pragma clang fp contract(off)
    const int col = xcd_remap(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (col >= c.nx * c.ny) return;
    const int lane = lane_id();
    float* ocol = vol + (size_t)col * c.nz;
    float img[NZL];
#pragma unroll
    for (int k = 0; k < NZL; ++k) img[k] = 0.0f;
    if (map[col]) {
        const int ix = col % c.nx, iy = col / c.nx;
        const float xc = ((float)ix + c.wx) * c.dx;
        const float yc = ((float)iy + c.wy) * c.dy;
        for (int b = 0; b < c.nviews; ++b) {
            const PmatView pv = views[b];
            const float* pm = pv.pm;
            const float u0 = xc * pm[0] + yc * pm[3] + c.wz * c.dz * pm[6] + pm[9];
            const float v0 = xc * pm[1] + yc * pm[4] + c.wz * c.dz * pm[7] + pm[10];
            const float f0 = xc * pm[2] + yc * pm[5] + c.wz * c.dz * pm[8] + pm[11];
            const float w2s = pv.sdd * pv.sdd / ((pv.cx - xc) * (pv.cx - xc) + (pv.cy - yc) * (pv.cy - yc));
            float uu[NZL], vv[NZL];
            bool uok[NZL];
            // back,pd breaks out of the column at the first voxel (in z order) whose u is on
            // the detector and whose v is past its top: the lowest such iz over the lanes
            int first_break = c.nz;
#pragma unroll
            for (int k = 0; k < NZL; ++k) {
                const int iz = k * kWave + lane;
                const float fz = (float)iz * c.dz;
                const float den = f0 + fz * pm[8];
                uu[k] = (u0 + fz * pm[6]) / den;
                vv[k] = (v0 + fz * pm[7]) / den;
                uok[k] = iz < c.nz && uu[k] >= c.lo && uu[k] <= c.u_hi;
                if (OFF) {
                    const unsigned long long brk = __ballot(uok[k] && vv[k] > c.v_hi);
                    if (brk && first_break == c.nz) first_break = k * kWave + __ffsll((long long)brk) - 1;
                }
            }
#pragma unroll
            for (int k = 0; k < NZL; ++k) {
                const int iz = k * kWave + lane;
                if (!uok[k] || iz >= first_break) continue;
                float v = vv[k];
                if (OFF) {
                    if (!(v >= c.lo)) continue;
                } else {
                    v = (double)v > 0.5 ? v : 0.5f;
                    v = (double)v < (double)c.nv - 1.5 ? v : (float)((double)c.nv - 1.5);
                }
                const float u = uu[k];
                const int iu = (int)floorf(u);
                const float wul = u - (float)iu;
                const int iv = max(0, min((int)floorf(v), c.nv - 2));
                const float wvl = v - (float)iv;
                const float* p = sino + ((size_t)b * c.nu + iu) * c.nv + iv;
                img[k] = img[k] + w2s * ((1.0f - wul) * ((1.0f - wvl) * p[0] + wvl * p[1]) +
                                         wul * ((1.0f - wvl) * p[c.nv] + wvl * p[c.nv + 1]));
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NZL; ++k) {
        const int iz = k * kWave + lane;
        if (iz < c.nz) ocol[iz] = img[k];
    }
}

template <int NZL>
static hipError_t pmat_back_nzl(int kind, const PmatConst& c, const PmatView* views, const float* sino,
                                const uint8_t* map, float* vol, hipStream_t st) {
    const unsigned blocks = (unsigned)((c.nx * c.ny + 3) / 4);
    if (kind == PMAT_BACK_DD)
        hipLaunchKernelGGL(pmat_back_dd_kernel<NZL>, dim3(blocks), dim3(256), 0, st, c, views, sino, map, vol);
    else if (kind == PMAT_BACK_PD)
        hipLaunchKernelGGL((pmat_back_pd_kernel<NZL, 1>), dim3(blocks), dim3(256), 0, st, c, views, sino, map, vol);
    else
        hipLaunchKernelGGL((pmat_back_pd_kernel<NZL, 0>), dim3(blocks), dim3(256), 0, st, c, views, sino, map, vol);
    return hipGetLastError();
}

hipError_t pmat_launch(int kind, const PmatConst& c, const PmatView* views, const uint8_t* map, const float* in,
                       float* out, float* tmp, hipStream_t st) {
    if (c.nz > 8 * kWave) return hipErrorNotSupported;
    const size_t nsino = (size_t)c.nviews * c.nu * c.nv;
    if (kind == PMAT_PROJ_DD) {
        hipError_t e = fill_launch(out, DT_F32, nsino, 0.0, st);
        if (e != hipSuccess) return e;
        const int ntx = (c.nx + kPtX - 1) / kPtX, nty = (c.ny + kPtY - 1) / kPtY;
        const long long blocks = (long long)c.nviews * ntx * nty * ((c.nz + kWave - 1) / kWave);
        if (blocks > INT_MAX) return hipErrorInvalidValue;
        hipLaunchKernelGGL(pmat_proj_kernel, dim3((unsigned)blocks), dim3(256), 0, st, c, views, in, map, out, ntx,
                           nty);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(pmat_amp_kernel, dim3((unsigned)std::min<size_t>((nsino + 255) / 256, 65536)), dim3(256), 0,
                           st, c, views, out, out, nsino);
        return hipGetLastError();
    }
    const float* src = in;
    if (kind == PMAT_BACK_DD) {
        // the reference scales its input in place (cbct_back_pmat.c:190-202); here a copy
        hipLaunchKernelGGL(pmat_amp_kernel, dim3((unsigned)std::min<size_t>((nsino + 255) / 256, 65536)), dim3(256), 0,
                           st, c, views, in, tmp, nsino);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        src = tmp;
    }
    const int nzl = (c.nz + kWave - 1) / kWave;
    if (nzl <= 1) return pmat_back_nzl<1>(kind, c, views, src, map, out, st);
    if (nzl <= 2) return pmat_back_nzl<2>(kind, c, views, src, map, out, st);
    if (nzl <= 4) return pmat_back_nzl<4>(kind, c, views, src, map, out, st);
    return pmat_back_nzl<8>(kind, c, views, src, map, out, st);
}

}  // namespace ctsim
