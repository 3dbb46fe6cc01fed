// Elementwise, stencil and reduction kernels of the PWLS ordered-subset loop
// (statsrecon/pwls/reconPwlsSeNesterovSqs.m:59-137, statsrecon/pwls/reconPwlsSeSqs.m:51-110)
// and the Huber penalty (penalty/huberPenalty.m:101-161, helpers :222-238).
//
// Volumes are float32 in device ZXY order (z fastest).  Sums that the loop reports
// (data-fit cost, penalty cost, RMSD) accumulate in FP64 with a fixed-order two-level
// reduction: every block writes a partial, the last block to finish adds the partials
// in index order.  Results are therefore bit-reproducible run to run.
#include <mutex>

#include "internal.h"
#include "penalty.h"

namespace ctsim {

constexpr int kRB = 256;        // reduction block size
constexpr int kRedBlocks = 1024;  // fixed grid of reduction kernels

// ---------------------------------------------------------------------------------
// fill / convert
// ---------------------------------------------------------------------------------
template <class D>
__global__ void fill_kernel(D* __restrict__ p, size_t n, double value) {
    const D v = cvt_store<D>((float)value);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_f64_kernel(double* __restrict__ p, size_t n, double value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = value;
}

template <class S, class D>
__global__ void convert_kernel(const S* __restrict__ s, D* __restrict__ d, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = cvt_store<D>(ld_f(s + i));
}
__global__ void f64_to_f32_kernel(const double* __restrict__ s, float* __restrict__ d, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}
__global__ void f32_to_f64_kernel(const float* __restrict__ s, double* __restrict__ d, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = (double)s[i];
}

static unsigned grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b == 0 ? 1 : b) : 8192);
}

// Voxel kernels with a stencil run on a 3D grid: a block is (bz = 2^lbz lanes along z) x
// (256 / bz along x), the grid's z index is y.  bz = min(64, next power of two >= nz), so
// a wave walks one voxel column and no thread divides to find its voxel.
struct VoxGrid {
    dim3 grid;
    int lbz;
};
static VoxGrid vox_grid(int nx, int ny, int nz) {
    int lbz = 0;
    while ((1 << lbz) < nz && lbz < 6) ++lbz;
    const int bz = 1 << lbz, bx = 256 >> lbz;
    return VoxGrid{dim3((unsigned)((nz + bz - 1) / bz), (unsigned)((nx + bx - 1) / bx), (unsigned)ny), lbz};
}
// Blocks in dispatch order (the dispatcher deals them to the 8 XCDs round robin): giving
// each XCD a contiguous run of y planes, so that a stencil's y +-1 planes come from its own
// L2, measured 3% slower on the 512 x 512 x 128 Nesterov update.
__device__ __forceinline__ bool vox_index(int nx, int nz, int lbz, int& ix, int& iy, int& iz, size_t& i) {
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int b = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
    const int bzi = b % gx, r = b / gx;
    iz = (bzi << lbz) + (int)(threadIdx.x & ((1u << lbz) - 1));
    ix = (r % gy) * (256 >> lbz) + (int)(threadIdx.x >> lbz);
    iy = r / gy;
    i = ((size_t)iy * nx + ix) * nz + iz;
    return iz < nz && ix < nx;
}

hipError_t fill_launch(void* p, int dt, size_t n, double value, hipStream_t st) {
    if (n == 0) return hipSuccess;
    switch (dt) {
        case DT_F32: hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, st, (float*)p, n, value); break;
        case DT_F16: hipLaunchKernelGGL(fill_kernel<__half>, dim3(grid_for(n)), dim3(256), 0, st, (__half*)p, n, value); break;
        case DT_BF16: hipLaunchKernelGGL(fill_kernel<__hip_bfloat16>, dim3(grid_for(n)), dim3(256), 0, st, (__hip_bfloat16*)p, n, value); break;
        case DT_F64: hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, (double*)p, n, value); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class S>
static hipError_t convert_from(const void* s, void* d, int ddt, size_t n, hipStream_t st) {
    switch (ddt) {
        case DT_F32: hipLaunchKernelGGL((convert_kernel<S, float>), dim3(grid_for(n)), dim3(256), 0, st, (const S*)s, (float*)d, n); break;
        case DT_F16: hipLaunchKernelGGL((convert_kernel<S, __half>), dim3(grid_for(n)), dim3(256), 0, st, (const S*)s, (__half*)d, n); break;
        case DT_BF16: hipLaunchKernelGGL((convert_kernel<S, __hip_bfloat16>), dim3(grid_for(n)), dim3(256), 0, st, (const S*)s, (__hip_bfloat16*)d, n); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t convert_launch(const void* s, int sdt, void* d, int ddt, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (sdt == DT_F64 && ddt == DT_F32) {
        hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const double*)s, (float*)d, n);
        return hipGetLastError();
    }
    if (sdt == DT_F32 && ddt == DT_F64) {
        hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const float*)s, (double*)d, n);
        return hipGetLastError();
    }
    if (sdt == DT_F64 && ddt == DT_F64) return hipMemcpyAsync(d, s, n * 8, hipMemcpyDeviceToDevice, st);
    switch (sdt) {
        case DT_F32: return convert_from<float>(s, d, ddt, n, st);
        case DT_F16: return convert_from<__half>(s, d, ddt, n, st);
        case DT_BF16: return convert_from<__hip_bfloat16>(s, d, ddt, n, st);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------
// deterministic FP64 reduction tail: block partial -> last block sums in order
// ---------------------------------------------------------------------------------
struct RedScratch {
    double* partial;       // kRedBlocks doubles
    unsigned int* count;   // one counter, returns to 0 after every reduction
};

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kRB / 64];
    // wave reduction
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kRB / 64; ++i) t += sh[i];
    return t;  // valid in thread 0
}

// called by every block; the last one adds `sum(partials)` to *out
__device__ __forceinline__ void finish_reduction(double block_value, RedScratch rs, double* out, int accumulate) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
        rs.partial[blockIdx.x] = block_value;
        __threadfence();
        const unsigned prev = atomicAdd(rs.count, 1u);
        last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double v = 0.0;
    const volatile double* part = rs.partial;  // bypass this CU's L1: written by other blocks
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kRB) v += part[i];
    // fixed-order tree over the strided partial sums
    const double tot = block_sum(v);
    if (threadIdx.x == 0) {
        *out = accumulate ? *out + tot : tot;
        *rs.count = 0u;
    }
}

// ---------------------------------------------------------------------------------
// residual r = w .* (p - y) and phi += sum(w .* (p - y).^2) / 2
// ---------------------------------------------------------------------------------
template <class P, class Y, class W, class R>
__global__ __launch_bounds__(kRB) void residual_kernel(const P* __restrict__ p, const Y* __restrict__ y,
                                                       const W* __restrict__ w, R* __restrict__ r, size_t n,
                                                       double wscale, RedScratch rs, double* __restrict__ phi) {
    // FP32 arithmetic for FP32 / half storage, FP64 when every operand is FP64
    using C = typename std::conditional<std::is_same<P, double>::value && std::is_same<Y, double>::value &&
                                            std::is_same<W, double>::value,
                                        double, float>::type;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (size_t)gridDim.x * kRB) {
        const C d = ld<C>(p + i) - ld<C>(y + i);
        const C wi = ld<C>(w + i);
        r[i] = cvt_store<R>(wi * d);
        acc += (double)wi * (double)d * (double)d;
    }
    const double bs = block_sum(acc);
    finish_reduction(0.5 * wscale * bs, rs, phi, 1);
}

// ---------------------------------------------------------------------------------
// Edge-preserving penalties on a ZXY volume (penalty/huberPenalty.m:101-161,
// quadPenalty.m, hyperbolaPenalty.m).  For voxel i with neighbours k (weights 1 along x
// and y, 1/sqrt(2) on the two xy diagonals, 1/2 along z) and potential psi:
//   S(x)_i = sum_k w_k psi'(x_i - x_k)                (mode 1)
//   T(x)_i = sum_k w_k c(x_i - x_k)                   (mode 2, the curvature bound)
//   R(x)   = sum over pairs w_k psi(x_i - x_k)        (mode 0)
// which is what the reference builds by adding the pair terms into both voxels.
//   Huber:     psi' = clip(d, +-delta), c = [|d| <= delta], psi = d^2/2 or delta|d| - delta^2/2
//   quadratic: psi' = d, psi = d^2/2, and T is the constant 4 + 4/sqrt2 (+ 2 * 1/2 in 3D)
//              of quadPenalty.m for every voxel, boundary voxels included
//   hyperbola: psi' = d / sqrt(1 + (d/delta)^2), c = (1 + (d/delta)^2)^(-3/2),
//              psi = delta^2 (sqrt(1 + (d/delta)^2) - 1)
//   aniso:     c = exp(-d^2/delta^2), psi' = d c, psi = d^2 c / 2 (anisotropicPenalty.m
//              :144-155; its S is d c as written, not the derivative of its R)
//   isohuber:  Huber over the axis neighbours only, every weight 1 (huberIsoPenalty.m:73-104)
// Mode 3 of the Huber penalty is the curvature H of reconPwlsSeNusqs.m
// (huberPenalty.m:163-185): H(x)_i = sum_k w_k [|x_i - x_k| <= delta] (x_i + x_k) / x_i.
//
// Total variation (totalVariationPenalty.m:53-92, totalVariationIsoPenalty.m:53-92) is not a
// pair sum: on the cells k of the grid cropped by one in every dimension (x and y only for
// a 2D volume, nz == 1), c_k = 1 / sqrt(sum_e d_e(k)^2 + 1e-8) with d_e(k) = x_k - x_{k+e}
// over e = x, y (and z for the isotropic form of a 3D volume).  S adds sum_e d_e c at k and
// -d_e c at k + e; T adds c at k once per direction and at every k + e; R = sum_k 1 / c_k.
// ---------------------------------------------------------------------------------
template <int MODE, int POT, class V>
__global__ __launch_bounds__(256) void pen_map_kernel(const V* __restrict__ x, int nx, int ny, int nz, V delta,
                                                      V* __restrict__ out, int lbz) {
    {
        int ix, iy, iz;
        size_t i;
        if (!vox_index(nx, nz, lbz, ix, iy, iz, i)) return;
        out[i] = pen_point<MODE, POT>(x, nx, ny, nz, ix, iy, iz, delta);
    }
}

// penalty cost: each pair counted once (forward neighbours only)
template <int POT, class V>
__global__ __launch_bounds__(kRB) void pen_cost_kernel(const V* __restrict__ x, int nx, int ny, int nz, double delta,
                                                       RedScratch rs, double* __restrict__ out) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t sx = nz, sy = (size_t)nx * nz;
    const double axy = 0.70710678118654752440;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (size_t)gridDim.x * kRB) {
        const int iz = (int)(i % nz);
        const size_t c = i / nz;
        const int ix = (int)(c % nx), iy = (int)(c / nx);
        const double xi = x[i];
        if (POT == POT_TV || POT == POT_ISOTV) {
            const bool three = nz > 1;
            if (ix <= nx - 2 && iy <= ny - 2 && (!three || iz <= nz - 2)) {
                const double d1 = xi - (double)x[i + sx], d2 = xi - (double)x[i + sy];
                const double d3 = (POT == POT_ISOTV && three) ? xi - (double)x[i + 1] : 0.0;
                acc += sqrt(d1 * d1 + d2 * d2 + d3 * d3 + kTvEps);
            }
            continue;
        }
        if (POT == POT_ISOHUBER) {
            if (ix + 1 < nx) acc += cost_pot<POT>(xi - x[i + sx], delta);
            if (iy + 1 < ny) acc += cost_pot<POT>(xi - x[i + sy], delta);
            if (iz + 1 < nz) acc += cost_pot<POT>(xi - x[i + 1], delta);
            continue;
        }
        if (ix + 1 < nx) acc += cost_pot<POT>(xi - x[i + sx], delta);
        if (iy + 1 < ny) acc += cost_pot<POT>(xi - x[i + sy], delta);
        if (ix + 1 < nx && iy + 1 < ny) acc += axy * cost_pot<POT>(xi - x[i + sx + sy], delta);
        // anti-diagonal pair {(ix, iy), (ix-1, iy+1)}, counted at its (ix, iy) member
        if (ix > 0 && iy + 1 < ny) acc += axy * cost_pot<POT>(xi - x[i - sx + sy], delta);
        if (iz + 1 < nz) acc += 0.5 * cost_pot<POT>(xi - x[i + 1], delta);
    }
    finish_reduction(block_sum(acc), rs, out, 0);
}


// fused Nesterov-SQS subset step; reads x, z (stencils) and writes x_out, z_out
// t and t_next come either from the arguments or, when tseq is not null, from
// tseq[*base + k] and tseq[*base + k + 1] (device memory, so a captured graph can be
// replayed across iterations while the momentum sequence advances)
template <class V, class ZC, int POT>
__global__ __launch_bounds__(256) void nesterov_kernel(const V* __restrict__ x, const V* __restrict__ z,
                                                       V* __restrict__ s, const V* __restrict__ g,
                                                       const V* __restrict__ precom, V* __restrict__ x_out,
                                                       V* __restrict__ z_out, ZC* __restrict__ zcopy,
                                                       const V* __restrict__ zref, int nx, int ny,
                                                       int nz, double beta, V delta, double t, double t_next,
                                                       const double* __restrict__ tseq, const int* __restrict__ base,
                                                       int k, int lbz) {
    if (tseq) {
        const int i0 = *base + k;
        t = tseq[i0];
        t_next = tseq[i0 + 1];
    }
    const double a = 1.0 - 1.0 / t_next, b = 1.0 / t_next;
    {
        int ix, iy, iz;
        size_t i;
        if (!vox_index(nx, nz, lbz, ix, iy, iz, i)) return;
        const V sz = pen_point<1, POT>(z, nx, ny, nz, ix, iy, iz, delta);
        const V tx = pen_point<2, POT>(x, nx, ny, nz, ix, iy, iz, delta);
        const V grad = g[i] + (V)(beta * (double)sz);
        const double den = (double)precom[i] + beta * (double)tx;
        const V du = (V)((double)grad / den);
        const V xn = z[i] - du;
        const V sn = s[i] - (V)(t * (double)du);
        const V zn = (V)(a * (double)xn + b * (double)sn);
        x_out[i] = scrub(xn);
        s[i] = scrub(sn);
        const V zs = scrub(zn);
        z_out[i] = zs;
        if (zcopy) zcopy[i] = cvt_store<ZC>(zref ? zs - zref[i] : zs);
    }
}

// Four consecutive z voxels per thread (nz % 4 == 0, the 13-neighbour penalties): the
// eight in-plane neighbour columns are float4 loads and the z neighbours come from the
// segment itself plus one scalar below and above it, so a wave issues about a fifth of the
// memory instructions of the one-voxel form.  Terms are added in pen_point's order, so the
// results are bit-identical to nesterov_kernel.
struct Seg4 {
    float4 v[9];  // centre, x+1, x-1, y+1, y-1, (+1,+1), (-1,-1), (-1,+1), (+1,-1)
    unsigned in;  // bit k: neighbour column k exists
    float zlo, zhi;
};

__device__ __forceinline__ float f4(const float4& a, int q) { return q == 0 ? a.x : (q == 1 ? a.y : (q == 2 ? a.z : a.w)); }

__device__ __forceinline__ void load_seg4(const float* __restrict__ a, int nx, int ny, int nz, int ix, int iy, int iz0,
                                          size_t i, Seg4& g) {
    const long long sx = nz, sy = (long long)nx * nz;
    const bool xp = ix + 1 < nx, xm = ix > 0, yp = iy + 1 < ny, ym = iy > 0;
    const long long off[9] = {0, sx, -sx, sy, -sy, sx + sy, -sx - sy, -sx + sy, sx - sy};
    const bool ok[9] = {true, xp, xm, yp, ym, xp && yp, xm && ym, xm && yp, xp && ym};
    g.in = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        g.v[k] = ok[k] ? *reinterpret_cast<const float4*>(a + (long long)i + off[k]) : make_float4(0.f, 0.f, 0.f, 0.f);
        g.in |= ok[k] ? (1u << k) : 0u;
    }
    g.zlo = iz0 > 0 ? a[i - 1] : 0.f;
    g.zhi = iz0 + 4 < nz ? a[i + 4] : 0.f;
}

template <int MODE, int POT>
__device__ __forceinline__ float pen_seg4(const Seg4& g, int q, int iz, int nz, float delta) {
    const float axy = 0.70710678118654752440f;
    if (MODE == 2 && POT == POT_QUAD) return nz > 1 ? 4.f + 4.f * axy + 1.f : 4.f + 4.f * axy;
    const float xi = f4(g.v[0], q);
    float acc = 0.f;
    auto term = [&](bool inside, float xj, float w) {
        if (!inside) return;
        const float d = xi - xj;
        acc += MODE == 1 ? w * dpot<POT>(d, delta) : w * cpot<POT>(d, delta);
    };
    const float wk[9] = {0.f, 1.f, 1.f, 1.f, 1.f, axy, axy, axy, axy};
#pragma unroll
    for (int k = 1; k < 9; ++k) term((g.in >> k) & 1u, f4(g.v[k], q), wk[k]);
    term(iz + 1 < nz, q == 3 ? g.zhi : f4(g.v[0], q + 1), 0.5f);
    term(iz > 0, q == 0 ? g.zlo : f4(g.v[0], q - 1), 0.5f);
    return acc;
}

template <class ZC>
__device__ __forceinline__ void store4(ZC* p, const float v[4]) {
    if constexpr (sizeof(ZC) == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        ZC h[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) h[q] = cvt_store<ZC>(v[q]);
        uint2 u;
        __builtin_memcpy(&u, h, 8);
        *reinterpret_cast<uint2*>(p) = u;
    }
}

template <class ZC, int POT>
__global__ __launch_bounds__(256) void nesterov4_kernel(const float* __restrict__ x, const float* __restrict__ z,
                                                        float* __restrict__ s, const float* __restrict__ g,
                                                        const float* __restrict__ precom, float* __restrict__ x_out,
                                                        float* __restrict__ z_out, ZC* __restrict__ zcopy,
                                                        const float* __restrict__ zref, int nx,
                                                        int ny, int nz, double beta, float delta, double t,
                                                        double t_next, const double* __restrict__ tseq,
                                                        const int* __restrict__ base, int k, int lbz) {
    if (tseq) {
        const int i0 = *base + k;
        t = tseq[i0];
        t_next = tseq[i0 + 1];
    }
    const double a = 1.0 - 1.0 / t_next, b = 1.0 / t_next;
    int ix, iy, iq;
    size_t i4;
    if (!vox_index(nx, nz >> 2, lbz, ix, iy, iq, i4)) return;  // iq: segment index along z
    const int iz0 = iq * 4;
    const size_t i = ((size_t)iy * nx + ix) * nz + iz0;
    Seg4 sz, sx;
    load_seg4(z, nx, ny, nz, ix, iy, iz0, i, sz);
    load_seg4(x, nx, ny, nz, ix, iy, iz0, i, sx);
    const float4 s4 = *reinterpret_cast<const float4*>(s + i);
    const float4 g4 = *reinterpret_cast<const float4*>(g + i);
    const float4 p4 = *reinterpret_cast<const float4*>(precom + i);
    float xo[4], so[4], zo[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float szv = pen_seg4<1, POT>(sz, q, iz0 + q, nz, delta);
        const float txv = pen_seg4<2, POT>(sx, q, iz0 + q, nz, delta);
        const float grad = f4(g4, q) + (float)(beta * (double)szv);
        const double den = (double)f4(p4, q) + beta * (double)txv;
        const float du = (float)((double)grad / den);
        const float xn = f4(sz.v[0], q) - du;
        const float sn = f4(s4, q) - (float)(t * (double)du);
        const float zn = (float)(a * (double)xn + b * (double)sn);
        xo[q] = scrub(xn);
        so[q] = scrub(sn);
        zo[q] = scrub(zn);
    }
    store4(x_out + i, xo);
    store4(s + i, so);
    store4(z_out + i, zo);
    if (zcopy) {
        if (zref) {
            const float4 r4 = *reinterpret_cast<const float4*>(zref + i);
            const float d[4] = {zo[0] - r4.x, zo[1] - r4.y, zo[2] - r4.z, zo[3] - r4.w};
            store4(zcopy + i, d);
        } else {
            store4(zcopy + i, zo);
        }
    }
}

template <int POT>
__global__ __launch_bounds__(256) void sqs_kernel(const float* __restrict__ x, float* __restrict__ x_out,
                                                  const float* __restrict__ g, const float* __restrict__ precom, int nx,
                                                  int ny, int nz, double beta, float delta, int lbz) {
    {
        int ix, iy, iz;
        size_t i;
        if (!vox_index(nx, nz, lbz, ix, iy, iz, i)) return;
        const float sx = pen_point<1, POT>(x, nx, ny, nz, ix, iy, iz, delta);
        const float tx = pen_point<2, POT>(x, nx, ny, nz, ix, iy, iz, delta);
        const float num = g[i] + (float)(beta * (double)sx);
        const double den = (double)precom[i] + beta * (double)tx;
        x_out[i] = scrub((float)((double)x[i] - (double)num / den));
    }
}

// LALM-OS (statsrecon/pwls/reconPwlsLALMOs.m:66-90): after the subset gradient dL,
//   g = rho/(rho+1) dL + 1/(rho+1) g;   s = rho dL + (1 - rho) g   (NaN/Inf in s -> 0)
__global__ __launch_bounds__(256) void lalm_sg_kernel(const float* __restrict__ dl, float* __restrict__ g,
                                                      float* __restrict__ s, size_t n, double rho) {
    const float a = (float)(rho / (rho + 1.0)), b = (float)(1.0 / (rho + 1.0)), r = (float)rho,
                c = (float)(1.0 - rho);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = a * dl[i] + b * g[i];
        g[i] = gi;
        const float si = r * dl[i] + c * gi;
        s[i] = isfinite(si) ? si : 0.f;
    }
}

// one inner step: x' = x - (beta S(x) + rho precom (x - xold) + s) / (beta T(x) + rho precom),
// clamped to >= 0 with NaN/Inf -> 0 (the denominator in FP64)
template <int POT>
__global__ __launch_bounds__(256) void lalm_inner_kernel(const float* __restrict__ x, const float* __restrict__ xold,
                                                         const float* __restrict__ s,
                                                         const float* __restrict__ precom, float* __restrict__ x_out,
                                                         int nx, int ny, int nz, double beta, float delta, double rho, int lbz) {
    {
        int ix, iy, iz;
        size_t i;
        if (!vox_index(nx, nz, lbz, ix, iy, iz, i)) return;
        const float sx = pen_point<1, POT>(x, nx, ny, nz, ix, iy, iz, delta);
        const float tx = pen_point<2, POT>(x, nx, ny, nz, ix, iy, iz, delta);
        const double num = beta * (double)sx + rho * (double)precom[i] * ((double)x[i] - (double)xold[i]) + (double)s[i];
        const double den = beta * (double)tx + rho * (double)precom[i];
        x_out[i] = scrub((float)((double)x[i] - num / den));
    }
}

// ---------------------------------------------------------------------------------
// NU-SQS (statsrecon/pwls/reconPwlsSeNusqs.m:47-90)
// ---------------------------------------------------------------------------------
// v = 2 (|G1 * x| + |G2 * x|) + |x|, at least eps (:47-51).  G1 = [1 2 1; 0 0 0; -1 -2 1] and
// G2 = [1 0 -1; 2 0 -2; 1 0 -1] correlate each z slice (filter/imageFilter2D.m: imfilter,
// 'replicate' edges); MATLAB's rows are x, its columns y.
__global__ __launch_bounds__(256) void nusqs_seed_kernel(const float* __restrict__ x, float* __restrict__ v, int nx,
                                                         int ny, int nz, float eps, int lbz) {
    {
        int ix, iy, iz;
        size_t i;
        if (!vox_index(nx, nz, lbz, ix, iy, iz, i)) return;
        auto at = [&](int a, int b) {
            a = a < 0 ? 0 : (a >= nx ? nx - 1 : a);
            b = b < 0 ? 0 : (b >= ny ? ny - 1 : b);
            return x[((size_t)b * nx + a) * nz + iz];
        };
        const float g1 = (at(ix - 1, iy - 1) + 2.f * at(ix - 1, iy) + at(ix - 1, iy + 1)) -
                         (at(ix + 1, iy - 1) + 2.f * at(ix + 1, iy) + at(ix + 1, iy + 1));
        const float g2 = (at(ix - 1, iy - 1) + 2.f * at(ix, iy - 1) + at(ix + 1, iy - 1)) -
                         (at(ix - 1, iy + 1) + 2.f * at(ix, iy + 1) + at(ix + 1, iy + 1));
        const float val = (fabsf(g1) + fabsf(g2)) * 2.f + fabsf(x[i]);
        v[i] = val < eps ? eps : val;
    }
}

// out = a / b
__global__ void div_kernel(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ o, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = a[i] / b[i];
}

// v = |x - xref|, at least eps (:72-73)
__global__ void absdiff_floor_kernel(const float* __restrict__ x, const float* __restrict__ xref,
                                     float* __restrict__ o, size_t n, float eps) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float d = fabsf(x[i] - xref[i]);
        o[i] = d < eps ? eps : d;
    }
}

// x' = x - (g + beta S(x)) / (dL + beta H), NaN/Inf and negatives -> 0 (:81-90); the
// denominator in FP64
template <int POT>
__global__ __launch_bounds__(256) void nusqs_kernel(const float* __restrict__ x, float* __restrict__ x_out,
                                                    const float* __restrict__ g, const float* __restrict__ dl,
                                                    const float* __restrict__ h, int nx, int ny, int nz, double beta,
                                                    float delta, int lbz) {
    {
        int ix, iy, iz;
        size_t i;
        if (!vox_index(nx, nz, lbz, ix, iy, iz, i)) return;
        const float sx = pen_point<1, POT>(x, nx, ny, nz, ix, iy, iz, delta);
        const float num = g[i] + (float)(beta * (double)sx);
        const double den = (double)dl[i] + beta * (double)h[i];
        x_out[i] = scrub((float)((double)x[i] - (double)num / den));
    }
}

// dynamicRangeAdjustment (statsrecon/pwls/dynamicRangeAdjustment.m): z = max(F(x)^t, eta)
// with F the piecewise-linear empirical CDF through min, (min+mean)/2, mean, (mean+max)/2,
// max.  Three passes: min / max / sum, the three counts, the map.  The input is
// non-negative (|x - xref| floored at eps), so the order of the float bit patterns is the
// order of the values and min / max are exact integer atomics; the sum is the
// deterministic FP64 reduction and the counts are exact.
struct DraStats {
    unsigned int minbits, maxbits;
    unsigned long long cnt[3];
    double sum;
};

__global__ void dra_init_kernel(DraStats* s) {
    s->minbits = 0x7f800000u;  // +inf
    s->maxbits = 0u;
    s->cnt[0] = s->cnt[1] = s->cnt[2] = 0ull;
    s->sum = 0.0;
}

__global__ __launch_bounds__(kRB) void dra_stats_kernel(const float* __restrict__ x, size_t n, DraStats* s,
                                                        RedScratch rs) {
    unsigned int mn = 0x7f800000u, mx = 0u;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (size_t)gridDim.x * kRB) {
        const float xi = x[i];
        const unsigned int b = __float_as_uint(xi);
        mn = b < mn ? b : mn;
        mx = b > mx ? b : mx;
        acc += (double)xi;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned int a = __shfl_down(mn, off, 64), c = __shfl_down(mx, off, 64);
        mn = a < mn ? a : mn;
        mx = c > mx ? c : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&s->minbits, mn);
        atomicMax(&s->maxbits, mx);
    }
    finish_reduction(block_sum(acc), rs, &s->sum, 0);
}

__device__ __forceinline__ void dra_thresholds(const DraStats* s, size_t n, float& a, float& t1, float& t2, float& t3,
                                               float& c) {
    a = __uint_as_float(s->minbits);
    c = __uint_as_float(s->maxbits);
    const float b = (float)(s->sum / (double)n);
    t1 = (a + b) / 2.f;
    t2 = b;
    t3 = (b + c) / 2.f;
}

__global__ __launch_bounds__(kRB) void dra_count_kernel(const float* __restrict__ x, size_t n, DraStats* s) {
    float a, t1, t2, t3, c;
    dra_thresholds(s, n, a, t1, t2, t3, c);
    unsigned long long k1 = 0, k2 = 0, k3 = 0;
    for (size_t i = (size_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (size_t)gridDim.x * kRB) {
        const float xi = x[i];
        k1 += xi <= t1;
        k2 += xi <= t2;
        k3 += xi <= t3;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        k1 += __shfl_down(k1, off, 64);
        k2 += __shfl_down(k2, off, 64);
        k3 += __shfl_down(k3, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&s->cnt[0], k1);
        atomicAdd(&s->cnt[1], k2);
        atomicAdd(&s->cnt[2], k3);
    }
}

__global__ void dra_map_kernel(const float* __restrict__ x, size_t n, const DraStats* s, float t, float eta,
                               float* __restrict__ out) {
    float a, t1, t2, t3, c;
    dra_thresholds(s, n, a, t1, t2, t3, c);
    const float p1 = (float)((double)s->cnt[0] / (double)n), p2 = (float)((double)s->cnt[1] / (double)n),
                p3 = (float)((double)s->cnt[2] / (double)n);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float xi = x[i];
        float y;
        if (xi <= t1) y = p1 * (xi - a) / (t1 - a);
        else if (xi <= t2) y = (p2 - p1) * (xi - t1) / (t2 - t1) + p1;
        else if (xi <= t3) y = (p3 - p2) * (xi - t2) / (t3 - t2) + p2;
        else y = (1.f - p3) * (xi - t3) / (c - t3) + p3;
        const float z = powf(y, t);
        out[i] = z > eta ? z : eta;  // MATLAB max: a NaN loses to eta
    }
}

template <class V>
__global__ __launch_bounds__(kRB) void sqdiff_kernel(const V* __restrict__ a, const V* __restrict__ b, size_t n,
                                                     RedScratch rs, double* __restrict__ out) {
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (size_t)gridDim.x * kRB) {
        const double d = (double)a[i] - (double)b[i];
        acc += d * d;
    }
    finish_reduction(block_sum(acc), rs, out, 0);
}

template <class A, class B, class O>
__global__ void mul_kernel(const A* __restrict__ a, const B* __restrict__ b, O* __restrict__ o, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = cvt_store<O>(ld_f(a + i) * ld_f(b + i));
}

__global__ void mul_f64_kernel(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ o,
                               size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = a[i] * b[i];
}

// ---------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------
static RedScratch g_scratch[16];

hipError_t reduction_scratch_init();

static hipError_t lib_scratch(RedScratch& rs) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 16) return hipErrorInvalidDevice;
    if (!g_scratch[dev].partial) {
        // allocated once per device and kept until the process ends; the counter
        // starts at zero and every reduction returns it to zero
        void* p = nullptr;
        e = hipMalloc(&p, kRedBlocks * sizeof(double) + 256);
        if (e != hipSuccess) return e;
        e = hipMemset(p, 0, kRedBlocks * sizeof(double) + 256);
        if (e != hipSuccess) return e;
        g_scratch[dev].partial = (double*)p;
        g_scratch[dev].count = (unsigned int*)((char*)p + kRedBlocks * sizeof(double));
    }
    rs = g_scratch[dev];
    return hipSuccess;
}

hipError_t reduction_scratch_init() {
    RedScratch rs;
    return lib_scratch(rs);
}

static unsigned red_grid(size_t n) {
    size_t b = (n + kRB - 1) / kRB;
    return (unsigned)(b < (size_t)kRedBlocks ? (b == 0 ? 1 : b) : kRedBlocks);
}

template <class P, class Y, class W>
static hipError_t residual_r(const void* p, const void* y, const void* w, void* r, int rdt, size_t n, double ws,
                             RedScratch rs, double* phi, hipStream_t st) {
    const unsigned gb = red_grid(n);
    switch (rdt) {
        case DT_F32: hipLaunchKernelGGL((residual_kernel<P, Y, W, float>), dim3(gb), dim3(kRB), 0, st, (const P*)p, (const Y*)y, (const W*)w, (float*)r, n, ws, rs, phi); break;
        case DT_F16: hipLaunchKernelGGL((residual_kernel<P, Y, W, __half>), dim3(gb), dim3(kRB), 0, st, (const P*)p, (const Y*)y, (const W*)w, (__half*)r, n, ws, rs, phi); break;
        case DT_BF16: hipLaunchKernelGGL((residual_kernel<P, Y, W, __hip_bfloat16>), dim3(gb), dim3(kRB), 0, st, (const P*)p, (const Y*)y, (const W*)w, (__hip_bfloat16*)r, n, ws, rs, phi); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class P, class Y>
static hipError_t residual_w(const void* p, const void* y, const void* w, int wdt, void* r, int rdt, size_t n,
                             double ws, RedScratch rs, double* phi, hipStream_t st) {
    switch (wdt) {
        case DT_F32: return residual_r<P, Y, float>(p, y, w, r, rdt, n, ws, rs, phi, st);
        case DT_F16: return residual_r<P, Y, __half>(p, y, w, r, rdt, n, ws, rs, phi, st);
        case DT_BF16: return residual_r<P, Y, __hip_bfloat16>(p, y, w, r, rdt, n, ws, rs, phi, st);
        default: return hipErrorInvalidValue;
    }
}

template <class P>
static hipError_t residual_y(const void* p, const void* y, int ydt, const void* w, int wdt, void* r, int rdt, size_t n,
                             double ws, RedScratch rs, double* phi, hipStream_t st) {
    switch (ydt) {
        case DT_F32: return residual_w<P, float>(p, y, w, wdt, r, rdt, n, ws, rs, phi, st);
        case DT_F16: return residual_w<P, __half>(p, y, w, wdt, r, rdt, n, ws, rs, phi, st);
        case DT_BF16: return residual_w<P, __hip_bfloat16>(p, y, w, wdt, r, rdt, n, ws, rs, phi, st);
        default: return hipErrorInvalidValue;
    }
}

static hipError_t residual_launch_rs(const void* p, int pdt, const void* y, int ydt, const void* w, int wdt, void* r,
                                     int rdt, size_t n, double ws, double* phi, RedScratch rs, hipStream_t st) {
    if (pdt == DT_F64 || ydt == DT_F64 || wdt == DT_F64 || rdt == DT_F64) {
        if (pdt != DT_F64 || ydt != DT_F64 || wdt != DT_F64 || rdt != DT_F64) return hipErrorInvalidValue;
        hipLaunchKernelGGL((residual_kernel<double, double, double, double>), dim3(red_grid(n)), dim3(kRB), 0, st,
                           (const double*)p, (const double*)y, (const double*)w, (double*)r, n, ws, rs, phi);
        return hipGetLastError();
    }
    switch (pdt) {
        case DT_F32: return residual_y<float>(p, y, ydt, w, wdt, r, rdt, n, ws, rs, phi, st);
        case DT_F16: return residual_y<__half>(p, y, ydt, w, wdt, r, rdt, n, ws, rs, phi, st);
        case DT_BF16: return residual_y<__hip_bfloat16>(p, y, ydt, w, wdt, r, rdt, n, ws, rs, phi, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t residual_launch(const void* p, int pdt, const void* y, int ydt, const void* w, int wdt, void* r, int rdt,
                           size_t n, double ws, double* phi, hipStream_t st) {
    RedScratch rs;
    hipError_t e = lib_scratch(rs);
    if (e != hipSuccess) return e;
    return residual_launch_rs(p, pdt, y, ydt, w, wdt, r, rdt, n, ws, phi, rs, st);
}

// ---------------------------------------------------------------------------------
// Allocation-free reduction scratch: the same layout as the library-allocated one
// (kRedBlocks FP64 partial sums, then the counter) in a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes the array's name
// to a kernel: the first call on a device looks its address up with
// hipGetSymbolAddress and keeps it, and every launch takes that address.
// ---------------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = kRedBlocks * sizeof(double) + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static void* g_noalloc_ptr[16];
static std::mutex g_noalloc_mutex;

__global__ void zero_words_kernel(unsigned int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

static hipError_t noalloc_scratch(RedScratch& rs, hipStream_t st) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 16) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        // the counter must start at zero (every reduction returns it to zero); cleared by a
        // kernel on the caller's stream, ahead of this call's own reduction
        hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, st,
                           (unsigned int*)((char*)ptr + kRedBlocks * sizeof(double)), 64);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        g_noalloc_ptr[dev] = ptr;
    }
    rs.partial = (double*)g_noalloc_ptr[dev];
    rs.count = (unsigned int*)((char*)g_noalloc_ptr[dev] + kRedBlocks * sizeof(double));
    return hipSuccess;
}

hipError_t residual_launch_noalloc(const void* p, int pdt, const void* y, int ydt, const void* w, int wdt, void* r,
                                   int rdt, size_t n, double ws, double* phi, hipStream_t st) {
    // the reduction grid is at most kRedBlocks blocks whatever n is, so the array always holds
    // its partial sums
    if (red_grid(n) > (unsigned)kRedBlocks) return hipErrorInvalidValue;
    RedScratch rs;
    hipError_t e = noalloc_scratch(rs, st);
    if (e != hipSuccess) return e;
    return residual_launch_rs(p, pdt, y, ydt, w, wdt, r, rdt, n, ws, phi, rs, st);
}

template <int POT, class V>
static hipError_t pen_launch_p(const V* x, int nx, int ny, int nz, double delta, int mode, V* out, double* cost,
                               hipStream_t st) {
    const size_t n = (size_t)nx * ny * nz;
    const VoxGrid vg = vox_grid(nx, ny, nz);
    if (mode == 1) {
        hipLaunchKernelGGL((pen_map_kernel<1, POT, V>), vg.grid, dim3(256), 0, st, x, nx, ny, nz, (V)delta, out,
                           vg.lbz);
    } else if (mode == 2) {
        hipLaunchKernelGGL((pen_map_kernel<2, POT, V>), vg.grid, dim3(256), 0, st, x, nx, ny, nz, (V)delta, out,
                           vg.lbz);
    } else if (mode == 3) {
        if constexpr (POT == POT_HUBER) {
            hipLaunchKernelGGL((pen_map_kernel<3, POT, V>), vg.grid, dim3(256), 0, st, x, nx, ny, nz, (V)delta, out,
                               vg.lbz);
        } else {
            return hipErrorInvalidValue;  // mode 3 is defined for the Huber penalty only
        }
    } else if (mode == 0) {
        RedScratch rs;
        hipError_t e = lib_scratch(rs);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((pen_cost_kernel<POT, V>), dim3(red_grid(n)), dim3(kRB), 0, st, x, nx, ny, nz, delta, rs,
                           cost);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// pfun: one of Pot
#define CT_POT_DISPATCH(pfun, CALL)                     \
    switch (pfun) {                                     \
        case POT_HUBER: return CALL(POT_HUBER);          \
        case POT_QUAD: return CALL(POT_QUAD);            \
        case POT_HYPER: return CALL(POT_HYPER);          \
        case POT_ANISO: return CALL(POT_ANISO);          \
        case POT_ISOHUBER: return CALL(POT_ISOHUBER);    \
        case POT_TV: return CALL(POT_TV);                \
        case POT_ISOTV: return CALL(POT_ISOTV);          \
        default: return hipErrorInvalidValue;           \
    }

hipError_t penalty_launch(const float* x, int nx, int ny, int nz, int pfun, double delta, int mode, float* out,
                          double* cost, hipStream_t st) {
#define CALL(P) pen_launch_p<P, float>(x, nx, ny, nz, delta, mode, out, cost, st)
    CT_POT_DISPATCH(pfun, CALL)
#undef CALL
}

hipError_t penalty_launch(const double* x, int nx, int ny, int nz, int pfun, double delta, int mode, double* out,
                          double* cost, hipStream_t st) {
#define CALL(P) pen_launch_p<P, double>(x, nx, ny, nz, delta, mode, out, cost, st)
    CT_POT_DISPATCH(pfun, CALL)
#undef CALL
}

template <int POT>
static hipError_t nesterov_p(const float* x, const float* z, float* s, const float* g, const float* precom,
                             float* x_out, float* z_out, void* zcopy, int zcdt, const float* zref, int nx, int ny,
                             int nz, double beta,
                             double delta, double t, double t_next, const double* tseq, const int* base, int k,
                             hipStream_t st) {
    if (nz % 4 == 0 && (POT == POT_HUBER || POT == POT_QUAD || POT == POT_HYPER || POT == POT_ANISO)) {
        const VoxGrid vg = vox_grid(nx, ny, nz / 4);
        const dim3 gr = vg.grid, bl(256);
        const int lbz = vg.lbz;
        switch (zcopy ? zcdt : DT_F32) {
            case DT_F32: hipLaunchKernelGGL((nesterov4_kernel<float, POT>), gr, bl, 0, st, x, z, s, g, precom, x_out, z_out, (float*)zcopy, zref, nx, ny, nz, beta, (float)delta, t, t_next, tseq, base, k, lbz); break;
            case DT_F16: hipLaunchKernelGGL((nesterov4_kernel<__half, POT>), gr, bl, 0, st, x, z, s, g, precom, x_out, z_out, (__half*)zcopy, zref, nx, ny, nz, beta, (float)delta, t, t_next, tseq, base, k, lbz); break;
            case DT_BF16: hipLaunchKernelGGL((nesterov4_kernel<__hip_bfloat16, POT>), gr, bl, 0, st, x, z, s, g, precom, x_out, z_out, (__hip_bfloat16*)zcopy, zref, nx, ny, nz, beta, (float)delta, t, t_next, tseq, base, k, lbz); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const VoxGrid vg = vox_grid(nx, ny, nz);
    const dim3 gr = vg.grid, bl(256);
    const int lbz = vg.lbz;
    switch (zcopy ? zcdt : DT_F32) {
        case DT_F32: hipLaunchKernelGGL((nesterov_kernel<float, float, POT>), gr, bl, 0, st, x, z, s, g, precom, x_out, z_out, (float*)zcopy, zref, nx, ny, nz, beta, (float)delta, t, t_next, tseq, base, k, lbz); break;
        case DT_F16: hipLaunchKernelGGL((nesterov_kernel<float, __half, POT>), gr, bl, 0, st, x, z, s, g, precom, x_out, z_out, (__half*)zcopy, zref, nx, ny, nz, beta, (float)delta, t, t_next, tseq, base, k, lbz); break;
        case DT_BF16: hipLaunchKernelGGL((nesterov_kernel<float, __hip_bfloat16, POT>), gr, bl, 0, st, x, z, s, g, precom, x_out, z_out, (__hip_bfloat16*)zcopy, zref, nx, ny, nz, beta, (float)delta, t, t_next, tseq, base, k, lbz); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t nesterov_launch(const float* x, const float* z, float* s, const float* g, const float* precom, float* x_out,
                           float* z_out, void* zcopy, int zcdt, const float* zref, int nx, int ny, int nz, double beta,
                           double delta,
                           int pfun, double t, double t_next, const double* tseq, const int* base, int k,
                           hipStream_t st) {
#define CALL(P) nesterov_p<P>(x, z, s, g, precom, x_out, z_out, zcopy, zcdt, zref, nx, ny, nz, beta, delta, t, t_next, tseq, base, k, st)
    CT_POT_DISPATCH(pfun, CALL)
#undef CALL
}

template <int POT>
static hipError_t nesterov64_p(const double* x, const double* z, double* s, const double* g, const double* precom,
                               double* x_out, double* z_out, int nx, int ny, int nz, double beta, double delta,
                               double t, double t_next, hipStream_t st) {
    const VoxGrid vg = vox_grid(nx, ny, nz);
    hipLaunchKernelGGL((nesterov_kernel<double, double, POT>), vg.grid, dim3(256), 0, st, x, z, s, g, precom, x_out,
                       z_out, (double*)nullptr, (const double*)nullptr, nx, ny, nz, beta, delta, t, t_next,
                       (const double*)nullptr,
                       (const int*)nullptr, 0, vg.lbz);
    return hipGetLastError();
}

hipError_t nesterov_launch(const double* x, const double* z, double* s, const double* g, const double* precom,
                           double* x_out, double* z_out, int nx, int ny, int nz, double beta, double delta, int pfun,
                           double t, double t_next, hipStream_t st) {
#define CALL(P) nesterov64_p<P>(x, z, s, g, precom, x_out, z_out, nx, ny, nz, beta, delta, t, t_next, st)
    CT_POT_DISPATCH(pfun, CALL)
#undef CALL
}

// ---- NU-SQS launchers ---------------------------------------------------------------
hipError_t nusqs_seed_launch(const float* x, float* v, int nx, int ny, int nz, double eps, hipStream_t st) {
    const VoxGrid vg = vox_grid(nx, ny, nz);
    hipLaunchKernelGGL(nusqs_seed_kernel, vg.grid, dim3(256), 0, st, x, v, nx, ny, nz, (float)eps, vg.lbz);
    return hipGetLastError();
}

hipError_t div_launch(const float* a, const float* b, float* o, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(div_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, b, o, n);
    return hipGetLastError();
}

hipError_t absdiff_floor_launch(const float* x, const float* xref, float* o, size_t n, double eps, hipStream_t st) {
    hipLaunchKernelGGL(absdiff_floor_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, xref, o, n, (float)eps);
    return hipGetLastError();
}

template <int POT>
static hipError_t nusqs_p(const float* x, float* x_out, const float* g, const float* dl, const float* h, int nx,
                          int ny, int nz, double beta, double delta, hipStream_t st) {
    const VoxGrid vg = vox_grid(nx, ny, nz);
    hipLaunchKernelGGL((nusqs_kernel<POT>), vg.grid, dim3(256), 0, st, x, x_out, g, dl, h, nx, ny, nz, beta,
                       (float)delta, vg.lbz);
    return hipGetLastError();
}

hipError_t nusqs_launch(const float* x, float* x_out, const float* g, const float* dl, const float* h, int nx, int ny,
                        int nz, double beta, double delta, int pfun, hipStream_t st) {
#define CALL(P) nusqs_p<P>(x, x_out, g, dl, h, nx, ny, nz, beta, delta, st)
    CT_POT_DISPATCH(pfun, CALL)
#undef CALL
}

static DraStats* g_dra[16];

hipError_t dynamic_range_launch(const float* x, size_t n, double t, double eta, float* out, hipStream_t st) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 16) return hipErrorInvalidDevice;
    if (!g_dra[dev]) {  // allocated once per device, kept until the process ends
        void* p = nullptr;
        e = hipMalloc(&p, sizeof(DraStats));
        if (e != hipSuccess) return e;
        g_dra[dev] = (DraStats*)p;
    }
    RedScratch rs;
    e = lib_scratch(rs);
    if (e != hipSuccess) return e;
    DraStats* s = g_dra[dev];
    hipLaunchKernelGGL(dra_init_kernel, dim3(1), dim3(1), 0, st, s);
    hipLaunchKernelGGL(dra_stats_kernel, dim3(red_grid(n)), dim3(kRB), 0, st, x, n, s, rs);
    hipLaunchKernelGGL(dra_count_kernel, dim3(red_grid(n)), dim3(kRB), 0, st, x, n, s);
    hipLaunchKernelGGL(dra_map_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, n, (const DraStats*)s, (float)t,
                       (float)eta, out);
    return hipGetLastError();
}

__global__ void add_i32_kernel(int* p, int inc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += inc;
}

hipError_t add_i32_launch(int* p, int inc, hipStream_t st) {
    hipLaunchKernelGGL(add_i32_kernel, dim3(1), dim3(64), 0, st, p, inc);
    return hipGetLastError();
}

template <int POT>
static hipError_t sqs_p(const float* x, float* x_out, const float* g, const float* precom, int nx, int ny, int nz,
                        double beta, double delta, hipStream_t st) {
    const VoxGrid vg = vox_grid(nx, ny, nz);
    hipLaunchKernelGGL(sqs_kernel<POT>, vg.grid, dim3(256), 0, st, x, x_out, g, precom, nx, ny, nz, beta, (float)delta,
                       vg.lbz);
    return hipGetLastError();
}

hipError_t sqs_launch(const float* x, float* x_out, const float* g, const float* precom, int nx, int ny, int nz,
                      double beta, double delta, int pfun, hipStream_t st) {
#define CALL(P) sqs_p<P>(x, x_out, g, precom, nx, ny, nz, beta, delta, st)
    CT_POT_DISPATCH(pfun, CALL)
#undef CALL
}

hipError_t lalm_sg_launch(const float* dl, float* g, float* s, size_t n, double rho, hipStream_t st) {
    hipLaunchKernelGGL(lalm_sg_kernel, dim3(grid_for(n)), dim3(256), 0, st, dl, g, s, n, rho);
    return hipGetLastError();
}

template <int POT>
static hipError_t lalm_inner_p(const float* x, const float* xold, const float* s, const float* precom, float* x_out,
                               int nx, int ny, int nz, double beta, double delta, double rho, hipStream_t st) {
    const VoxGrid vg = vox_grid(nx, ny, nz);
    hipLaunchKernelGGL(lalm_inner_kernel<POT>, vg.grid, dim3(256), 0, st, x, xold, s, precom, x_out, nx, ny, nz, beta,
                       (float)delta, rho, vg.lbz);
    return hipGetLastError();
}

hipError_t lalm_inner_launch(const float* x, const float* xold, const float* s, const float* precom, float* x_out,
                             int nx, int ny, int nz, double beta, double delta, int pfun, double rho, hipStream_t st) {
#define CALL(P) lalm_inner_p<P>(x, xold, s, precom, x_out, nx, ny, nz, beta, delta, rho, st)
    CT_POT_DISPATCH(pfun, CALL)
#undef CALL
}

template <class V>
static hipError_t sqdiff_launch_t(const V* a, const V* b, size_t n, double* out, hipStream_t st) {
    RedScratch rs;
    hipError_t e = lib_scratch(rs);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sqdiff_kernel<V>, dim3(red_grid(n)), dim3(kRB), 0, st, a, b, n, rs, out);
    return hipGetLastError();
}

hipError_t sqdiff_launch(const float* a, const float* b, size_t n, double* out, hipStream_t st) {
    return sqdiff_launch_t<float>(a, b, n, out, st);
}

hipError_t sqdiff_launch(const double* a, const double* b, size_t n, double* out, hipStream_t st) {
    return sqdiff_launch_t<double>(a, b, n, out, st);
}

template <class A, class B>
static hipError_t mul_o(const void* a, const void* b, void* o, int odt, size_t n, hipStream_t st) {
    switch (odt) {
        case DT_F32: hipLaunchKernelGGL((mul_kernel<A, B, float>), dim3(grid_for(n)), dim3(256), 0, st, (const A*)a, (const B*)b, (float*)o, n); break;
        case DT_F16: hipLaunchKernelGGL((mul_kernel<A, B, __half>), dim3(grid_for(n)), dim3(256), 0, st, (const A*)a, (const B*)b, (__half*)o, n); break;
        case DT_BF16: hipLaunchKernelGGL((mul_kernel<A, B, __hip_bfloat16>), dim3(grid_for(n)), dim3(256), 0, st, (const A*)a, (const B*)b, (__hip_bfloat16*)o, n); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class A>
static hipError_t mul_b(const void* a, const void* b, int bdt, void* o, int odt, size_t n, hipStream_t st) {
    switch (bdt) {
        case DT_F32: return mul_o<A, float>(a, b, o, odt, n, st);
        case DT_F16: return mul_o<A, __half>(a, b, o, odt, n, st);
        case DT_BF16: return mul_o<A, __hip_bfloat16>(a, b, o, odt, n, st);
        default: return hipErrorInvalidValue;
    }
}

template <class O>
__global__ void sub_kernel(const float* __restrict__ a, const float* __restrict__ b, O* __restrict__ o, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = cvt_store<O>(a[i] - b[i]);
}

hipError_t sub_launch(const float* a, const float* b, void* o, int odt, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    switch (odt) {
        case DT_F32: hipLaunchKernelGGL((sub_kernel<float>), dim3(grid_for(n)), dim3(256), 0, st, a, b, (float*)o, n); break;
        case DT_F16: hipLaunchKernelGGL((sub_kernel<__half>), dim3(grid_for(n)), dim3(256), 0, st, a, b, (__half*)o, n); break;
        case DT_BF16: hipLaunchKernelGGL((sub_kernel<__hip_bfloat16>), dim3(grid_for(n)), dim3(256), 0, st, a, b, (__hip_bfloat16*)o, n); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t mul_launch(const void* a, int adt, const void* b, int bdt, void* o, int odt, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (adt == DT_F64 || bdt == DT_F64 || odt == DT_F64) {
        if (adt != DT_F64 || bdt != DT_F64 || odt != DT_F64) return hipErrorInvalidValue;
        hipLaunchKernelGGL(mul_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const double*)a, (const double*)b,
                           (double*)o, n);
        return hipGetLastError();
    }
    switch (adt) {
        case DT_F32: return mul_b<float>(a, b, bdt, o, odt, n, st);
        case DT_F16: return mul_b<__half>(a, b, bdt, o, odt, n, st);
        case DT_BF16: return mul_b<__hip_bfloat16>(a, b, bdt, o, odt, n, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ctsim
