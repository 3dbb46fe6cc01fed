// Ray-driven forward projection, op "proj,rd" (geom/fbct_os/fbct_proj.c:3-153,
// geom/cbct_os/cbct_proj.c:10-198).  Each detector cell's ray is marched from its entry
// into the image box in steps of dx/2, the parameter accumulated by repeated float
// addition, and the image is sampled by bilinear (2D) or trilinear (3D) interpolation at
// interior positions (the voxels of the outer shell are skipped, as in the reference).
//
// The sample positions are set by that float accumulation, so the result is a
// discontinuous function of last-ulp differences: the march reproduces the reference's
// operations one for one.  The per-(view, column) set-up (double sin / atan, float sinf,
// the slab tests) comes from a host-built table (RayRec); in the kernels FMA contraction
// is off and divisions and square roots are the correctly rounded ones (the per-sample
// divisions by the voxel size through div_rn).
//
// 3D: one wavefront per (view, detector column, 64 detector rows); lanes are rows.  The
// ray's in-plane position is the same on every lane (x, y depend on t only), so the
// bounds test is wave-uniform, and in the ZXY volume the lanes' trilinear reads fall on
// neighbouring z of the same columns.  2D: one lane per (view, detector cell).
#include "internal.h"

namespace ctsim {

constexpr int kRdWave = 64;
constexpr int kRdBlock = 256;

// x / d rounded to float, from the double reciprocal of d: the double product is within
// 2^-52 (relative) of x / d, while a quotient of two floats is never within 2^-50 of a
// float rounding midpoint (a midpoint's odd 25-bit significand times d's cannot be x's
// 24-bit one), so rounding the product to float gives the correctly rounded division.
// Three instructions in place of the division's scale / reciprocal / fix-up sequence.
__device__ __forceinline__ float div_rn(float x, double inv_d) { return (float)((double)x * inv_d); }

template <class Tin, class Tout>
__global__ __launch_bounds__(kRdBlock) void cbct_rd_kernel(const RayRec* __restrict__ rays, int view0, int vstep,
                                                           int nviews, RdConst c, const Tin* __restrict__ vol,
                                                           Tout* __restrict__ proj) {
#pragma clang fp contract(off)
    const int nvc = (c.nv + kRdWave - 1) / kRdWave;
    const long long items = (long long)nviews * c.nu * nvc;
    const long long item = (long long)blockIdx.x * (kRdBlock / kRdWave) + threadIdx.x / kRdWave;
    if (item >= items) return;
    const int vc = (int)(item % nvc);
    const int iu = (int)((item / nvc) % c.nu);
    const int k = (int)(item / ((long long)nvc * c.nu));
    const int lane = (int)(threadIdx.x % kRdWave);
    const int iv = vc * kRdWave + lane;
    const RayRec r = rays[(size_t)(view0 + (size_t)k * vstep) * c.nu + iu];
    const float rdz = ((float)iv * c.dv + c.wv - 0.0f) / r.spd;
    const double zlim = (double)c.nz - 1.1;
    const size_t sy = (size_t)c.nx * c.nz;
    float temp = 0.0f;
    bool done = iv >= c.nv;
    for (float t = r.tn; t < r.tf; t = t + c.delta) {
        if (__ballot(!done) == 0ull) break;  // every row of the wave has left the volume
        if (done) continue;
        const float z = div_rn(0.0f + t * rdz - r.czmin, c.inv_dz);
        if (z < 1.0f) continue;
        if ((double)z > zlim) {
            done = true;
            continue;
        }
        const float x = div_rn(r.sx + t * r.rdx - c.cmin_x, c.inv_dx);
        const float y = div_rn(r.sy + t * r.rdy - c.cmin_y, c.inv_dy);
        const int ix = (int)floorf(x), iy = (int)floorf(y), iz = (int)floorf(z);
        const float wzl = z - (float)iz;
        if (ix > 0 && ix < c.nx - 1 && iy > 0 && iy < c.ny - 1) {
            const float wxl = x - (float)ix, wyl = y - (float)iy;
            const Tin* p = vol + ((size_t)iy * c.nx + ix) * c.nz + iz;  // (ix, iy, iz); x +nz, y +nx nz, z +1
            const float v000 = ld<float>(p), v100 = ld<float>(p + c.nz), v010 = ld<float>(p + sy),
                        v110 = ld<float>(p + sy + c.nz);
            const float v001 = ld<float>(p + 1), v101 = ld<float>(p + c.nz + 1), v011 = ld<float>(p + sy + 1),
                        v111 = ld<float>(p + sy + c.nz + 1);
            temp = temp +
                   (1.0f - wzl) * ((1.0f - wyl) * ((1.0f - wxl) * v000 + wxl * v100) +
                                   wyl * ((1.0f - wxl) * v010 + wxl * v110)) +
                   wzl * ((1.0f - wyl) * ((1.0f - wxl) * v001 + wxl * v101) + wyl * ((1.0f - wxl) * v011 + wxl * v111));
        }
    }
    if (iv < c.nv)
        proj[((size_t)k * c.nu + iu) * c.nv + iv] = cvt_store<Tout>(temp * c.delta * sqrtf(1.0f + rdz * rdz) / 10.0f);
}

template <class Tin, class Tout>
__global__ __launch_bounds__(kRdBlock) void fbct_rd_kernel(const RayRec* __restrict__ rays, int view0, int vstep,
                                                           int nviews, RdConst c, const Tin* __restrict__ img,
                                                           Tout* __restrict__ proj) {
#pragma clang fp contract(off)
    const long long idx = (long long)blockIdx.x * kRdBlock + threadIdx.x;
    if (idx >= (long long)nviews * c.nu) return;
    const int iu = (int)(idx % c.nu);
    const int k = (int)(idx / c.nu);
    const RayRec r = rays[(size_t)(view0 + (size_t)k * vstep) * c.nu + iu];
    float temp = 0.0f;
    for (float t = r.tn; t < r.tf; t = t + c.delta) {
        const float x = div_rn(r.sx + t * r.rdx - c.cmin_x, c.inv_dx);
        const float y = div_rn(r.sy + t * r.rdy - c.cmin_y, c.inv_dy);
        const int ix = (int)floorf(x), iy = (int)floorf(y);
        if (ix > 0 && ix < c.nx - 1 && iy > 0 && iy < c.ny - 1) {
            const float wxl = x - (float)ix, wyl = y - (float)iy;
            const Tin* p = img + (size_t)iy * c.nx + ix;
            temp = temp + (1.0f - wyl) * ((1.0f - wxl) * ld<float>(p) + wxl * ld<float>(p + 1)) +
                   wyl * ((1.0f - wxl) * ld<float>(p + c.nx) + wxl * ld<float>(p + c.nx + 1));
        }
    }
    proj[(size_t)k * c.nu + iu] = cvt_store<Tout>(temp * c.delta / 10.0f);
}

template <class Tin>
static hipError_t cbct_rd_out(const RayRec* rays, int view0, int vstep, int nviews, const RdConst& c, const void* vol,
                              void* proj, int proj_dt, hipStream_t st) {
    const long long items = (long long)nviews * c.nu * ((c.nv + kRdWave - 1) / kRdWave);
    const unsigned blocks = (unsigned)((items + (kRdBlock / kRdWave) - 1) / (kRdBlock / kRdWave));
    switch (proj_dt) {
        case DT_F32: hipLaunchKernelGGL((cbct_rd_kernel<Tin, float>), dim3(blocks), dim3(kRdBlock), 0, st, rays, view0, vstep, nviews, c, (const Tin*)vol, (float*)proj); break;
        case DT_F16: hipLaunchKernelGGL((cbct_rd_kernel<Tin, __half>), dim3(blocks), dim3(kRdBlock), 0, st, rays, view0, vstep, nviews, c, (const Tin*)vol, (__half*)proj); break;
        case DT_BF16: hipLaunchKernelGGL((cbct_rd_kernel<Tin, __hip_bfloat16>), dim3(blocks), dim3(kRdBlock), 0, st, rays, view0, vstep, nviews, c, (const Tin*)vol, (__hip_bfloat16*)proj); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t cbct_rd_launch(const RayRec* rays, int view0, int vstep, int nviews, const RdConst& c, const void* vol,
                          int vol_dt, void* proj, int proj_dt, hipStream_t st) {
    if (nviews <= 0) return hipSuccess;
    switch (vol_dt) {
        case DT_F32: return cbct_rd_out<float>(rays, view0, vstep, nviews, c, vol, proj, proj_dt, st);
        case DT_F16: return cbct_rd_out<__half>(rays, view0, vstep, nviews, c, vol, proj, proj_dt, st);
        case DT_BF16: return cbct_rd_out<__hip_bfloat16>(rays, view0, vstep, nviews, c, vol, proj, proj_dt, st);
        default: return hipErrorInvalidValue;
    }
}

template <class Tin>
static hipError_t fbct_rd_out(const RayRec* rays, int view0, int vstep, int nviews, const RdConst& c, const void* img,
                              void* proj, int proj_dt, hipStream_t st) {
    const long long n = (long long)nviews * c.nu;
    const unsigned blocks = (unsigned)((n + kRdBlock - 1) / kRdBlock);
    switch (proj_dt) {
        case DT_F32: hipLaunchKernelGGL((fbct_rd_kernel<Tin, float>), dim3(blocks), dim3(kRdBlock), 0, st, rays, view0, vstep, nviews, c, (const Tin*)img, (float*)proj); break;
        case DT_F16: hipLaunchKernelGGL((fbct_rd_kernel<Tin, __half>), dim3(blocks), dim3(kRdBlock), 0, st, rays, view0, vstep, nviews, c, (const Tin*)img, (__half*)proj); break;
        case DT_BF16: hipLaunchKernelGGL((fbct_rd_kernel<Tin, __hip_bfloat16>), dim3(blocks), dim3(kRdBlock), 0, st, rays, view0, vstep, nviews, c, (const Tin*)img, (__hip_bfloat16*)proj); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t fbct_rd_launch(const RayRec* rays, int view0, int vstep, int nviews, const RdConst& c, const void* img,
                          int img_dt, void* proj, int proj_dt, hipStream_t st) {
    if (nviews <= 0) return hipSuccess;
    switch (img_dt) {
        case DT_F32: return fbct_rd_out<float>(rays, view0, vstep, nviews, c, img, proj, proj_dt, st);
        case DT_F16: return fbct_rd_out<__half>(rays, view0, vstep, nviews, c, img, proj, proj_dt, st);
        case DT_BF16: return fbct_rd_out<__hip_bfloat16>(rays, view0, vstep, nviews, c, img, proj, proj_dt, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ctsim
