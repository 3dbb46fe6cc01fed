// FDK pre-processing on the device (fbp/reconFDK.m:33-76): cosine weight, optional
// per-view short-scan weight, and the ramp filter applied row by row in the frequency
// domain (fbp/filter/filterFreq.m:42-51: zero-pad each detector row to the filter
// length, multiply its FFT by H, keep the real part of the inverse FFT, truncate).
//
// Data flow for a chunk of views, sinogram [view][u][v] (v fastest, MATLAB order):
//   weight_pad_kernel   [view][u][v] -> rows[view][v][0..L)   weighted, zero padded,
//                                        transposed through LDS so rows are u-fastest
//   hipFFT R2C          rows -> spectrum[view][v][0..L/2]
//   spectrum_mul_kernel spectrum *= H[k]                       (H real, even)
//   hipFFT C2R          spectrum -> rows                        (unnormalised)
//   crop_kernel         rows[view][v][u] / L -> out[view][u][v] for u < nu, stored in
//                                        the output dtype (F32, F16, BF16 or F64)
#include <hipfft/hipfft.h>

#include "internal.h"

namespace ctsim {

constexpr int kT = 64;  // transpose tile (u x v)

// cosine weight of detector cell (u, v) in mm: flat 1 = reconFDK.m:39, arc (flat 0) =
// reconFDK.m:44 (at v = 0 also the fan-beam arc weight of reconFBP2d.m:38), flat 2 = the
// fan-beam flat-panel weight of reconFBP2d.m:32, SAD / sqrt(u^2 + SDD^2)
template <class C>
__device__ __forceinline__ C fdk_weight(C u, C v, C sad, C sdd, int flat) {
    if (flat == 2) return sad / sqrt(u * u + sdd * sdd);
    if (flat) return sad * sqrt(C(1) + u * u / (sdd * sdd)) / sqrt(u * u + v * v + sdd * sdd);
    return sad / sdd * cos(u / sqrt(sdd * sdd + v * v));
}

template <class Tin, class C>
__global__ __launch_bounds__(256) void weight_pad_kernel(const Tin* __restrict__ sino, const float* __restrict__ vw,
                                                         C* __restrict__ rows, int nu, int nv, int L, FdkGeom g) {
    __shared__ C tile[kT][kT + 1];
    const int u0 = blockIdx.x * kT, v0 = blockIdx.y * kT, b = blockIdx.z;
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;  // 64 x 4 threads
    // read: lanes along v (contiguous), rows of the tile along u
    for (int k = ty; k < kT; k += 4) {
        const int u = u0 + k, v = v0 + tx;
        C val = C(0);
        if (u < nu && v < nv) {
            const C um = (C(u) - C(0.5) * C(nu - 1) + C(g.ou)) * C(g.du);
            const C vm = (C(v) - C(0.5) * C(nv - 1) + C(g.ov)) * C(g.dv);
            val = ld<C>(sino + ((size_t)b * nu + u) * nv + v) * fdk_weight<C>(um, vm, C(g.sad), C(g.sdd), g.flat);
            if (vw) val *= C(vw[(size_t)b * nu + u]);
        }
        tile[k][tx] = val;
    }
    __syncthreads();
    // write: lanes along u (contiguous in the padded row)
    for (int k = ty; k < kT; k += 4) {
        const int v = v0 + k, u = u0 + tx;
        if (v < nv && u < L) rows[((size_t)b * nv + v) * L + u] = tile[tx][k];
    }
}

template <class CX, class C>
__global__ void spectrum_mul_kernel(CX* __restrict__ spec, const C* __restrict__ H, size_t nrows, int nk) {
    const size_t n = nrows * (size_t)nk;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const C h = H[i % nk];
        spec[i].x *= h;
        spec[i].y *= h;
    }
}

template <class C, class Tout>
__global__ __launch_bounds__(256) void crop_kernel(const C* __restrict__ rows, Tout* __restrict__ out, int nu, int nv,
                                                   int L, C inv_l) {
    __shared__ C tile[kT][kT + 1];
    const int u0 = blockIdx.x * kT, v0 = blockIdx.y * kT, b = blockIdx.z;
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
    for (int k = ty; k < kT; k += 4) {  // read along u
        const int v = v0 + k, u = u0 + tx;
        tile[k][tx] = (v < nv && u < nu) ? rows[((size_t)b * nv + v) * L + u] * inv_l : C(0);
    }
    __syncthreads();
    for (int k = ty; k < kT; k += 4) {  // write along v
        const int u = u0 + k, v = v0 + tx;
        if (u < nu && v < nv) out[((size_t)b * nu + u) * nv + v] = cvt_store<Tout>(tile[tx][k]);
    }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
struct FdkPlanImpl {
    FdkGeom g{};
    int nu = 0, nv = 0, nviews = 0, L = 0, chunk = 0, f64 = 0;
    void* rows = nullptr;      // chunk x nv x L real
    void* spec = nullptr;      // chunk x nv x (L/2+1) complex
    void* H = nullptr;         // L/2+1 real
    float* vw = nullptr;       // nviews x nu, or null
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false;
};

static hipError_t fft_err(hipfftResult r) { return r == HIPFFT_SUCCESS ? hipSuccess : hipErrorUnknown; }

void fdk_plan_destroy(FdkPlanImpl* p) {
    if (!p) return;
    if (p->have_fwd) hipfftDestroy(p->fwd);
    if (p->have_inv) hipfftDestroy(p->inv);
    if (p->rows) (void)hipFree(p->rows);
    if (p->spec) (void)hipFree(p->spec);
    if (p->H) (void)hipFree(p->H);
    if (p->vw) (void)hipFree(p->vw);
    delete p;
}

hipError_t fdk_plan_create(FdkPlanImpl** out, const FdkGeom& g, int nu, int nv, int nviews, const double* H, int L,
                           const float* view_weights, int f64, size_t max_workspace) {
    *out = nullptr;
    FdkPlanImpl* p = new FdkPlanImpl;
    p->g = g;
    p->nu = nu;
    p->nv = nv;
    p->nviews = nviews;
    p->L = L;
    p->f64 = f64;
    const size_t esz = f64 ? 8 : 4;
    const int nk = L / 2 + 1;
    const size_t per_view = (size_t)nv * ((size_t)L * esz + (size_t)nk * 2 * esz);
    long long chunk = (long long)(max_workspace / per_view);
    if (chunk < 1) chunk = 1;
    if (chunk > nviews) chunk = nviews;
    p->chunk = (int)chunk;
    hipError_t e = hipMalloc(&p->rows, (size_t)chunk * nv * L * esz);
    if (e == hipSuccess) e = hipMalloc(&p->spec, (size_t)chunk * nv * nk * 2 * esz);
    if (e == hipSuccess) e = hipMalloc(&p->H, (size_t)nk * esz);
    if (e == hipSuccess) {
        if (f64) {
            e = hipMemcpy(p->H, H, (size_t)nk * 8, hipMemcpyHostToDevice);
        } else {
            float* hf = new float[nk];
            for (int k = 0; k < nk; ++k) hf[k] = (float)H[k];
            e = hipMemcpy(p->H, hf, (size_t)nk * 4, hipMemcpyHostToDevice);
            delete[] hf;
        }
    }
    if (e == hipSuccess && view_weights) {
        e = hipMalloc((void**)&p->vw, (size_t)nviews * nu * 4);
        if (e == hipSuccess) e = hipMemcpy(p->vw, view_weights, (size_t)nviews * nu * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        int n[1] = {L};
        const int batch = (int)chunk * nv;
        hipfftResult r = hipfftPlanMany(&p->fwd, 1, n, nullptr, 1, L, nullptr, 1, nk, f64 ? HIPFFT_D2Z : HIPFFT_R2C,
                                        batch);
        p->have_fwd = r == HIPFFT_SUCCESS;
        if (r == HIPFFT_SUCCESS)
            r = hipfftPlanMany(&p->inv, 1, n, nullptr, 1, nk, nullptr, 1, L, f64 ? HIPFFT_Z2D : HIPFFT_C2R, batch);
        p->have_inv = p->have_fwd && r == HIPFFT_SUCCESS;
        e = fft_err(r);
    }
    if (e != hipSuccess) {
        fdk_plan_destroy(p);
        return e;
    }
    *out = p;
    return hipSuccess;
}

// views b0 .. b0+nb-1 of d_sino / d_out; their per-view weights start at vw (or null)
template <class C, class Tin>
static hipError_t run_chunk(FdkPlanImpl* p, int b0, int nb, const Tin* sino, const float* vw, void* out, int odt,
                            hipStream_t st) {
    const int nu = p->nu, nv = p->nv, L = p->L, nk = L / 2 + 1;
    C* rows = (C*)p->rows;
    const dim3 blk(256);
    hipLaunchKernelGGL((weight_pad_kernel<Tin, C>), dim3((L + kT - 1) / kT, (nv + kT - 1) / kT, nb), blk, 0, st,
                       sino + (size_t)b0 * nu * nv, vw, rows, nu, nv, L, p->g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the FFT plans are made for full chunks; a short last chunk transforms the
    // whole buffer (rows past nb*nv hold the previous chunk's data and are not used)
    if (hipfftSetStream(p->fwd, st) != HIPFFT_SUCCESS || hipfftSetStream(p->inv, st) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    const size_t nrows = (size_t)nb * nv;
    if constexpr (sizeof(C) == 8) {
        e = fft_err(hipfftExecD2Z(p->fwd, rows, (hipfftDoubleComplex*)p->spec));
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((spectrum_mul_kernel<hipfftDoubleComplex, double>), dim3(2048), dim3(256), 0, st,
                           (hipfftDoubleComplex*)p->spec, (const double*)p->H, nrows, nk);
        e = fft_err(hipfftExecZ2D(p->inv, (hipfftDoubleComplex*)p->spec, rows));
    } else {
        e = fft_err(hipfftExecR2C(p->fwd, rows, (hipfftComplex*)p->spec));
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((spectrum_mul_kernel<hipfftComplex, float>), dim3(2048), dim3(256), 0, st,
                           (hipfftComplex*)p->spec, (const float*)p->H, nrows, nk);
        e = fft_err(hipfftExecC2R(p->inv, (hipfftComplex*)p->spec, rows));
    }
    if (e != hipSuccess) return e;
    const dim3 grid((nu + kT - 1) / kT, (nv + kT - 1) / kT, nb);
    const C inv_l = C(1) / C(L);
    const size_t off = (size_t)b0 * nu * nv;
    switch (odt) {
        case DT_F32: hipLaunchKernelGGL((crop_kernel<C, float>), grid, blk, 0, st, rows, (float*)out + off, nu, nv, L, inv_l); break;
        case DT_F16: hipLaunchKernelGGL((crop_kernel<C, __half>), grid, blk, 0, st, rows, (__half*)out + off, nu, nv, L, inv_l); break;
        case DT_BF16: hipLaunchKernelGGL((crop_kernel<C, __hip_bfloat16>), grid, blk, 0, st, rows, (__hip_bfloat16*)out + off, nu, nv, L, inv_l); break;
        case DT_F64: hipLaunchKernelGGL((crop_kernel<C, double>), grid, blk, 0, st, rows, (double*)out + off, nu, nv, L, inv_l); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class C, class Tin>
static hipError_t run_all(FdkPlanImpl* p, int view0, int nsub, const void* sino, void* out, int odt, hipStream_t st) {
    // d_sino / d_out hold views view0 .. view0+nsub-1; the per-view weights are
    // indexed by the plan's view number
    for (int b = 0; b < nsub; b += p->chunk) {
        const int nb = nsub - b < p->chunk ? nsub - b : p->chunk;
        const float* vw = p->vw ? p->vw + (size_t)(view0 + b) * p->nu : nullptr;
        hipError_t e = run_chunk<C, Tin>(p, b, nb, (const Tin*)sino, vw, out, odt, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t fdk_filter_launch(FdkPlanImpl* p, int view0, int nsub, const void* sino, int sdt, void* out, int odt,
                             hipStream_t st) {
    if (p->f64) {
        if (sdt != DT_F64) return hipErrorInvalidValue;
        return run_all<double, double>(p, view0, nsub, sino, out, odt, st);
    }
    switch (sdt) {
        case DT_F32: return run_all<float, float>(p, view0, nsub, sino, out, odt, st);
        case DT_F16: return run_all<float, __half>(p, view0, nsub, sino, out, odt, st);
        case DT_BF16: return run_all<float, __hip_bfloat16>(p, view0, nsub, sino, out, odt, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ctsim
