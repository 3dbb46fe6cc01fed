// FDK pre-processing on the device (fbp/reconFDK.m:33-76): cosine weight, optional
// per-view short-scan weight, and the ramp filter applied row by row in the frequency
// domain (fbp/filter/filterFreq.m:42-51: zero-pad each detector row to the filter
// length, multiply its FFT by H, keep the real part of the inverse FFT, truncate).
//
// FP32 plans whose filter length is a power of two from 256 to 2048 (and covers the row)
// take the fused form below (one kernel, one read of the views and one write); the others
// (FP64, other lengths) the hipFFT chain.
//
// hipFFT chain, for a chunk of views, sinogram [view][u][v] (v fastest, MATLAB order):
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

// the weighted sample (u, v) of view b: the cosine weight and, when given, the view's
// short-scan weight of cell u
template <class Tin, class C>
__device__ __forceinline__ C fdk_weighted(const Tin* __restrict__ sino, const float* __restrict__ vw, int b, int u,
                                          int v, int nu, int nv, const FdkGeom& g) {
    const C um = (C(u) - C(0.5) * C(nu - 1) + C(g.ou)) * C(g.du);
    const C vm = (C(v) - C(0.5) * C(nv - 1) + C(g.ov)) * C(g.dv);
    C val = ld<C>(sino + ((size_t)b * nu + u) * nv + v) * fdk_weight<C>(um, vm, C(g.sad), C(g.sdd), g.flat);
    if (vw) val *= C(vw[(size_t)b * nu + u]);
    return val;
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
        tile[k][tx] = (u < nu && v < nv) ? fdk_weighted<Tin, C>(sino, vw, b, u, v, nu, nv, g) : C(0);
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
// Fused form (FP32, filter lengths L = 2^8 .. 2^11 with nu <= L): one block per (view,
// kFV detector rows).  The block loads its rows' weighted samples once into an LDS tile
// [u][v], filters them two rows at a time as one complex sequence (row a in the real part,
// row b in the imaginary part, zero padded to L: H is real and even, so the filtered
// sequence is filtered(a) + i filtered(b)), and stores the tile once.  The transforms are
// Stockham autosort FFTs in LDS: L/8 threads, stages of radix 8 (one butterfly per thread)
// and a last one of radix 4 or 2, natural order in and out; the first forward stage reads
// the tile, the first inverse stage multiplies by H / L as it loads, the last inverse stage
// writes the tile.  Twiddles exp(-2 pi i m / L) come from a table formed in FP64 on the host.
// Against the hipFFT chain (five passes over padded rows in HBM) this reads the views and
// writes the filtered views once.
// ---------------------------------------------------------------------------------
constexpr int kFV = 8;  // detector rows per block (four row pairs)

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (sign i)
__device__ __forceinline__ float2 cmul_i(float2 a, float sign) { return make_float2(-sign * a.y, sign * a.x); }

// size-R DFT of v[0..R) with exponent sign (-1 forward, +1 inverse), decimation in
// frequency; outputs in natural order
template <int R>
__device__ __forceinline__ void small_dft(float2* v, float sign);
template <>
__device__ __forceinline__ void small_dft<2>(float2* v, float) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void small_dft<4>(float2* v, float sign) {
    const float2 a0 = cadd(v[0], v[2]), a2 = csub(v[0], v[2]);
    const float2 a1 = cadd(v[1], v[3]), a3 = cmul_i(csub(v[1], v[3]), sign);
    v[0] = cadd(a0, a1);
    v[1] = cadd(a2, a3);
    v[2] = csub(a0, a1);
    v[3] = csub(a2, a3);
}
template <>
__device__ __forceinline__ void small_dft<8>(float2* v, float sign) {
    const float r = 0.70710678118654752f;
    const float2 w1 = make_float2(r, sign * r), w3 = make_float2(-r, sign * r);
    const float2 a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
    const float2 a1 = cadd(v[1], v[5]), a5 = cmul(csub(v[1], v[5]), w1);
    const float2 a2 = cadd(v[2], v[6]), a6 = cmul_i(csub(v[2], v[6]), sign);
    const float2 a3 = cadd(v[3], v[7]), a7 = cmul(csub(v[3], v[7]), w3);
    const float2 b0 = cadd(a0, a2), b2 = csub(a0, a2);
    const float2 b1 = cadd(a1, a3), b3 = cmul_i(csub(a1, a3), sign);
    const float2 b4 = cadd(a4, a6), b6 = csub(a4, a6);
    const float2 b5 = cadd(a5, a7), b7 = cmul_i(csub(a5, a7), sign);
    v[0] = cadd(b0, b1);
    v[4] = csub(b0, b1);
    v[2] = cadd(b2, b3);
    v[6] = csub(b2, b3);
    v[1] = cadd(b4, b5);
    v[5] = csub(b4, b5);
    v[3] = cadd(b6, b7);
    v[7] = csub(b6, b7);
}

// one Stockham stage of radix R after stages whose radices multiply to NS: butterfly j
// (8 / R of them per thread) takes x[j + r L/R], r < R, times W_{NS R}^{(j mod NS) r},
// and its DFT goes to (j / NS) NS R + j mod NS + r NS.  load(m) gives x[m]; store(m, y).
template <int L, int R, int NS, class Load, class Store>
__device__ __forceinline__ void fft_stage(int tid, float sign, const float2* __restrict__ tw, Load load, Store store) {
    constexpr int NB = L / R, PER = 8 / R;
    float2 v[8];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = tid + q * (L / 8);
        const int k = j & (NS - 1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 x = load(j + r * NB);
            if (NS > 1 && r > 0) {
                float2 w = tw[k * r * (L / (NS * R))];
                w.y *= -sign;  // the table holds exp(-2 pi i m / L); sign +1 conjugates it
                x = cmul(x, w);
            }
            v[q * R + r] = x;
        }
        small_dft<R>(v + q * R, sign);
    }
    __syncthreads();  // every load of this stage before any store
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = tid + q * (L / 8);
        const int base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
        for (int r = 0; r < R; ++r) store(base + r * NS, v[q * R + r]);
    }
    __syncthreads();
}

// the whole transform: stages of radix 8, then one of L / 8^k (4 or 2) if L is not a
// power of 8.  in / out: the first stage's load and the last stage's store; the stages
// between read and write W.
template <int L, int S, int NS, class In, class Out>
__device__ __forceinline__ void fft_run(int tid, float sign, const float2* __restrict__ tw, float2* W, In in,
                                       Out out) {
    constexpr int REM = L / NS;  // points per butterfly still to combine
    constexpr int R = REM >= 8 ? 8 : REM;
    constexpr bool first = NS == 1, last = R == REM;
    auto ldW = [&](int m) { return W[m]; };
    auto stW = [&](int m, float2 y) { W[m] = y; };
    if constexpr (first && last) {
        fft_stage<L, R, NS>(tid, sign, tw, in, out);
    } else if constexpr (first) {
        fft_stage<L, R, NS>(tid, sign, tw, in, stW);
    } else if constexpr (last) {
        fft_stage<L, R, NS>(tid, sign, tw, ldW, out);
    } else {
        fft_stage<L, R, NS>(tid, sign, tw, ldW, stW);
    }
    if constexpr (!last) fft_run<L, S + 1, NS * R>(tid, sign, tw, W, in, out);
}

template <class Tin, class Tout, int LOG2L>
__global__ __launch_bounds__((1 << LOG2L) / 8) void fdk_fused_kernel(const Tin* __restrict__ sino,
                                                                    const float* __restrict__ vw,
                                                                    Tout* __restrict__ out, int nu, int nv,
                                                                    const float* __restrict__ hs,
                                                                    const float2* __restrict__ tw, FdkGeom g) {
    constexpr int L = 1 << LOG2L, NT = L / 8;
    extern __shared__ __align__(16) float fdk_lds[];
    float2* W = reinterpret_cast<float2*>(fdk_lds);  // L complex
    float* T = fdk_lds + 2 * L;                      // [u][kFV]
    const int tid = threadIdx.x;
    const int b = blockIdx.y, v0 = blockIdx.x * kFV;
    const int nvt = min(kFV, nv - v0);  // rows of this tile
    for (int e = tid; e < nu * kFV; e += NT) {
        const int u = e / kFV, vt = e % kFV;
        T[e] = vt < nvt ? fdk_weighted<Tin, float>(sino, vw, b, u, v0 + vt, nu, nv, g) : 0.f;
    }
    __syncthreads();
    for (int p = 0; 2 * p < nvt; ++p) {
        const int c0 = 2 * p, c1 = 2 * p + 1;
        // forward: rows c0 + i c1, zero past nu
        fft_run<L, 0, 1>(
            tid, -1.f, tw, W,
            [&](int m) { return m < nu ? make_float2(T[m * kFV + c0], T[m * kFV + c1]) : make_float2(0.f, 0.f); },
            [&](int m, float2 y) { W[m] = y; });
        // inverse of H / L times the spectrum (H even: bin m and L - m share H[min(m, L - m)]); the
        // first u < nu of the result go back to the tile
        fft_run<L, 0, 1>(
            tid, 1.f, tw, W,
            [&](int m) {
                const float h = hs[m <= L / 2 ? m : L - m];
                const float2 x = W[m];
                return make_float2(x.x * h, x.y * h);
            },
            [&](int m, float2 y) {
                if (m < nu) {
                    T[m * kFV + c0] = y.x;
                    T[m * kFV + c1] = y.y;
                }
            });
    }
    for (int e = tid; e < nu * kFV; e += NT) {
        const int u = e / kFV, vt = e % kFV;
        if (vt < nvt) out[((size_t)b * nu + u) * nv + v0 + vt] = cvt_store<Tout>(T[e]);
    }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
struct FdkPlanImpl {
    FdkGeom g{};
    int nu = 0, nv = 0, nviews = 0, L = 0, chunk = 0, f64 = 0;
    int fused_log2 = 0;        // log2 L of the fused form, 0 for the hipFFT chain
    void* rows = nullptr;      // chunk x nv x L real
    void* spec = nullptr;      // chunk x nv x (L/2+1) complex
    void* H = nullptr;         // L/2+1 real (the fused form: H / L in FP32)
    float2* tw = nullptr;      // fused form: exp(-2 pi i m / L), m < L
    float* vw = nullptr;       // nviews x nu, or null
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false;
};

// dynamic LDS of the fused form: the complex work row and the [nu][kFV] tile
static size_t fused_lds_bytes(int L, int nu) { return (size_t)L * 8 + (size_t)nu * kFV * 4; }

static hipError_t fft_err(hipfftResult r) { return r == HIPFFT_SUCCESS ? hipSuccess : hipErrorUnknown; }

void fdk_plan_destroy(FdkPlanImpl* p) {
    if (!p) return;
    if (p->have_fwd) hipfftDestroy(p->fwd);
    if (p->have_inv) hipfftDestroy(p->inv);
    if (p->rows) (void)hipFree(p->rows);
    if (p->spec) (void)hipFree(p->spec);
    if (p->H) (void)hipFree(p->H);
    if (p->tw) (void)hipFree(p->tw);
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
    const int nk = L / 2 + 1;
    // FP32 plans with a power-of-two filter of 256 .. 2048 taps covering the row take the
    // fused form (no workspace, no hipFFT plans)
    int lg = 0;
    while ((1 << lg) < L) ++lg;
    if (!f64 && (1 << lg) == L && lg >= 8 && lg <= 11 && nu <= L && fused_lds_bytes(L, nu) <= 65536) {
        p->fused_log2 = lg;
        hipError_t e = hipMalloc(&p->H, (size_t)nk * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&p->tw, (size_t)L * sizeof(float2));
        if (e == hipSuccess) {
            float* hf = new float[nk];
            for (int k = 0; k < nk; ++k) hf[k] = (float)(H[k] / (double)L);  // L is a power of two: exact
            e = hipMemcpy(p->H, hf, (size_t)nk * 4, hipMemcpyHostToDevice);
            delete[] hf;
        }
        if (e == hipSuccess) {
            float2* t = new float2[L];
            const double pi = 3.14159265358979323846;
            for (int m = 0; m < L; ++m) t[m] = make_float2((float)cos(2 * pi * m / L), (float)-sin(2 * pi * m / L));
            e = hipMemcpy(p->tw, t, (size_t)L * sizeof(float2), hipMemcpyHostToDevice);
            delete[] t;
        }
        if (e == hipSuccess && view_weights) {
            e = hipMalloc((void**)&p->vw, (size_t)nviews * nu * 4);
            if (e == hipSuccess) e = hipMemcpy(p->vw, view_weights, (size_t)nviews * nu * 4, hipMemcpyHostToDevice);
        }
        p->chunk = nviews;
        if (e != hipSuccess) {
            fdk_plan_destroy(p);
            return e;
        }
        *out = p;
        return hipSuccess;
    }
    const size_t esz = f64 ? 8 : 4;
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

template <class Tin, class Tout>
static hipError_t run_fused_out(FdkPlanImpl* p, int nsub, const Tin* sino, const float* vw, Tout* out,
                                hipStream_t st) {
    const dim3 grid((unsigned)((p->nv + kFV - 1) / kFV), (unsigned)nsub), blk((unsigned)(p->L / 8));
    const size_t lds = fused_lds_bytes(p->L, p->nu);
    const float* hs = (const float*)p->H;
    switch (p->fused_log2) {
        case 8: hipLaunchKernelGGL((fdk_fused_kernel<Tin, Tout, 8>), grid, blk, lds, st, sino, vw, out, p->nu, p->nv, hs, p->tw, p->g); break;
        case 9: hipLaunchKernelGGL((fdk_fused_kernel<Tin, Tout, 9>), grid, blk, lds, st, sino, vw, out, p->nu, p->nv, hs, p->tw, p->g); break;
        case 10: hipLaunchKernelGGL((fdk_fused_kernel<Tin, Tout, 10>), grid, blk, lds, st, sino, vw, out, p->nu, p->nv, hs, p->tw, p->g); break;
        case 11: hipLaunchKernelGGL((fdk_fused_kernel<Tin, Tout, 11>), grid, blk, lds, st, sino, vw, out, p->nu, p->nv, hs, p->tw, p->g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class Tin>
static hipError_t run_fused(FdkPlanImpl* p, int view0, int nsub, const void* sino, void* out, int odt,
                            hipStream_t st) {
    if (nsub == 0) return hipSuccess;
    const float* vw = p->vw ? p->vw + (size_t)view0 * p->nu : nullptr;
    const Tin* si = (const Tin*)sino;
    switch (odt) {
        case DT_F32: return run_fused_out<Tin, float>(p, nsub, si, vw, (float*)out, st);
        case DT_F16: return run_fused_out<Tin, __half>(p, nsub, si, vw, (__half*)out, st);
        case DT_BF16: return run_fused_out<Tin, __hip_bfloat16>(p, nsub, si, vw, (__hip_bfloat16*)out, st);
        case DT_F64: return run_fused_out<Tin, double>(p, nsub, si, vw, (double*)out, st);
        default: return hipErrorInvalidValue;
    }
}

template <class C, class Tin>
static hipError_t run_all(FdkPlanImpl* p, int view0, int nsub, const void* sino, void* out, int odt, hipStream_t st) {
    if constexpr (sizeof(C) == 4)
        if (p->fused_log2) return run_fused<Tin>(p, view0, nsub, sino, out, odt, st);
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
