// Data side of the reconstruction loop on the device: Poisson photon counts from line
// integrals, the log attenuation of sinogram/log/computeSimplePoissonSinogramAttunation.m
// (-log(max(I, dummyLowPhotonCount) / I0), :14-16), and the PWLS weights of
// statsrecon/pwls/computeWeightsPwls.m:8-9 (w = (s - r)^2 / (s + v), 0 where s <= r).
//
// Random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC 2011), counter = (element, attempt), key = seed, so every
// element's draw is independent of the launch shape.  Poisson variates: inversion by
// sequential search for lambda < 10, and PTRS, the transformed rejection with squeeze of
// W. Hoermann, "The transformed rejection method for generating Poisson random
// variables", Insurance: Mathematics and Economics 12 (1993), for lambda >= 10.
#include "internal.h"

namespace ctsim {

struct U4 {
    unsigned x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, unsigned k0, unsigned k1) {
    const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long)M0 * c.x;
        const unsigned long long p1 = (unsigned long long)M1 * c.z;
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// uniform in (0, 1) from 32 random bits (53-bit double not needed here)
__device__ __forceinline__ double u01(unsigned b) { return ((double)b + 0.5) * 2.3283064365386963e-10; }

// one Poisson variate with mean lam for element i; attempts draw fresh counters
__device__ double poisson_sample(double lam, unsigned long long i, unsigned k0, unsigned k1) {
    if (!(lam > 0.0)) return 0.0;
    const unsigned lo = (unsigned)i, hi = (unsigned)(i >> 32);
    if (lam < 10.0) {
        const U4 r = philox4x32_10(U4{lo, hi, 0u, 0u}, k0, k1);
        const double u = u01(r.x);
        double p = exp(-lam), f = p;
        int k = 0;
        while (u > f && k < 1000) {
            ++k;
            p *= lam / k;
            f += p;
        }
        return (double)k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (unsigned att = 0; att < 1024; ++att) {  // expected about 1.1 attempts
        const U4 r = philox4x32_10(U4{lo, hi, att, 1u}, k0, k1);
        const double U = u01(r.x) - 0.5, V = u01(r.y);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
    return floor(lam + 0.5);  // not reached in practice (probability < 1e-300)
}

template <class P>
__global__ void poisson_kernel(const P* __restrict__ p, size_t n, double i0, unsigned k0, unsigned k1,
                               float* __restrict__ counts) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        counts[i] = (float)poisson_sample(i0 * exp(-(double)ld<double>(p + i)), i, k0, k1);
}

template <class Y>
__global__ void neglog_kernel(const float* __restrict__ counts, size_t n, double i0, double low, Y* __restrict__ y) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double c = (double)counts[i];
        y[i] = cvt_store<Y>(-log((c > low ? c : low) / i0));
    }
}

// General form of the raw-data log (sinogram/log/processCTRawData.m:14-33 with
// computeSimplePoissonSinogramAttunation.m:14-52 or computeCompoundPoissonSinogramAttunation.m
// :14-38): element i of view i / ncell and detector cell i % ncell,
//   c = counts / profile[view]                        (tube-current profile, optional)
//   simple:   y = -log((max(c, low) / i0) / ff[cell])  (flat-field ratio of the bowtie, optional)
//   compound: y = -log(max((c / i0) / ff[cell], low / i0))
template <class Y>
__global__ void neglog_ex_kernel(const float* __restrict__ counts, size_t n, int ncell, double i0, double low,
                                 int compound, const float* __restrict__ ff, const float* __restrict__ profile,
                                 Y* __restrict__ y) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double c = (double)counts[i];
        if (profile) c = c / (double)profile[i / (size_t)ncell];
        double r;
        if (compound) {
            r = c / i0;
            if (ff) r = r / (double)ff[i % (size_t)ncell];
            const double rl = low / i0;
            r = r > rl ? r : rl;
        } else {
            r = (c > low ? c : low) / i0;
            if (ff) r = r / (double)ff[i % (size_t)ncell];
        }
        y[i] = cvt_store<Y>(-log(r));
    }
}

template <class W>
__global__ void weights_kernel(const float* __restrict__ s, size_t n, double r, double v, W* __restrict__ w) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float si = s[i];
        const float d = si - (float)r;
        const float val = d * d / (si + (float)v);  // single precision, as the reference's single()
        w[i] = cvt_store<W>(si <= (float)r ? 0.f : val);
    }
}

// Water beam-hardening correction (artifact/beamhardening/beamHardeningWarterCorrection.m:
// 40-69): out = sum_k c_k(u) p^k, k = 1 .. order, evaluated as the reference writes it
// (p c_1, then + p^k c_k in turn, single precision), with the coefficients of detector
// channel u (the sinogram's column, [view][u][v] order; one set for all channels when
// per_channel is 0).  coeffs[u * order + k - 1] = c_k(u).  In place is allowed.
template <class Tin, class Tout>
__global__ void water_poly_kernel(const Tin* in, Tout* out, size_t n, int nv, int nu,
                                  const float* __restrict__ coeffs, int order, int per_channel) {
#pragma clang fp contract(off)
    const size_t cells = (size_t)nu * nv;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int u = per_channel ? (int)((i % cells) / (size_t)nv) : 0;
        const float* c = coeffs + (size_t)u * order;
        const float p = ld_f(in + i);
        float acc = p * c[0];
        float pk = p;
        for (int k = 1; k < order; ++k) {
            pk = pk * p;
            acc = acc + pk * c[k];
        }
        out[i] = cvt_store<Tout>(acc);
    }
}

static unsigned grid_sim(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}

hipError_t poisson_launch(const void* p, int pdt, size_t n, double i0, unsigned long long seed, float* counts,
                          hipStream_t st) {
    const unsigned k0 = (unsigned)seed, k1 = (unsigned)(seed >> 32);
    switch (pdt) {
        case DT_F32: hipLaunchKernelGGL(poisson_kernel<float>, dim3(grid_sim(n)), dim3(256), 0, st, (const float*)p, n, i0, k0, k1, counts); break;
        case DT_F64: hipLaunchKernelGGL(poisson_kernel<double>, dim3(grid_sim(n)), dim3(256), 0, st, (const double*)p, n, i0, k0, k1, counts); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t neglog_launch(const float* counts, size_t n, double i0, double low, void* y, int ydt, hipStream_t st) {
    switch (ydt) {
        case DT_F32: hipLaunchKernelGGL(neglog_kernel<float>, dim3(grid_sim(n)), dim3(256), 0, st, counts, n, i0, low, (float*)y); break;
        case DT_F16: hipLaunchKernelGGL(neglog_kernel<__half>, dim3(grid_sim(n)), dim3(256), 0, st, counts, n, i0, low, (__half*)y); break;
        case DT_BF16: hipLaunchKernelGGL(neglog_kernel<__hip_bfloat16>, dim3(grid_sim(n)), dim3(256), 0, st, counts, n, i0, low, (__hip_bfloat16*)y); break;
        case DT_F64: hipLaunchKernelGGL(neglog_kernel<double>, dim3(grid_sim(n)), dim3(256), 0, st, counts, n, i0, low, (double*)y); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t neglog_ex_launch(const float* counts, size_t n, int ncell, double i0, double low, int compound,
                            const float* ff, const float* profile, void* y, int ydt, hipStream_t st) {
    const dim3 g(grid_sim(n)), b(256);
    switch (ydt) {
        case DT_F32: hipLaunchKernelGGL(neglog_ex_kernel<float>, g, b, 0, st, counts, n, ncell, i0, low, compound, ff, profile, (float*)y); break;
        case DT_F16: hipLaunchKernelGGL(neglog_ex_kernel<__half>, g, b, 0, st, counts, n, ncell, i0, low, compound, ff, profile, (__half*)y); break;
        case DT_BF16: hipLaunchKernelGGL(neglog_ex_kernel<__hip_bfloat16>, g, b, 0, st, counts, n, ncell, i0, low, compound, ff, profile, (__hip_bfloat16*)y); break;
        case DT_F64: hipLaunchKernelGGL(neglog_ex_kernel<double>, g, b, 0, st, counts, n, ncell, i0, low, compound, ff, profile, (double*)y); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class Tin>
static hipError_t water_launch_in(const void* in, void* out, int odt, size_t n, int nv, int nu, const float* c,
                                  int order, int per_channel, hipStream_t st) {
    const dim3 g(grid_sim(n)), b(256);
    switch (odt) {
        case DT_F32: hipLaunchKernelGGL((water_poly_kernel<Tin, float>), g, b, 0, st, (const Tin*)in, (float*)out, n, nv, nu, c, order, per_channel); break;
        case DT_F16: hipLaunchKernelGGL((water_poly_kernel<Tin, __half>), g, b, 0, st, (const Tin*)in, (__half*)out, n, nv, nu, c, order, per_channel); break;
        case DT_BF16: hipLaunchKernelGGL((water_poly_kernel<Tin, __hip_bfloat16>), g, b, 0, st, (const Tin*)in, (__hip_bfloat16*)out, n, nv, nu, c, order, per_channel); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t water_launch(const void* in, int idt, void* out, int odt, size_t n, int nv, int nu, const float* coeffs,
                        int order, int per_channel, hipStream_t st) {
    switch (idt) {
        case DT_F32: return water_launch_in<float>(in, out, odt, n, nv, nu, coeffs, order, per_channel, st);
        case DT_F16: return water_launch_in<__half>(in, out, odt, n, nv, nu, coeffs, order, per_channel, st);
        case DT_BF16: return water_launch_in<__hip_bfloat16>(in, out, odt, n, nv, nu, coeffs, order, per_channel, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t weights_launch(const float* s, size_t n, double r, double v, void* w, int wdt, hipStream_t st) {
    switch (wdt) {
        case DT_F32: hipLaunchKernelGGL(weights_kernel<float>, dim3(grid_sim(n)), dim3(256), 0, st, s, n, r, v, (float*)w); break;
        case DT_F16: hipLaunchKernelGGL(weights_kernel<__half>, dim3(grid_sim(n)), dim3(256), 0, st, s, n, r, v, (__half*)w); break;
        case DT_BF16: hipLaunchKernelGGL(weights_kernel<__hip_bfloat16>, dim3(grid_sim(n)), dim3(256), 0, st, s, n, r, v, (__hip_bfloat16*)w); break;
        case DT_F64: hipLaunchKernelGGL(weights_kernel<double>, dim3(grid_sim(n)), dim3(256), 0, st, s, n, r, v, (double*)w); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ctsim
