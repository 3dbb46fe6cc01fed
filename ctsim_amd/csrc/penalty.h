// Penalty neighbourhood sums of the PWLS updates (penalty/huberPenalty.m:101-161 and the
// other penalties of penalty/loadPenaltyOperator.m:22-54), shared by the update kernels and
// the backprojector's fused update.
#pragma once

#include "geom.h"

namespace ctsim {

enum Pot { POT_HUBER = 0, POT_QUAD = 1, POT_HYPER = 2, POT_ANISO = 3, POT_ISOHUBER = 4, POT_TV = 5, POT_ISOTV = 6 };

template <int POT, class V>
__device__ __forceinline__ V dpot(V d, V delta) {
    if (POT == POT_HUBER || POT == POT_ISOHUBER) return d > delta ? delta : (d < -delta ? -delta : d);
    if (POT == POT_QUAD) return d;
    if (POT == POT_ANISO) return d * exp(-(d * d) / (delta * delta));
    const V q = d / delta;
    return d / sqrt(V(1) + q * q);
}
template <int POT, class V>
__device__ __forceinline__ V cpot(V d, V delta) {
    if (POT == POT_HUBER || POT == POT_ISOHUBER) return fabs_t(d) <= delta ? V(1) : V(0);
    if (POT == POT_ANISO) return exp(-(d * d) / (delta * delta));
    const V q = d / delta;
    const V r = sqrt(V(1) + q * q);
    return V(1) / (r * r * r);
}
template <int POT>
__device__ __forceinline__ double cost_pot(double d, double delta) {
    if (POT == POT_HUBER || POT == POT_ISOHUBER) {
        d = fabs(d);
        return d > delta ? delta * d - delta * delta / 2 : d * d / 2;
    }
    if (POT == POT_QUAD) return d * d / 2;
    if (POT == POT_ANISO) return d * d * exp(-(d * d) / (delta * delta)) / 2;
    const double q = d / delta;
    return delta * delta * (sqrt(1.0 + q * q) - 1.0);
}

constexpr double kTvEps = 1e-8;

// total-variation cell k = (a, b, c) of the cropped grid: differences and c_k
template <bool ISO, class V>
__device__ __forceinline__ V tv_cell(const V* __restrict__ v, int nx, int nz, int a, int b, int c, V& d1, V& d2,
                                     V& d3) {
    const size_t sx = nz, sy = (size_t)nx * nz;
    const size_t k = (size_t)b * sy + (size_t)a * sx + c;
    const V xk = v[k];
    d1 = xk - v[k + sx];
    d2 = xk - v[k + sy];
    d3 = (ISO && nz > 1) ? xk - v[k + 1] : V(0);
    return V(1) / sqrt(d1 * d1 + d2 * d2 + d3 * d3 + V(kTvEps));
}

// S (MODE 1) or T (MODE 2) of the total variation at voxel (ix, iy, iz): the voxel's own
// cell and the cells one step below it along each direction
template <int MODE, bool ISO, class V>
__device__ __forceinline__ V tv_point(const V* __restrict__ v, int nx, int ny, int nz, int ix, int iy, int iz) {
    const bool three = nz > 1;
    const bool zdir = ISO && three;
    auto inside = [&](int a, int b, int c) {
        return a >= 0 && b >= 0 && c >= 0 && a <= nx - 2 && b <= ny - 2 && (!three || c <= nz - 2);
    };
    V acc = V(0), d1, d2, d3;
    if (inside(ix, iy, iz)) {
        const V c = tv_cell<ISO>(v, nx, nz, ix, iy, iz, d1, d2, d3);
        acc += MODE == 1 ? d1 * c + d2 * c + d3 * c : V(zdir ? 3 : 2) * c;
    }
    if (inside(ix - 1, iy, iz)) {
        const V c = tv_cell<ISO>(v, nx, nz, ix - 1, iy, iz, d1, d2, d3);
        acc += MODE == 1 ? -d1 * c : c;
    }
    if (inside(ix, iy - 1, iz)) {
        const V c = tv_cell<ISO>(v, nx, nz, ix, iy - 1, iz, d1, d2, d3);
        acc += MODE == 1 ? -d2 * c : c;
    }
    if (zdir && inside(ix, iy, iz - 1)) {
        const V c = tv_cell<ISO>(v, nx, nz, ix, iy, iz - 1, d1, d2, d3);
        acc += MODE == 1 ? -d3 * c : c;
    }
    return acc;
}

// S (MODE 1), T (MODE 2) or, for the Huber penalty, H (MODE 3) of voxel (ix, iy, iz) of
// volume v (ZXY)
template <int MODE, int POT, class V>
__device__ __forceinline__ V pen_point(const V* __restrict__ v, int nx, int ny, int nz, int ix, int iy, int iz,
                                       V delta) {
    static_assert(MODE != 3 || POT == POT_HUBER, "mode 3 is the Huber curvature H");
    if (POT == POT_TV || POT == POT_ISOTV) return tv_point<MODE == 1 ? 1 : 2, POT == POT_ISOTV>(v, nx, ny, nz, ix, iy, iz);
    const V axy = V(0.70710678118654752440);  // 1/sqrt(2)
    if (MODE == 2 && POT == POT_QUAD) return nz > 1 ? V(4) + V(4) * axy + V(1) : V(4) + V(4) * axy;
    const size_t sx = nz, sy = (size_t)nx * nz;
    const size_t i = (size_t)iy * sy + (size_t)ix * sx + iz;
    const V xi = v[i];
    V acc = V(0);
    auto term = [&](bool inside, size_t j, V w) {
        if (!inside) return;
        const V xj = v[j];
        const V d = xi - xj;
        if (MODE == 1) acc += w * dpot<POT>(d, delta);
        else if (MODE == 2) acc += w * cpot<POT>(d, delta);
        else acc += w * cpot<POT>(d, delta) * (xi + xj);
    };
    if (POT == POT_ISOHUBER) {
        term(ix + 1 < nx, i + sx, V(1));
        term(ix > 0, i - sx, V(1));
        term(iy + 1 < ny, i + sy, V(1));
        term(iy > 0, i - sy, V(1));
        term(iz + 1 < nz, i + 1, V(1));
        term(iz > 0, i - 1, V(1));
        return acc;
    }
    // x neighbours
    term(ix + 1 < nx, i + sx, V(1));
    term(ix > 0, i - sx, V(1));
    // y neighbours
    term(iy + 1 < ny, i + sy, V(1));
    term(iy > 0, i - sy, V(1));
    // diagonal (+1,+1) / (-1,-1)
    term(ix + 1 < nx && iy + 1 < ny, i + sx + sy, axy);
    term(ix > 0 && iy > 0, i - sx - sy, axy);
    // anti-diagonal (-1,+1) / (+1,-1)
    term(ix > 0 && iy + 1 < ny, i - sx + sy, axy);
    term(ix + 1 < nx && iy > 0, i + sx - sy, axy);
    // z neighbours
    term(iz + 1 < nz, i + 1, V(0.5));
    term(iz > 0, i - 1, V(0.5));
    return MODE == 3 ? acc / xi : acc;
}

// NaN, Inf and negative values to zero (the updates' x(x < 0 | ~isfinite(x)) = 0)
template <class V>
__device__ __forceinline__ V scrub(V v) { return (v < V(0) || !isfinite(v)) ? V(0) : v; }

}  // namespace ctsim
