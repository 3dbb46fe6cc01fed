// CPU oracle for the CTSim projector / backprojector path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker or the timed CPU
// baseline.  The product (ctsim_amd/) never links, loads or calls it.
//
// It restates the Linux CPU MEX kernels of CTSim (geom/fbct_os, geom/cbct_os) as
// voxel-driven scatter loops with the same loop order, clamps and rounding points:
//
//   T = float  : "fp32_faithful".  Every expression the reference evaluates in
//                double (the MATLAB-era `0.5`, `1.0`, `1E-6` literals promote) is
//                evaluated in double here too; everything else is float.  On
//                x86-64 without FMA contraction this is bit-identical to the
//                reference compiled by g++ -O2 (checked against oracle/_ref).
//   T = double : "fp64".  The same formulas evaluated in double throughout, with
//                double image / sinogram storage: the accuracy oracle.
//
// Reference functions restated (file:line):
//   fbct_proj_dd   geom/fbct_os/fbct_proj.c:155-291
//   fbct_proj_tf   geom/fbct_os/fbct_proj.c:294-466
//   fbct_back_pd   geom/fbct_os/fbct_back.c:3-80
//   fbct_back_dd   geom/fbct_os/fbct_back.c:82-225
//   fbct_back_tf   geom/fbct_os/fbct_back.c:227-399
//   cbct_proj1_dd  geom/cbct_os/cbct_proj.c:427-642   (cbct_proj_dd :208-418 = map all true)
//   cbct_proj1_tf  geom/cbct_os/cbct_proj.c:912-1169  (cbct_proj_tf :652-903)
//   cbct_back1_pd  geom/cbct_os/cbct_back.c:135-262   (cbct_back_pd :9-132; pd/hpd flag)
//   cbct_back1_dd  geom/cbct_os/cbct_back.c:479-692   (cbct_back_dd :267-475)
//   cbct_back1_tf  geom/cbct_os/cbct_back.c:945-1197  (cbct_back_tf :695-942)
//   fbct_proj_rd   geom/fbct_os/fbct_proj.c:3-153
//   cbct_proj_rd   geom/cbct_os/cbct_proj.c:10-198
//   op dispatch    geom/fbct_os/fbct_geom_mex.c:271-305, geom/cbct_os/cbct_geom_mex.c:527-583
//
// Layouts are MATLAB column-major: image [nx ny (nz)] with x fastest, 2D
// sinogram [nu nviews], 3D sinogram [nv nu nviews] with v fastest.
// Outputs must be zero-filled by the caller (the MEX gateway's
// mxCreateNumericArray zero-fills).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

struct Geom2 {
    int nx, ny;
    float dx, dy, ox, oy;
    int nu;
    float du, ou;
    float sad, add, sdd;
    int nviews;
    int dettype;
};

struct Geom3 {
    int nx, ny, nz;
    float dx, dy, dz, ox, oy, oz;
    int nu, nv;
    float du, dv, ou, ov;
    float sad, add, sdd;
    int nviews;
    int dettype;
};

// The reference's max/min macros: ((a) > (b)) ? (a) : (b), ((a) < (b)) ? (a) : (b).
template <class A>
inline A mx(A a, A b) { return (a > b) ? a : b; }
template <class A>
inline A mn(A a, A b) { return (a < b) ? a : b; }

// Per-view trigonometry: sin/cos evaluated in double and rounded to T, the
// half-pixel projections dx*sin/2 etc. (geom/cbct_os/cbct_proj.c:467-474).
template <class T>
struct ViewTrig {
    T s, c, dxs2, dxc2, dys2, ndyc2;
    ViewTrig(double beta, T dx, T dy) {
        s = (T)std::sin(beta);
        c = (T)std::cos(beta);
        dxs2 = dx * s / T(2);
        dxc2 = dx * c / T(2);
        dys2 = dy * s / T(2);
        ndyc2 = -dy * c / T(2);
    }
};

// In-plane quantities of the column whose centre is (xc, yc):
// t (detector-parallel coordinate), d (source distance along the central ray),
// r = |(t,d)|, and the amplitude dx / max(|cos phi|, |sin phi|) / 10.
template <class T>
struct ColumnGeom {
    T t, d, r, amp;
    bool xdom;  // cos(phi) > sin(phi): footprint from the x-edges
    ColumnGeom(T xc, T yc, const ViewTrig<T>& v, T sad, T dx) {
        t = xc * v.c - yc * v.s;
        d = sad + xc * v.s + yc * v.c;
        r = std::sqrt(t * t + d * d);
        T cg = d / r;
        T sg = t / r;
        T cphi = std::fabs(v.c * cg - v.s * sg);
        T sphi = std::fabs(v.s * cg + v.c * sg);
        xdom = cphi > sphi;
        amp = xdom ? dx / cphi / T(10) : dx / sphi / T(10);
    }
};

// Distance-driven u boundaries, before the +0.5 shift (fbct_proj.c:216-248).
template <class T>
inline void dd_u_bounds(const ColumnGeom<T>& cg, const ViewTrig<T>& v, T sdd_du, T wu,
                        int dettype, T& umin, T& umax) {
    T a0, b0, a1, b1;
    if (cg.xdom) {
        a0 = cg.t + v.dxc2; b0 = cg.d + v.dxs2;
        a1 = cg.t - v.dxc2; b1 = cg.d - v.dxs2;
    } else {
        a0 = cg.t + v.dys2; b0 = cg.d + v.ndyc2;
        a1 = cg.t - v.dys2; b1 = cg.d - v.ndyc2;
    }
    if (dettype == 1) {
        umin = sdd_du * a0 / b0 - wu;
        umax = sdd_du * a1 / b1 - wu;
    } else {
        umin = sdd_du * std::atan(a0 / b0) - wu;
        umax = sdd_du * std::atan(a1 / b1) - wu;
    }
    if (umin >= umax) std::swap(umin, umax);
}

// Separable-footprint trapezoid corners tau[0..3], sorted (fbct_proj.c:365-412).
template <class T>
inline void sf_taus(const ColumnGeom<T>& cg, const ViewTrig<T>& v, T sdd_du, T wu, int dettype,
                    T tau[4]) {
    T n0 = cg.t + v.dxc2 + v.dys2, q0 = cg.d + v.dxs2 + v.ndyc2;
    T n1 = cg.t - v.dxc2 + v.dys2, q1 = cg.d - v.dxs2 + v.ndyc2;
    T n2 = cg.t + v.dxc2 - v.dys2, q2 = cg.d + v.dxs2 - v.ndyc2;
    T n3 = cg.t - v.dxc2 - v.dys2, q3 = cg.d - v.dxs2 - v.ndyc2;
    if (dettype == 1) {
        tau[0] = sdd_du * n0 / q0 - wu;
        tau[1] = sdd_du * n1 / q1 - wu;
        tau[2] = sdd_du * n2 / q2 - wu;
        tau[3] = sdd_du * n3 / q3 - wu;
    } else {
        tau[0] = sdd_du * std::atan(n0 / q0) - wu;
        tau[1] = sdd_du * std::atan(n1 / q1) - wu;
        tau[2] = sdd_du * std::atan(n2 / q2) - wu;
        tau[3] = sdd_du * std::atan(n3 / q3) - wu;
    }
    // five compare-exchanges, strict '>' as in the reference
    if (tau[0] > tau[1]) std::swap(tau[0], tau[1]);
    if (tau[2] > tau[3]) std::swap(tau[2], tau[3]);
    if (tau[0] > tau[2]) std::swap(tau[0], tau[2]);
    if (tau[1] > tau[3]) std::swap(tau[1], tau[3]);
    if (tau[1] > tau[2]) std::swap(tau[1], tau[2]);
}

// Trapezoid integral over the cell [iu-1/2, iu+1/2] (fbct_proj.c:428-453).
template <class T>
inline T sf_weight(int iu, const T tau[4]) {
    T w = 0;
    T bl = (T)(iu - 0.5);
    T br = (T)(iu + 0.5);
    T xl = mx(bl, tau[0]), xr = mn(br, tau[1]);
    if (xr > xl)
        w = w + (((xr - tau[0]) * (xr - tau[0]) - (xl - tau[0]) * (xl - tau[0])) / (tau[1] - tau[0])) / 2;
    xl = mx(bl, tau[1]);
    xr = mn(br, tau[2]);
    if (xr > xl) w = w + (xr - xl);
    xl = mx(bl, tau[2]);
    xr = mn(br, tau[3]);
    if (xr > xl)
        w = w + (((xl - tau[3]) * (xl - tau[3]) - (xr - tau[3]) * (xr - tau[3])) / (tau[3] - tau[2])) / 2;
    return w;
}

// Detector-edge clamp of a [lo, hi] footprint after the +0.5 shift.
// Returns false when the footprint misses the detector (lo > hi).
template <class T>
inline bool clamp_cells(T lo, T hi, int n, T& clo, T& chi, int& ilo, int& ihi, int cap) {
    clo = mx(lo + T(0.5), T(0));
    chi = mn(hi + T(0.5), (T)n);
    if (clo > chi) return false;
    ilo = (int)std::floor(clo);
    ihi = (int)std::floor(chi);
    ihi = std::min(ihi, n - 1);
    if (cap > 0) ihi = std::min(ihi, ilo + cap - 1);
    return true;
}

// DD box weights of a clamped footprint (cbct_proj.c:553-567); capped footprints
// give the last cell (chi - ihi), which can exceed 1, as the reference does.
template <class T>
inline void dd_weights(T clo, T chi, int ilo, int ihi, T* w) {
    if (ilo == ihi) {
        w[0] = chi - clo;
        return;
    }
    for (int i = ilo; i <= ihi; ++i) {
        if (i == ilo)
            w[0] = (T)(ilo + 1.0 - clo);
        else if (i == ihi)
            w[i - ilo] = (T)(chi - ihi);
        else
            w[i - ilo] = T(1);
    }
}

inline int nthreads_or_default(int nthreads) { return nthreads > 0 ? nthreads : 1; }

// ------------------------------------------------------------------------------------
// 2D fan beam
// ------------------------------------------------------------------------------------

template <class T>
void fbct_proj_dd(const Geom2& g, const double* betas, const T* img, T* proj, int nthreads) {
    const T dx = g.dx, dy = g.dy, sad = g.sad;
    const T wx = (T)(-(g.nx - 1.0) / 2.0 + g.ox);
    const T wy = (T)(-(g.ny - 1.0) / 2.0 + g.oy);
    const T wu = (T)(-(g.nu - 1.0) / 2.0 + g.ou);
    const T sdd_du = (T)g.sdd / (T)g.du;
    const int nu = g.nu;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int ib = 0; ib < g.nviews; ++ib) {
        ViewTrig<T> v(betas[ib], dx, dy);
        T* row = proj + (size_t)ib * nu;
        for (int iy = 0; iy < g.ny; ++iy) {
            for (int ix = 0; ix < g.nx; ++ix) {
                T xc = (ix + wx) * dx, yc = (iy + wy) * dy;
                ColumnGeom<T> cg(xc, yc, v, sad, dx);
                T umin, umax;
                dd_u_bounds(cg, v, sdd_du, wu, g.dettype, umin, umax);
                T clo, chi;
                int ilo, ihi;
                if (!clamp_cells(umin, umax, nu, clo, chi, ilo, ihi, 0)) continue;
                const T px = img[iy * g.nx + ix];
                if (ilo == ihi) {
                    row[ilo] = row[ilo] + cg.amp * (chi - clo) * px;
                    continue;
                }
                for (int iu = ilo; iu <= ihi; ++iu) {
                    if (iu == ilo)  // weight kept in double, as in the reference
                        row[iu] = (T)((double)row[iu] + (double)cg.amp * (ilo + 1.0 - (double)clo) * (double)px);
                    else if (iu == ihi)
                        row[iu] = row[iu] + cg.amp * (chi - ihi) * px;
                    else
                        row[iu] = row[iu] + cg.amp * px;
                }
            }
        }
    }
}

template <class T>
void fbct_back_dd(const Geom2& g, const double* betas, T* img, const T* proj, int nthreads) {
    const T dx = g.dx, dy = g.dy, sad = g.sad;
    const T wx = (T)(-(g.nx - 1.0) / 2.0 + g.ox);
    const T wy = (T)(-(g.ny - 1.0) / 2.0 + g.oy);
    const T wu = (T)(-(g.nu - 1.0) / 2.0 + g.ou);
    const T sdd_du = (T)g.sdd / (T)g.du;
    const int nu = g.nu;
    // parallel over image rows; every pixel still sums its views in ascending order
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int iy = 0; iy < g.ny; ++iy) {
        for (int ib = 0; ib < g.nviews; ++ib) {
            ViewTrig<T> v(betas[ib], dx, dy);
            const T* row = proj + (size_t)ib * nu;
            for (int ix = 0; ix < g.nx; ++ix) {
                T xc = (ix + wx) * dx, yc = (iy + wy) * dy;
                ColumnGeom<T> cg(xc, yc, v, sad, dx);
                T umin, umax;
                dd_u_bounds(cg, v, sdd_du, wu, g.dettype, umin, umax);
                T clo, chi;
                int ilo, ihi;
                if (!clamp_cells(umin, umax, nu, clo, chi, ilo, ihi, 0)) continue;
                T acc = 0;
                if (ilo == ihi) {
                    acc = (chi - clo) * row[ilo];
                } else {
                    for (int iu = ilo; iu <= ihi; ++iu) {
                        if (iu == ilo)
                            acc = (T)((double)acc + (ilo + 1.0 - (double)clo) * (double)row[iu]);
                        else if (iu == ihi)
                            acc = acc + (chi - ihi) * row[iu];
                        else
                            acc = acc + row[iu];
                    }
                }
                img[iy * g.nx + ix] = img[iy * g.nx + ix] + cg.amp * acc;
            }
        }
    }
}

template <class T>
void fbct_proj_tf(const Geom2& g, const double* betas, const T* img, T* proj, int nthreads) {
    const T dx = g.dx, dy = g.dy, sad = g.sad;
    const T wx = (T)(-(g.nx - 1.0) / 2.0 + g.ox);
    const T wy = (T)(-(g.ny - 1.0) / 2.0 + g.oy);
    const T wu = (T)(-(g.nu - 1.0) / 2.0 + g.ou);
    const T sdd_du = (T)g.sdd / (T)g.du;
    const int nu = g.nu;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int ib = 0; ib < g.nviews; ++ib) {
        ViewTrig<T> v(betas[ib], dx, dy);
        T* row = proj + (size_t)ib * nu;
        for (int iy = 0; iy < g.ny; ++iy) {
            for (int ix = 0; ix < g.nx; ++ix) {
                T xc = (ix + wx) * dx, yc = (iy + wy) * dy;
                ColumnGeom<T> cg(xc, yc, v, sad, dx);
                T tau[4];
                sf_taus(cg, v, sdd_du, wu, g.dettype, tau);
                T clo, chi;
                int ilo, ihi;
                if (!clamp_cells(tau[0], tau[3], nu, clo, chi, ilo, ihi, 0)) continue;
                const T val = cg.amp * img[iy * g.nx + ix];
                for (int iu = ilo; iu <= ihi; ++iu) row[iu] = row[iu] + sf_weight(iu, tau) * val;
            }
        }
    }
}

template <class T>
void fbct_back_tf(const Geom2& g, const double* betas, T* img, const T* proj, int nthreads) {
    const T dx = g.dx, dy = g.dy, sad = g.sad;
    const T wx = (T)(-(g.nx - 1.0) / 2.0 + g.ox);
    const T wy = (T)(-(g.ny - 1.0) / 2.0 + g.oy);
    const T wu = (T)(-(g.nu - 1.0) / 2.0 + g.ou);
    const T sdd_du = (T)g.sdd / (T)g.du;
    const int nu = g.nu;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int iy = 0; iy < g.ny; ++iy) {
        for (int ib = 0; ib < g.nviews; ++ib) {
            ViewTrig<T> v(betas[ib], dx, dy);
            const T* row = proj + (size_t)ib * nu;
            for (int ix = 0; ix < g.nx; ++ix) {
                T xc = (ix + wx) * dx, yc = (iy + wy) * dy;
                ColumnGeom<T> cg(xc, yc, v, sad, dx);
                T tau[4];
                sf_taus(cg, v, sdd_du, wu, g.dettype, tau);
                // the 2D back pass bounds the cells with floor/ceil (fbct_back.c:347-352)
                int ilo = (int)mx(std::floor(tau[0] + T(0.5)), T(0));
                int ihi = (int)mn(std::ceil(tau[3] + T(0.5)), (T)(nu - 1));
                if (ilo > ihi) continue;
                T acc = 0;
                for (int iu = ilo; iu <= ihi; ++iu) acc = acc + cg.amp * sf_weight(iu, tau) * row[iu];
                img[iy * g.nx + ix] = img[iy * g.nx + ix] + acc;
            }
        }
    }
}

template <class T>
void fbct_back_pd(const Geom2& g, const double* betas, T* img, const T* proj, int nthreads) {
    const T dx = g.dx, dy = g.dy, sad = g.sad, sdd = g.sdd, du = g.du;
    const T wx = (T)(-(g.nx - 1.0) / 2.0 + g.ox);
    const T wy = (T)(-(g.ny - 1.0) / 2.0 + g.oy);
    const T wu = (T)(-(g.nu - 1.0) / 2.0 + g.ou);
    const int nu = g.nu;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int iy = 0; iy < g.ny; ++iy) {
        for (int ib = 0; ib < g.nviews; ++ib) {
            const T s = (T)std::sin(betas[ib]), c = (T)std::cos(betas[ib]);
            const T* row = proj + (size_t)ib * nu;
            for (int ix = 0; ix < g.nx; ++ix) {
                T xc = (ix + wx) * dx, yc = (iy + wy) * dy;
                T xr = xc * c - yc * s;
                T yr = xc * s + yc * c;
                T u, w2s;
                if (g.dettype == 1) {
                    T w2 = sdd / (sad + yr);
                    w2s = w2 * w2;
                    u = xr * w2 / du - wu;
                } else {
                    T roff = (T)g.ou * sad / sdd;
                    u = std::atan((xr - roff) / (sad + yr)) * sdd / du - wu;
                    w2s = sdd * sdd / ((sad + yr) * (sad + yr) + (xr - roff) * (xr - roff));
                }
                if ((double)u < 1e-6 || (double)u > nu - 1 - 1e-6) continue;
                int iu = (int)std::floor(u);
                T wl = u - iu;
                img[iy * g.nx + ix] = img[iy * g.nx + ix] + w2s * ((1 - wl) * row[iu] + wl * row[iu + 1]);
            }
        }
    }
}

// ------------------------------------------------------------------------------------
// 3D cone beam (map family; map == nullptr means every column is active)
// ------------------------------------------------------------------------------------

template <class T>
struct Cone {
    const Geom3& g;
    T wx, wy, wu, wv, sdd_du;
    explicit Cone(const Geom3& gg) : g(gg) {
        wx = (T)(-(g.nx - 1.0) / 2.0 + g.ox);
        wy = (T)(-(g.ny - 1.0) / 2.0 + g.oy);
        wu = (T)(-(g.nu - 1.0) / 2.0 + g.ou);
        wv = (T)(-(g.nv - 1.0) / 2.0 + g.ov);
        sdd_du = (T)g.sdd / (T)g.du;
    }
    T wz(const float* couchz, int ib) const {
        return (T)(-(g.nz - 1.0) / 2.0 + (double)(T)g.oz + (double)((T)couchz[ib] / (T)g.dz));
    }
};

// v-direction column parameters: magnified slice pitch zinc, start vstart and the
// iz window [iz0, iz1) (cbct_proj.c:569-580).
template <class T>
inline bool v_window(const Geom3& g, T d, T wz, T wv, T& zinc, T& vstart, int& iz0, int& iz1) {
    zinc = (T)g.dz / (T)g.dv * (T)g.sdd / d;
    vstart = (T)(((double)wz - 0.5) * (double)zinc - (double)wv);
    iz0 = (int)std::floor((-1 - vstart) / zinc);
    iz1 = (int)std::ceil((g.nv + 1 - vstart) / zinc);
    iz0 = std::max(iz0, 0);
    iz1 = std::min(iz1, g.nz);
    return iz0 <= iz1;
}

enum FootKind { FOOT_DD = 0, FOOT_TF = 1 };

// Column footprint along u: first cell and up to 16 weights.
template <class T>
inline bool u_footprint(FootKind kind, const Geom3& g, const Cone<T>& k, const ColumnGeom<T>& cg,
                        const ViewTrig<T>& v, int& ilo, int& ihi, T* wu16) {
    T clo, chi;
    if (kind == FOOT_DD) {
        T umin, umax;
        dd_u_bounds(cg, v, k.sdd_du, k.wu, g.dettype, umin, umax);
        if (!clamp_cells(umin, umax, g.nu, clo, chi, ilo, ihi, 16)) return false;
        dd_weights(clo, chi, ilo, ihi, wu16);
    } else {
        T tau[4];
        sf_taus(cg, v, k.sdd_du, k.wu, g.dettype, tau);
        if (!clamp_cells(tau[0], tau[3], g.nu, clo, chi, ilo, ihi, 16)) return false;
        for (int iu = ilo; iu <= ihi; ++iu) wu16[iu - ilo] = sf_weight(iu, tau);
    }
    return true;
}

template <class T>
void cbct_proj(FootKind kind, const Geom3& g, const double* betas, const float* couchz, const T* img,
               const uint8_t* map, T* proj, int nthreads) {
    Cone<T> k(g);
    const T dx = g.dx, dy = g.dy, dz = g.dz, sad = g.sad;
    const size_t nxy = (size_t)g.nx * g.ny;
    const int nu = g.nu, nv = g.nv;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int ib = 0; ib < g.nviews; ++ib) {
        const T wz = k.wz(couchz, ib);
        ViewTrig<T> v(betas[ib], dx, dy);
        T* view = proj + (size_t)ib * nu * nv;
        T wu16[16], wv16[16];
        for (int iy = 0; iy < g.ny; ++iy) {
            for (int ix = 0; ix < g.nx; ++ix) {
                if (map && !map[iy * g.nx + ix]) continue;
                T xc = (ix + k.wx) * dx, yc = (iy + k.wy) * dy;
                ColumnGeom<T> cg(xc, yc, v, sad, dx);
                int iu0, iu1;
                if (!u_footprint(kind, g, k, cg, v, iu0, iu1, wu16)) continue;
                T zinc, vstart;
                int iz0, iz1;
                if (!v_window(g, cg.d, wz, k.wv, zinc, vstart, iz0, iz1)) continue;
                for (int iz = iz0; iz < iz1; ++iz) {
                    const T px = img[iz * nxy + (size_t)iy * g.nx + ix];
                    if (px == 0) continue;
                    T zc = (iz + wz) * dz;
                    T ampv = std::sqrt(cg.r * cg.r + zc * zc) / cg.r;
                    T val = cg.amp * ampv * px;
                    T vmin = vstart + iz * zinc;
                    T vmax = vmin + zinc;
                    T clo, chi;
                    int iv0, iv1;
                    if (!clamp_cells(vmin, vmax, nv, clo, chi, iv0, iv1, 16)) continue;
                    dd_weights(clo, chi, iv0, iv1, wv16);
                    for (int iu = iu0; iu <= iu1; ++iu) {
                        T* cell = view + (size_t)iu * nv;
                        for (int iv = iv0; iv <= iv1; ++iv)
                            cell[iv] = cell[iv] + wu16[iu - iu0] * wv16[iv - iv0] * val;
                    }
                }
            }
        }
    }
}

template <class T>
void cbct_back(FootKind kind, const Geom3& g, const double* betas, const float* couchz, T* img,
               const uint8_t* map, const T* proj, int nthreads) {
    Cone<T> k(g);
    const T dx = g.dx, dy = g.dy, dz = g.dz, sad = g.sad;
    const size_t nxy = (size_t)g.nx * g.ny;
    const int nu = g.nu, nv = g.nv;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int iy = 0; iy < g.ny; ++iy) {
        T wu16[16], wv16[16];
        for (int ib = 0; ib < g.nviews; ++ib) {
            const T wz = k.wz(couchz, ib);
            ViewTrig<T> v(betas[ib], dx, dy);
            const T* view = proj + (size_t)ib * nu * nv;
            for (int ix = 0; ix < g.nx; ++ix) {
                if (map && !map[iy * g.nx + ix]) continue;
                T xc = (ix + k.wx) * dx, yc = (iy + k.wy) * dy;
                ColumnGeom<T> cg(xc, yc, v, sad, dx);
                int iu0, iu1;
                if (!u_footprint(kind, g, k, cg, v, iu0, iu1, wu16)) continue;
                T zinc, vstart;
                int iz0, iz1;
                if (!v_window(g, cg.d, wz, k.wv, zinc, vstart, iz0, iz1)) continue;
                for (int iz = iz0; iz < iz1; ++iz) {
                    T zc = (iz + wz) * dz;
                    T ampv = std::sqrt(cg.r * cg.r + zc * zc) / cg.r;
                    T vmin = vstart + iz * zinc;
                    T vmax = vmin + zinc;
                    T clo, chi;
                    int iv0, iv1;
                    if (!clamp_cells(vmin, vmax, nv, clo, chi, iv0, iv1, 16)) continue;
                    dd_weights(clo, chi, iv0, iv1, wv16);
                    T acc = 0;
                    for (int iu = iu0; iu <= iu1; ++iu) {
                        const T* cell = view + (size_t)iu * nv;
                        for (int iv = iv0; iv <= iv1; ++iv)
                            acc = acc + cell[iv] * wu16[iu - iu0] * wv16[iv - iv0];
                    }
                    T& out = img[iz * nxy + (size_t)iy * g.nx + ix];
                    out = out + cg.amp * ampv * acc;
                }
            }
        }
    }
}

// Voxel-driven bilinear FDK backprojection; hpd clamps v into the detector,
// pd skips / stops at the detector edge (cbct_back.c:221-242).
template <class T>
void cbct_back_pd(bool off_boundary, const Geom3& g, const double* betas, const float* couchz, T* img,
                  const uint8_t* map, const T* proj, int nthreads) {
    Cone<T> k(g);
    const T dx = g.dx, dy = g.dy, dz = g.dz, dv = g.dv, du = g.du, sad = g.sad, sdd = g.sdd;
    const size_t nxy = (size_t)g.nx * g.ny;
    const int nu = g.nu, nv = g.nv;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int iy = 0; iy < g.ny; ++iy) {
        for (int ib = 0; ib < g.nviews; ++ib) {
            const T wz = k.wz(couchz, ib);
            const T s = (T)std::sin(betas[ib]), c = (T)std::cos(betas[ib]);
            const T* view = proj + (size_t)ib * nu * nv;
            for (int ix = 0; ix < g.nx; ++ix) {
                if (map && !map[iy * g.nx + ix]) continue;
                T xc = (ix + k.wx) * dx, yc = (iy + k.wy) * dy;
                T xr = xc * c - yc * s;
                T yr = xc * s + yc * c;
                T u, w2, w2s;
                if (g.dettype == 1) {
                    w2 = sdd / (sad + yr);
                    w2s = w2 * w2;
                    u = xr * w2 / du - k.wu;
                } else {
                    T roff = (T)g.ou * sad / sdd;
                    u = std::atan((xr - roff) / (sad + yr)) * sdd / du - k.wu;
                    w2s = sdd * sdd / ((sad + yr) * (sad + yr) + (xr - roff) * (xr - roff));
                    w2 = sdd / (sad + yr);
                }
                if ((double)u < 1e-6 || (double)u > nu - 1 - 1e-6) continue;
                int iu = (int)std::floor(u);
                T wul = u - iu;
                for (int iz = 0; iz < g.nz; ++iz) {
                    T zc = (iz + wz) * dz;
                    T vv = zc * w2 / dv - k.wv;
                    if ((double)vv < 1e-6) {
                        if (off_boundary) continue;
                        vv = (T)0.5;
                    }
                    if ((double)vv > nv - 1 - 1e-6) {
                        if (off_boundary) break;
                        vv = (T)(nv - 1.5);
                    }
                    int iv = (int)std::floor(vv);
                    T wvl = vv - iv;
                    const T* p = view + (size_t)iu * nv + iv;
                    T& out = img[iz * nxy + (size_t)iy * g.nx + ix];
                    out = out + w2s * ((1 - wul) * ((1 - wvl) * p[0] + wvl * p[1]) +
                                       wul * ((1 - wvl) * p[nv] + wvl * p[nv + 1]));
                }
            }
        }
    }
}

// Ray-driven forward projection (geom/fbct_os/fbct_proj.c:3-153, geom/cbct_os/cbct_proj.c:10-198):
// every detector cell's ray is marched from its entry into the image box in steps of dx/2,
// t accumulated by repeated addition as the reference does, sampling the image by
// bilinear (2D) / trilinear (3D) interpolation at interior positions.  The reference's
// test `raydir_x < -1E-12 && raydir_x > 1E-12` is never true, so the slab bounds are
// always formed by division (an axis-parallel ray gives +-inf, as there).  In the float
// form: sin(beta) and atan in double, sin(phi) and sqrt of floats in float (C++
// overloads), the box corners and dx/2 in double rounded to float.
template <class T>
struct RayRow {
    bool hit;
    T tn, tf, rdx, rdy, spd, sx, sy;
};

template <class T>
RayRow<T> ray_row(double beta, int iu, T du, T wu, T sad, T sdd, int dettype, T cmin_x, T cmax_x, T cmin_y,
                  T cmax_y) {
    RayRow<T> r{};
    const T sin_a = (T)std::sin(beta), cos_a = (T)std::cos(beta);
    r.sx = -sad * sin_a;
    r.sy = -sad * cos_a;
    const T u = iu * du + wu;
    T phi;
    if (dettype == 1) {
        r.spd = std::sqrt(sdd * sdd + u * u);
        phi = (T)(beta + std::atan((double)(u / sdd)));
    } else {
        r.spd = sdd;
        phi = (T)(beta + (double)(u / sdd));
    }
    r.rdx = (T)std::sin(phi);
    r.rdy = (T)std::cos(phi);
    T tn = (T)-1E10, tf = (T)1E10;
    {
        T t1 = (cmin_x - r.sx) / r.rdx, t2 = (cmax_x - r.sx) / r.rdx;
        if (t1 > t2) std::swap(t1, t2);
        if (t1 > tn) tn = t1;
        if (t2 < tf) tf = t2;
        if (tn > tf || tf < 0) return r;
    }
    {
        T t1 = (cmin_y - r.sy) / r.rdy, t2 = (cmax_y - r.sy) / r.rdy;
        if (t1 > t2) std::swap(t1, t2);
        if (t1 > tn) tn = t1;
        if (t2 < tf) tf = t2;
        if (tn > tf || tf < 0) return r;
    }
    r.hit = true;
    r.tn = tn;
    r.tf = tf;
    return r;
}

template <class T>
void fbct_proj_rd(const Geom2& g, const double* betas, const T* img, T* proj, int nthreads) {
    const T wu = (T)((-(g.nu - 1.0) / 2.0 + g.ou) * g.du);
    const T cmin_x = (T)((-(g.nx - 1.0) / 2.0 + g.ox) * g.dx), cmax_x = (T)((+(g.nx - 1.0) / 2.0 + g.ox) * g.dx);
    const T cmin_y = (T)((-(g.ny - 1.0) / 2.0 + g.oy) * g.dy), cmax_y = (T)((+(g.ny - 1.0) / 2.0 + g.oy) * g.dy);
    const T delta = (T)(g.dx / 2.0);
    const T dx = g.dx, dy = g.dy;
    const int nx = g.nx, ny = g.ny;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int ib = 0; ib < g.nviews; ++ib)
        for (int iu = 0; iu < g.nu; ++iu) {
            const RayRow<T> r = ray_row<T>(betas[ib], iu, (T)g.du, wu, (T)g.sad, (T)g.sdd, g.dettype, cmin_x,
                                           cmax_x, cmin_y, cmax_y);
            if (!r.hit) continue;
            T temp = 0;
            for (T t = r.tn; t < r.tf; t = t + delta) {
                const T x = (r.sx + t * r.rdx - cmin_x) / dx;
                const T y = (r.sy + t * r.rdy - cmin_y) / dy;
                const int ix = (int)std::floor(x), iy = (int)std::floor(y);
                if (ix > 0 && ix < nx - 1 && iy > 0 && iy < ny - 1) {
                    const T wxl = x - ix, wyl = y - iy;
                    const T* p = img + iy * nx + ix;
                    temp = temp + (1 - wyl) * ((1 - wxl) * p[0] + wxl * p[1]) + wyl * ((1 - wxl) * p[nx] + wxl * p[nx + 1]);
                }
            }
            proj[ib * g.nu + iu] = temp * delta / 10;
        }
}

template <class T>
void cbct_proj_rd(const Geom3& g, const double* betas, const float* couchz, const T* img, T* proj, int nthreads) {
    const T wu = (T)((-(g.nu - 1.0) / 2.0 + g.ou) * g.du);
    const T wv = (T)((-(g.nv - 1.0) / 2.0 + g.ov) * g.dv);
    const T cmin_x = (T)((-(g.nx - 1.0) / 2.0 + g.ox) * g.dx), cmax_x = (T)((+(g.nx - 1.0) / 2.0 + g.ox) * g.dx);
    const T cmin_y = (T)((-(g.ny - 1.0) / 2.0 + g.oy) * g.dy), cmax_y = (T)((+(g.ny - 1.0) / 2.0 + g.oy) * g.dy);
    const T delta = (T)(g.dx / 2.0);
    const T dx = g.dx, dy = g.dy, dz = g.dz, dv = g.dv;
    const int nx = g.nx, ny = g.ny, nz = g.nz, nv = g.nv;
    const size_t nxy = (size_t)nx * ny;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic)
    for (int ib = 0; ib < g.nviews; ++ib) {
        const T cmin_z = (T)((-(nz - 1.0) / 2.0 + g.oz) * g.dz + couchz[ib]);
        const T source_z = 0;
        for (int iu = 0; iu < g.nu; ++iu) {
            const RayRow<T> r = ray_row<T>(betas[ib], iu, (T)g.du, wu, (T)g.sad, (T)g.sdd, g.dettype, cmin_x,
                                           cmax_x, cmin_y, cmax_y);
            if (!r.hit) continue;
            T* out = proj + ((size_t)ib * g.nu + iu) * nv;
            for (int iv = 0; iv < nv; ++iv) {
                const T rdz = (iv * dv + wv - source_z) / r.spd;
                T temp = 0;
                for (T t = r.tn; t < r.tf; t = t + delta) {
                    const T z = (source_z + t * rdz - cmin_z) / dz;
                    if (z < 1) continue;
                    if ((double)z > nz - 1.1) break;
                    const T x = (r.sx + t * r.rdx - cmin_x) / dx;
                    const T y = (r.sy + t * r.rdy - cmin_y) / dy;
                    const int ix = (int)std::floor(x), iy = (int)std::floor(y), iz = (int)std::floor(z);
                    const T wzl = z - iz;
                    if (ix > 0 && ix < nx - 1 && iy > 0 && iy < ny - 1) {
                        const T wxl = x - ix, wyl = y - iy;
                        const T* p = img + iz * nxy + (size_t)iy * nx + ix;
                        temp = temp +
                               (1 - wzl) * ((1 - wyl) * ((1 - wxl) * p[0] + wxl * p[1]) +
                                            wyl * ((1 - wxl) * p[nx] + wxl * p[nx + 1])) +
                               wzl * ((1 - wyl) * ((1 - wxl) * p[nxy] + wxl * p[nxy + 1]) +
                                      wyl * ((1 - wxl) * p[nxy + nx] + wxl * p[nxy + nx + 1]));
                    }
                }
                out[iv] = temp * delta * std::sqrt(1 + rdz * rdz) / 10;
            }
        }
    }
}

// Counts of the work a 3D DD/SF operator performs, independent of image values:
// U = active (voxel, view) pairs inside the iz window with a non-empty v footprint,
// C = (voxel, detector cell) weight pairs, K = active (column, view) pairs.
void cbct_count(FootKind kind, const Geom3& g, const double* betas, const float* couchz,
                const uint8_t* map, int nthreads, double out[3]) {
    Cone<double> k(g);
    double U = 0, C = 0, K = 0;
    const double dx = g.dx, dy = g.dy, sad = g.sad;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic) reduction(+ : U, C, K)
    for (int ib = 0; ib < g.nviews; ++ib) {
        const double wz = k.wz(couchz, ib);
        ViewTrig<double> v(betas[ib], dx, dy);
        double wu16[16];
        for (int iy = 0; iy < g.ny; ++iy)
            for (int ix = 0; ix < g.nx; ++ix) {
                if (map && !map[iy * g.nx + ix]) continue;
                ColumnGeom<double> cg((ix + k.wx) * dx, (iy + k.wy) * dy, v, sad, dx);
                int iu0, iu1;
                if (!u_footprint(kind, g, k, cg, v, iu0, iu1, wu16)) continue;
                double zinc, vstart;
                int iz0, iz1;
                if (!v_window(g, cg.d, wz, k.wv, zinc, vstart, iz0, iz1)) continue;
                K += 1;
                for (int iz = iz0; iz < iz1; ++iz) {
                    double vmin = vstart + iz * zinc, clo, chi;
                    int iv0, iv1;
                    if (!clamp_cells(vmin, vmin + zinc, g.nv, clo, chi, iv0, iv1, 16)) continue;
                    U += 1;
                    C += (double)(iu1 - iu0 + 1) * (iv1 - iv0 + 1);
                }
            }
    }
    out[0] = U;
    out[1] = C;
    out[2] = K;
}

Geom2 make2(const int32_t* nxy, const float* dxy, const float* oxy, int32_t nu, float du, float ou,
            float sad, float add, float sdd, int32_t nviews, int32_t dettype) {
    return Geom2{nxy[0], nxy[1], dxy[0], dxy[1], oxy[0], oxy[1], nu, du, ou, sad, add, sdd, nviews, dettype};
}

Geom3 make3(const int32_t* nxyz, const float* dxyz, const float* oxyz, const int32_t* nuv,
            const float* duv, const float* ouv, float sad, float add, float sdd, int32_t nviews,
            int32_t dettype) {
    return Geom3{nxyz[0], nxyz[1], nxyz[2], dxyz[0], dxyz[1], dxyz[2], oxyz[0], oxyz[1], oxyz[2],
                 nuv[0], nuv[1], duv[0], duv[1], ouv[0], ouv[1], sad, add, sdd, nviews, dettype};
}

template <class T>
int fbct_dispatch(const char* op, const Geom2& g, const double* betas, const void* in, void* out,
                  int nthreads) {
    const T* src = static_cast<const T*>(in);
    T* dst = static_cast<T*>(out);
    if (!std::strcmp(op, "proj,dd")) fbct_proj_dd<T>(g, betas, src, dst, nthreads);
    else if (!std::strcmp(op, "proj,rd")) fbct_proj_rd<T>(g, betas, src, dst, nthreads);
    else if (!std::strcmp(op, "proj,tf")) fbct_proj_tf<T>(g, betas, src, dst, nthreads);
    else if (!std::strcmp(op, "back,dd")) fbct_back_dd<T>(g, betas, dst, src, nthreads);
    else if (!std::strcmp(op, "back,tf")) fbct_back_tf<T>(g, betas, dst, src, nthreads);
    else if (!std::strcmp(op, "back,pd")) fbct_back_pd<T>(g, betas, dst, src, nthreads);
    else return 1;
    return 0;
}

template <class T>
int cbct_dispatch(const char* op, const Geom3& g, const double* betas, const float* couchz, const void* in,
                  const uint8_t* map, void* out, int nthreads) {
    const T* src = static_cast<const T*>(in);
    T* dst = static_cast<T*>(out);
    if (!std::strcmp(op, "proj,rd")) cbct_proj_rd<T>(g, betas, couchz, src, dst, nthreads);  // no map
    else if (!std::strcmp(op, "proj,dd")) cbct_proj<T>(FOOT_DD, g, betas, couchz, src, map, dst, nthreads);
    else if (!std::strcmp(op, "proj,tf")) cbct_proj<T>(FOOT_TF, g, betas, couchz, src, map, dst, nthreads);
    else if (!std::strcmp(op, "back,dd")) cbct_back<T>(FOOT_DD, g, betas, couchz, dst, map, src, nthreads);
    else if (!std::strcmp(op, "back,tf")) cbct_back<T>(FOOT_TF, g, betas, couchz, dst, map, src, nthreads);
    else if (!std::strcmp(op, "back,pd")) cbct_back_pd<T>(true, g, betas, couchz, dst, map, src, nthreads);
    else if (!std::strcmp(op, "back,hpd")) cbct_back_pd<T>(false, g, betas, couchz, dst, map, src, nthreads);
    else return 1;
    return 0;
}

}  // namespace

extern "C" {

// 2D fan beam: op is one of proj,dd proj,tf back,dd back,tf back,pd.
// prec 32 -> float in/out (faithful), 64 -> double in/out (fp64).  Returns 0 on success.
int oracle_fbct(const char* op, int prec, const int32_t* nxy, const float* dxy, const float* oxy,
                int32_t nu, float du, float ou, float sad, float add, float sdd, int32_t nviews,
                const double* betas, int32_t dettype, const void* in, void* out, int nthreads) {
    Geom2 g = make2(nxy, dxy, oxy, nu, du, ou, sad, add, sdd, nviews, dettype);
    if (prec == 32) return fbct_dispatch<float>(op, g, betas, in, out, nthreads);
    if (prec == 64) return fbct_dispatch<double>(op, g, betas, in, out, nthreads);
    return 2;
}

// 3D cone beam: op is one of proj,dd proj,tf back,dd back,tf back,pd back,hpd.
// map may be NULL (the 15-argument no-map family).
int oracle_cbct(const char* op, int prec, const int32_t* nxyz, const float* dxyz, const float* oxyz,
                const int32_t* nuv, const float* duv, const float* ouv, float sad, float add, float sdd,
                int32_t nviews, const double* betas, const float* couchz, int32_t dettype, const void* in,
                const uint8_t* map, void* out, int nthreads) {
    Geom3 g = make3(nxyz, dxyz, oxyz, nuv, duv, ouv, sad, add, sdd, nviews, dettype);
    if (prec == 32) return cbct_dispatch<float>(op, g, betas, couchz, in, map, out, nthreads);
    if (prec == 64) return cbct_dispatch<double>(op, g, betas, couchz, in, map, out, nthreads);
    return 2;
}

// Work counts (U, C, K) of a 3D DD ("dd") or SF ("tf") operator, in double.
int oracle_cbct_count(const char* kind, const int32_t* nxyz, const float* dxyz, const float* oxyz,
                      const int32_t* nuv, const float* duv, const float* ouv, float sad, float add,
                      float sdd, int32_t nviews, const double* betas, const float* couchz, int32_t dettype,
                      const uint8_t* map, int nthreads, double* out3) {
    Geom3 g = make3(nxyz, dxyz, oxyz, nuv, duv, ouv, sad, add, sdd, nviews, dettype);
    FootKind fk;
    if (!std::strcmp(kind, "dd")) fk = FOOT_DD;
    else if (!std::strcmp(kind, "tf")) fk = FOOT_TF;
    else return 1;
    cbct_count(fk, g, betas, couchz, map, nthreads, out3);
    return 0;
}

}  // extern "C"
