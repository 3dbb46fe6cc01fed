// CPU oracle for CTSim's projection-matrix (PMat / C-arm) cone-beam operators.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/ load this code (linked into liboracle.so), as
// the checker of the GPU operators; the product (ctsim_amd/) never links or calls it.
//
// Restates geom/pcbct (a C MEX: the gateway compiles the kernels as C, so `sqrt` and
// `fabs` take and return double, and the `0.5`, `1.0`, `1E-6` literals promote):
//   pmat_proj_dd   geom/pcbct/cbct_proj_pmat.c:8-215    (cbct_proj_dd)
//   pmat_back_dd   geom/pcbct/cbct_back_pmat.c:137-321  (cbct_back_dd)
//   pmat_back_pd   geom/pcbct/cbct_back_pmat.c:9-134    (cbct_back_pd: pd = offBoundary,
//                                                       hpd = the clamped form)
//   op dispatch    geom/pcbct/cbct_geom_pmat_mex.c:19-128, 171-253, 260-309
//   ("proj,tf" and "back,tf" run the DD code, as the gateway does after its warning)
//
//   T = float  : the reference's float / double mix, expression by expression.
//   T = double : the same formulas in double throughout (the accuracy oracle).
//
// Decisions where the reference has undefined behaviour:
//   * du, dv: the gateway reads duv[1] and duv[2] of a two-element vector
//     (cbct_geom_pmat_mex.c:49-50, :148-149).  Here du and dv are arguments; the MEX
//     shim passes duv[1], duv[2] for a three-element duv (the reference's reading) and
//     duv[0], duv[1] for a two-element one.
//   * the u footprint is capped at 16 cells (iumin + 15, cbct_proj_pmat.c:118) while
//     weights_u holds FOOTPRINT_SIZE_U = 10 (ct_def.h:81): footprints of 11..16 cells
//     write past the reference's array.  Here the array holds 16; footprints of at most
//     10 cells are computed exactly as the reference does.
// Reproduced as written: the v offset of the cell-amplitude scaling uses offset_uv[2 ib]
// (the u offset, cbct_proj_pmat.c:203), the forward f0 has no z term
// (cbct_proj_pmat.c:76) while the pd f0 has one (cbct_back_pmat.c:73), and the
// backprojectors' 10-row v cap.
//
// Layouts are MATLAB column-major: image [nx ny nz], sinogram [nv nu nviews], pmat
// [3 4 nviews], cameras [3 nviews], sdds [nviews], offset_uv [2 nviews], map [nx ny].

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// C semantics of the reference: D is the type its double-promoted expressions take
template <class T>
struct Prom {
    using D = double;
};

template <class A, class B>
inline auto cmax(A a, B b) -> decltype(a + b) { return a > b ? a : b; }
template <class A, class B>
inline auto cmin(A a, B b) -> decltype(a + b) { return a < b ? a : b; }

constexpr int kFootU = 16;
constexpr int kFootV = 10;

template <class T>
struct PmatGeom {
    int nx, ny, nz, nu, nv, nviews;
    T dx, dy, dz, du, dv;
    T wx, wy, wz, wu, wv;
    const float* pmat;
    const float* cameras;
    const float* sdds;
    const float* ouv;
    const uint8_t* map;
};

template <class T>
PmatGeom<T> make_pmat(const int32_t* n, const float* d, const float* o, const int32_t* nuv, float du, float dv,
                      int32_t nviews, const float* pmat, const float* cameras, const float* sdds, const float* ouv,
                      const uint8_t* map) {
    PmatGeom<T> g{};
    g.nx = n[0]; g.ny = n[1]; g.nz = n[2];
    g.nu = nuv[0]; g.nv = nuv[1];
    g.nviews = nviews;
    g.dx = d[0]; g.dy = d[1]; g.dz = d[2];
    g.du = du; g.dv = dv;
    // float wx = - ( nx - 1.0 ) / 2.0 + offset_x;   (cbct_proj_pmat.c:34-38)
    g.wx = (T)(-(g.nx - 1.0) / 2.0 + (double)o[0]);
    g.wy = (T)(-(g.ny - 1.0) / 2.0 + (double)o[1]);
    g.wz = (T)(-(g.nz - 1.0) / 2.0 + (double)o[2]);
    g.wu = (T)(-(g.nu - 1.0) / 2.0);
    g.wv = (T)(-(g.nv - 1.0) / 2.0);
    g.pmat = pmat; g.cameras = cameras; g.sdds = sdds; g.ouv = ouv; g.map = map;
    return g;
}

// u footprint of column (ix, iy) in view ib (cbct_proj_pmat.c:69-138); false when the
// column is off the detector.  Also returns v0 / f0 and dvdz.
template <class T>
bool pmat_u_footprint(const PmatGeom<T>& g, int ib, int ix, int iy, T& ampu, int& iumin, int& iumax, T* wu,
                      T& v0f, T& dvdz) {
    using D = typename Prom<T>::D;
    const float* pf = g.pmat + 12 * ib;
    T pm[12];
    for (int k = 0; k < 12; ++k) pm[k] = (T)pf[k];
    const T cx = (T)g.cameras[3 * ib], cy = (T)g.cameras[3 * ib + 1];
    const T xc = (T(ix) + g.wx) * g.dx;
    const T yc = (T(iy) + g.wy) * g.dy;
    const T u0 = xc * pm[0] + yc * pm[3] + g.wz * g.dz * pm[6] + pm[9];
    const T v0 = xc * pm[1] + yc * pm[4] + g.wz * g.dz * pm[7] + pm[10];
    const T f0 = xc * pm[2] + yc * pm[5] + pm[11];
    const T dsp_x = (T)std::fabs((D)(cx - xc));
    const T dsp_y = (T)std::fabs((D)(cy - yc));
    T umin, umax;
    if (std::fabs((D)dsp_x) > std::fabs((D)dsp_y)) {
        umin = (u0 - g.dy * pm[3] / T(2)) / (f0 - g.dy * pm[5] / T(2));
        umax = (u0 + g.dy * pm[3] / T(2)) / (f0 + g.dy * pm[5] / T(2));
        ampu = (T)((D)g.dx * std::sqrt((D)(dsp_x * dsp_x + dsp_y * dsp_y)) / (D)dsp_x / 10.0);
    } else {
        umin = (u0 - g.dx * pm[0] / T(2)) / (f0 - g.dx * pm[2] / T(2));
        umax = (u0 + g.dx * pm[0] / T(2)) / (f0 + g.dx * pm[2] / T(2));
        ampu = (T)((D)g.dx * std::sqrt((D)(dsp_x * dsp_x + dsp_y * dsp_y)) / (D)dsp_y / 10.0);
    }
    if (umin >= umax) std::swap(umin, umax);
    umin = (T)cmax((D)umin + 0.5, 0);
    umax = (T)cmin((D)umax + 0.5, g.nu);
    if (umin > umax) return false;
    iumin = (int)std::floor(umin);
    iumax = (int)std::floor(umax);
    iumax = std::min(iumax, g.nu - 1);
    iumax = std::min(iumax, iumin + kFootU - 1);
    if (iumin == iumax) {
        wu[0] = umax - umin;
    } else {
        for (int iu = iumin; iu <= iumax; ++iu) {
            if (iu == iumin) wu[0] = (T)((D)iumin + 1.0 - (D)umin);
            else if (iu == iumax) wu[iu - iumin] = (T)(umax - (T)iumax);
            else wu[iu - iumin] = T(1);
        }
    }
    v0f = v0 / f0;
    dvdz = pm[7] / f0;
    return true;
}

// v footprint of voxel iz (cbct_proj_pmat.c:155-190); false when it misses the detector
template <class T>
bool pmat_v_footprint(const PmatGeom<T>& g, int iz, T v0f, T dvdz, int& ivmin, int& ivmax, T* wv) {
    using D = typename Prom<T>::D;
    T vmin, vmax;
    if (dvdz > 0) {
        vmin = (T)((D)v0f + ((D)(T)iz - 0.5) * (D)g.dz * (D)dvdz);
        vmax = (T)((D)v0f + ((D)(T)iz + 0.5) * (D)g.dz * (D)dvdz);
    } else {
        vmin = (T)((D)v0f + ((D)(T)iz + 0.5) * (D)g.dz * (D)dvdz);
        vmax = (T)((D)v0f + ((D)(T)iz - 0.5) * (D)g.dz * (D)dvdz);
    }
    vmin = (T)cmax((D)vmin + 0.5, 0);
    vmax = (T)cmin((D)vmax + 0.5, g.nv);
    if (vmin > vmax) return false;
    ivmin = (int)std::floor(vmin);
    ivmax = (int)std::floor(vmax);
    ivmax = std::min(ivmax, g.nv - 1);
    ivmax = std::min(ivmax, ivmin + kFootV - 1);
    if (ivmin == ivmax) {
        wv[0] = vmax - vmin;
    } else {
        for (int iv = ivmin; iv <= ivmax; ++iv) {
            if (iv == ivmin) wv[0] = (T)((D)(T)ivmin + 1.0 - (D)vmin);
            else if (iv == ivmax) wv[iv - ivmin] = (T)(vmax - (T)ivmax);
            else wv[iv - ivmin] = T(1);
        }
    }
    return true;
}

// amplitude of detector cell (iu, iv) of view ib (cbct_proj_pmat.c:197-210)
template <class T>
T pmat_cell_amp(const PmatGeom<T>& g, int ib, int iu, int iv) {
    using D = typename Prom<T>::D;
    const T sdd = (T)g.sdds[ib];
    const T off = (T)g.ouv[2 * ib];  // the u offset, for both coordinates, as written
    const T u = (T(iu) + g.wu + off) * g.du;
    const T uu = sdd * sdd + u * u;
    const T v = (T(iv) + g.wv + off) * g.dv;
    return (T)std::sqrt((D)((uu + v * v) / uu));
}

template <class T>
void pmat_proj_dd(const PmatGeom<T>& g, const T* img, T* proj, int nthreads) {
    const size_t nxy = (size_t)g.nx * g.ny;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
    for (int ib = 0; ib < g.nviews; ++ib) {
        T* pv = proj + (size_t)g.nu * g.nv * ib;
        T wu[kFootU], wv[kFootV];
        for (int iy = 0; iy < g.ny; ++iy)
            for (int ix = 0; ix < g.nx; ++ix) {
                if (!g.map[iy * g.nx + ix]) continue;
                T ampu, v0f, dvdz;
                int iumin, iumax;
                if (!pmat_u_footprint(g, ib, ix, iy, ampu, iumin, iumax, wu, v0f, dvdz)) continue;
                for (int iz = 0; iz < g.nz; ++iz) {
                    const T val = img[iz * nxy + (size_t)iy * g.nx + ix];
                    if (val == 0) continue;
                    const T voxel = ampu * val;
                    int ivmin, ivmax;
                    if (!pmat_v_footprint(g, iz, v0f, dvdz, ivmin, ivmax, wv)) continue;
                    for (int iu = iumin; iu <= iumax; ++iu)
                        for (int iv = ivmin; iv <= ivmax; ++iv)
                            pv[iu * g.nv + iv] += wu[iu - iumin] * wv[iv - ivmin] * voxel;
                }
            }
        for (int iu = 0; iu < g.nu; ++iu)
            for (int iv = 0; iv < g.nv; ++iv) pv[iu * g.nv + iv] = pv[iu * g.nv + iv] * pmat_cell_amp(g, ib, iu, iv);
    }
}

template <class T>
void pmat_back_dd(const PmatGeom<T>& g, const T* proj_in, T* img, int nthreads) {
    const size_t nxy = (size_t)g.nx * g.ny;
    // the reference scales its input sinogram in place (cbct_back_pmat.c:190-202); here a copy
    std::vector<T> proj(proj_in, proj_in + (size_t)g.nu * g.nv * g.nviews);
    for (int ib = 0; ib < g.nviews; ++ib)
        for (int iu = 0; iu < g.nu; ++iu)
            for (int iv = 0; iv < g.nv; ++iv) {
                T& p = proj[((size_t)ib * g.nu + iu) * g.nv + iv];
                p = p * pmat_cell_amp(g, ib, iu, iv);
            }
    // views in order per voxel column (the reference's view loop is outermost; every voxel
    // adds its views in increasing order either way)
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
    for (int iy = 0; iy < g.ny; ++iy)
        for (int ib = 0; ib < g.nviews; ++ib) {
            const T* pv = proj.data() + (size_t)g.nu * g.nv * ib;
            T wu[kFootU], wv[kFootV];
            for (int ix = 0; ix < g.nx; ++ix) {
                if (!g.map[iy * g.nx + ix]) continue;
                T ampu, v0f, dvdz;
                int iumin, iumax;
                if (!pmat_u_footprint(g, ib, ix, iy, ampu, iumin, iumax, wu, v0f, dvdz)) continue;
                for (int iz = 0; iz < g.nz; ++iz) {
                    int ivmin, ivmax;
                    if (!pmat_v_footprint(g, iz, v0f, dvdz, ivmin, ivmax, wv)) continue;
                    T voxel = T(0);
                    for (int iu = iumin; iu <= iumax; ++iu)
                        for (int iv = ivmin; iv <= ivmax; ++iv)
                            voxel += pv[iu * g.nv + iv] * wu[iu - iumin] * wv[iv - ivmin];
                    img[iz * nxy + (size_t)iy * g.nx + ix] += ampu * voxel;
                }
            }
        }
}

template <class T>
void pmat_back_pd(const PmatGeom<T>& g, const T* proj, T* img, bool off_boundary, int nthreads) {
    using D = typename Prom<T>::D;
    const size_t nxy = (size_t)g.nx * g.ny;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
    for (int iy = 0; iy < g.ny; ++iy)
        for (int ib = 0; ib < g.nviews; ++ib) {
            const float* pf = g.pmat + 12 * ib;
            T pm[12];
            for (int k = 0; k < 12; ++k) pm[k] = (T)pf[k];
            const T sdd = (T)g.sdds[ib];
            const T cx = (T)g.cameras[3 * ib], cy = (T)g.cameras[3 * ib + 1];
            for (int ix = 0; ix < g.nx; ++ix) {
                if (!g.map[iy * g.nx + ix]) continue;
                const T xc = (T(ix) + g.wx) * g.dx;
                const T yc = (T(iy) + g.wy) * g.dy;
                const T u0 = xc * pm[0] + yc * pm[3] + g.wz * g.dz * pm[6] + pm[9];
                const T v0 = xc * pm[1] + yc * pm[4] + g.wz * g.dz * pm[7] + pm[10];
                const T f0 = xc * pm[2] + yc * pm[5] + g.wz * g.dz * pm[8] + pm[11];
                const T w2s = sdd * sdd / ((cx - xc) * (cx - xc) + (cy - yc) * (cy - yc));
                for (int iz = 0; iz < g.nz; ++iz) {
                    const T fz = T(iz) * g.dz;
                    const T u = (u0 + fz * pm[6]) / (f0 + fz * pm[8]);
                    T v = (v0 + fz * pm[7]) / (f0 + fz * pm[8]);
                    if ((D)u < 1E-6 || (D)u > g.nu - 1 - 1E-6) continue;
                    const int iu = (int)std::floor(u);
                    const T wul = u - T(iu);
                    if (off_boundary) {
                        if ((D)v < 1E-6) continue;
                        if ((D)v > g.nv - 1 - 1E-6) break;
                    } else {
                        v = (T)cmax((D)v, 0.5);
                        v = (T)cmin((D)v, g.nv - 1.5);
                    }
                    const int iv = (int)std::floor(v);
                    const T wvl = v - T(iv);
                    const size_t np = (size_t)ib * g.nu * g.nv + (size_t)iu * g.nv + iv;
                    T& out = img[iz * nxy + (size_t)iy * g.nx + ix];
                    out = out + w2s * ((T(1) - wul) * ((T(1) - wvl) * proj[np] + wvl * proj[np + 1]) +
                                       wul * ((T(1) - wvl) * proj[np + g.nv] + wvl * proj[np + g.nv + 1]));
                }
            }
        }
}

template <class T>
int pmat_dispatch(const char* op, const PmatGeom<T>& g, const void* in, void* out, int nthreads) {
    const T* src = (const T*)in;
    T* dst = (T*)out;
    if (!std::strcmp(op, "proj,dd") || !std::strcmp(op, "proj,tf")) pmat_proj_dd(g, src, dst, nthreads);
    else if (!std::strcmp(op, "back,dd") || !std::strcmp(op, "back,tf")) pmat_back_dd(g, src, dst, nthreads);
    else if (!std::strcmp(op, "back,pd")) pmat_back_pd(g, src, dst, true, nthreads);
    else if (!std::strcmp(op, "back,hpd")) pmat_back_pd(g, src, dst, false, nthreads);
    else return 1;
    return 0;
}

}  // namespace

extern "C" {

// Projection-matrix cone beam: op proj,dd | proj,tf (= dd) | back,dd | back,tf (= dd) |
// back,pd | back,hpd.  prec 32 -> float in/out, 64 -> double in/out.  map is required
// (the reference gateway's 13-argument form always passes it).  Outputs zero-filled by
// the caller.  Returns 0 on success.
int oracle_pmat(const char* op, int prec, const int32_t* nxyz, const float* dxyz, const float* oxyz,
                const int32_t* nuv, float du, float dv, int32_t nviews, const float* pmat, const float* cameras,
                const float* sdds, const float* ouv, const uint8_t* map, const void* in, void* out, int nthreads) {
    if (!map) return 3;
    if (prec == 32)
        return pmat_dispatch<float>(op, make_pmat<float>(nxyz, dxyz, oxyz, nuv, du, dv, nviews, pmat, cameras, sdds,
                                                         ouv, map), in, out, nthreads);
    if (prec == 64)
        return pmat_dispatch<double>(op, make_pmat<double>(nxyz, dxyz, oxyz, nuv, du, dv, nviews, pmat, cameras,
                                                           sdds, ouv, map), in, out, nthreads);
    return 2;
}

}  // extern "C"
