// Host-side launchers shared between the kernel translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geom.h"

namespace ctsim {

enum DType { DT_F32 = 0, DT_F16 = 1, DT_BF16 = 2, DT_F64 = 3 };

inline size_t dtype_size(int dt) { return dt == DT_F64 ? 8 : dt == DT_F32 ? 4 : 2; }
// below this many (view, 64-cell tile) items a 2D forward projection splits each item over
// the four waves of a block (fewer items would leave the chip with under eight waves per SIMD)
constexpr long long kFewItems2D = 8192;

// 3D cone beam.  foot: 0 = DD, 1 = SF.  Volume in ZXY order.
template <class T>
hipError_t cbct_proj_launch(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* vol,
                            int vol_dt, const uint8_t* map, void* proj, int proj_dt, int foot, int exhaustive,
                            hipStream_t st);
// buffers and scalars of a Nesterov-SQS update fused into the backprojection (FP32; the
// update kernel's arguments, the gradient being the backprojection itself)
struct NestArgs {
    const float* x = nullptr;
    const float* z = nullptr;
    float* s = nullptr;
    const float* precom = nullptr;
    float* x_out = nullptr;
    float* z_out = nullptr;
    float* zcopy = nullptr;       // z' (or z' - zref), FP32, or null
    const float* zref = nullptr;
    double beta = 0.0;
    float delta = 0.f;
    double t = 1.0, t_next = 1.0;
    const double* tseq = nullptr;  // when set, t and t_next are tseq[*base + k] and tseq[*base + k + 1]
    const int* base = nullptr;
    int k = 0;
};
hipError_t cbct_back_nesterov_launch(const Geom<float>& g, const View<float>* views, int vstride, int nviews,
                                     const void* proj, int proj_dt, const uint8_t* map, int foot, float scale,
                                     const NestArgs& na, hipStream_t st);
template <class T>
hipError_t cbct_back_launch(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* proj,
                            int proj_dt, const uint8_t* map, void* vol, int vol_dt, int foot, T scale, int accumulate,
                            hipStream_t st);
template <class T>
hipError_t cbct_back_pd_launch(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* proj,
                               int proj_dt, const uint8_t* map, void* vol, int vol_dt, int off_boundary, T scale,
                               int accumulate, hipStream_t st);
hipError_t xyz_zxy_launch(const void* src, int src_dt, void* dst, int dst_dt, int nxy, int nz, bool to_zxy,
                          hipStream_t st);

// Ray-driven forward projection (geom/fbct_os/fbct_proj.c:3-153, geom/cbct_os/cbct_proj.c:10-198).
// One record per (view, detector column), built on the host in float exactly as the
// reference forms it (double sin/atan, float sinf/sqrtf, box slab tests); a ray that misses
// the box has tn > tf.  czmin is the view's lower z bound of the box (3D).
struct RayRec {
    float tn, tf, rdx, rdy, spd, sx, sy, czmin;
};
// Constants of the march, formed on the host like the reference (double expressions
// rounded to float).
struct RdConst {
    float wv, cmin_x, cmin_y, delta, dx, dy, dz, dv;
    int nx, ny, nz, nu, nv;
    double inv_dx, inv_dy, inv_dz;  // 1 / dx, 1 / dy, 1 / dz in double
};
hipError_t cbct_rd_launch(const RayRec* rays, int view0, int vstep, int nviews, const RdConst& c, const void* vol,
                          int vol_dt, void* proj, int proj_dt, hipStream_t st);
hipError_t fbct_rd_launch(const RayRec* rays, int view0, int vstep, int nviews, const RdConst& c, const void* img,
                          int img_dt, void* proj, int proj_dt, hipStream_t st);

// 2D fan beam.  kind: 0 = DD, 1 = SF, 2 = pd (back only).
template <class T>
hipError_t fbct_proj_launch(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* img,
                            int img_dt, void* proj, int proj_dt, int kind, int exhaustive, hipStream_t st);
template <class T>
hipError_t fbct_back_launch(const Geom<T>& g, const View<T>* views, int vstride, int nviews, const void* proj,
                            int proj_dt, void* img, int img_dt, int kind, T scale, hipStream_t st);

// batched 2D: slice table on the device
// One 2D slice of a batch: geometry, view table and data offsets.
template <class T>
struct Slice2 {
    Geom<T> g;
    const View<T>* views;
    int nviews;
    long long img_off;   // element offset of the image
    long long sino_off;  // element offset of the sinogram
    int item0;           // first work item of this slice in the flat work list
    int pad;
};

template <class T>
hipError_t fbct_proj_batch_launch(const void* slices, int nslices, int nitems, const void* img, int img_dt,
                                  void* proj, int proj_dt, int kind, int split, hipStream_t st);
template <class T>
hipError_t fbct_back_batch_launch(const void* slices, int nslices, int nitems, const void* proj, int proj_dt,
                                  void* img, int img_dt, int kind, int split, hipStream_t st);
size_t slice2_size(bool dbl);

// utility kernels
hipError_t fill_launch(void* p, int dt, size_t n, double value, hipStream_t st);
hipError_t convert_launch(const void* src, int sdt, void* dst, int ddt, size_t n, hipStream_t st);

}  // namespace ctsim

namespace ctsim {
// PWLS / reduction launchers (pwls.hip)
hipError_t reduction_scratch_init();
hipError_t residual_launch(const void* p, int pdt, const void* y, int ydt, const void* w, int wdt, void* r, int rdt,
                           size_t n, double wscale, double* phi, hipStream_t st);
hipError_t residual_launch_noalloc(const void* p, int pdt, const void* y, int ydt, const void* w, int wdt, void* r, int rdt,
                           size_t n, double wscale, double* phi, hipStream_t st);
hipError_t add_i32_launch(int* p, int inc, hipStream_t st);
hipError_t sqdiff_launch(const float* a, const float* b, size_t n, double* out, hipStream_t st);
hipError_t sqdiff_launch(const double* a, const double* b, size_t n, double* out, hipStream_t st);
hipError_t mul_launch(const void* a, int adt, const void* b, int bdt, void* o, int odt, size_t n, hipStream_t st);
hipError_t sub_launch(const float* a, const float* b, void* o, int odt, size_t n, hipStream_t st);
// FDK pre-processing (fdk.hip)
struct FdkGeom {
    float du, dv, ou, ov, sad, sdd;
    int flat;
};
struct FdkPlanImpl;
hipError_t fdk_plan_create(FdkPlanImpl** out, const FdkGeom& g, int nu, int nv, int nviews, const double* H, int L,
                           const float* view_weights, int f64, size_t max_workspace);
void fdk_plan_destroy(FdkPlanImpl* p);
hipError_t fdk_filter_launch(FdkPlanImpl* p, int view0, int nsub, const void* sino, int sdt, void* out, int odt,
                             hipStream_t st);

// data side (sim.hip)
hipError_t poisson_launch(const void* p, int pdt, size_t n, double i0, unsigned long long seed, float* counts,
                          hipStream_t st);
hipError_t neglog_launch(const float* counts, size_t n, double i0, double low, void* y, int ydt, hipStream_t st);
hipError_t neglog_ex_launch(const float* counts, size_t n, int ncell, double i0, double low, int compound,
                            const float* ff, const float* profile, void* y, int ydt, hipStream_t st);
hipError_t water_launch(const void* in, int idt, void* out, int odt, size_t n, int nv, int nu, const float* coeffs,
                        int order, int per_channel, hipStream_t st);
hipError_t weights_launch(const float* s, size_t n, double r, double v, void* w, int wdt, hipStream_t st);

// penalties (pfun 0 Huber, 1 quadratic, 2 hyperbola) and the PWLS updates (pwls.hip)
hipError_t penalty_launch(const float* x, int nx, int ny, int nz, int pfun, double delta, int mode, float* out,
                          double* cost, hipStream_t st);
hipError_t penalty_launch(const double* x, int nx, int ny, int nz, int pfun, double delta, int mode, double* out,
                          double* cost, hipStream_t st);
hipError_t nesterov_launch(const float* x, const float* z, float* s, const float* g, const float* precom, float* x_out,
                           float* z_out, void* zcopy, int zcdt, const float* zref, int nx, int ny, int nz,
                           double beta, double delta, int pfun, double t, double t_next, const double* tseq,
                           const int* base, int k,
                           hipStream_t st);
hipError_t nesterov_launch(const double* x, const double* z, double* s, const double* g, const double* precom,
                           double* x_out, double* z_out, int nx, int ny, int nz, double beta, double delta, int pfun,
                           double t, double t_next, hipStream_t st);
hipError_t sqs_launch(const float* x, float* x_out, const float* g, const float* precom, int nx, int ny, int nz,
                      double beta, double delta, int pfun, hipStream_t st);
hipError_t lalm_sg_launch(const float* dl, float* g, float* s, size_t n, double rho, hipStream_t st);
hipError_t lalm_inner_launch(const float* x, const float* xold, const float* s, const float* precom, float* x_out,
                             int nx, int ny, int nz, double beta, double delta, int pfun, double rho, hipStream_t st);
hipError_t nusqs_seed_launch(const float* x, float* v, int nx, int ny, int nz, double eps, hipStream_t st);
hipError_t div_launch(const float* a, const float* b, float* o, size_t n, hipStream_t st);
hipError_t absdiff_floor_launch(const float* x, const float* xref, float* o, size_t n, double eps, hipStream_t st);
hipError_t dynamic_range_launch(const float* x, size_t n, double t, double eta, float* out, hipStream_t st);
hipError_t nusqs_launch(const float* x, float* x_out, const float* g, const float* dl, const float* h, int nx, int ny,
                        int nz, double beta, double delta, int pfun, hipStream_t st);


// ---------------------------------------------------------------------------------
// projection-matrix (PMat / C-arm) cone beam, pmat.hip
// ---------------------------------------------------------------------------------
struct PmatView {  // one view: the 3x4 matrix as stored by the gateway (column-major), camera x, y, SDD, u offset
    float pm[12];
    float cx, cy, sdd, ou;
};
struct PmatConst {
    int nx, ny, nz, nu, nv, nviews;
    float dx, dy, dz, du, dv;
    float wx, wy, wz, wu, wv;
    // the reference's double tests u, v >= 1e-6, u <= nu - 1 - 1e-6, v <= nv - 1 - 1e-6 as
    // float thresholds
    float lo, u_hi, v_hi;
};
enum { PMAT_PROJ_DD = 0, PMAT_BACK_DD = 1, PMAT_BACK_PD = 2, PMAT_BACK_HPD = 3 };
hipError_t pmat_launch(int kind, const PmatConst& c, const PmatView* views, const uint8_t* map, const float* in,
                       float* out, float* tmp, hipStream_t st);
}  // namespace ctsim
