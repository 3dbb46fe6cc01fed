// C ABI of libctsim_hip.so (declared in include/ctsim_amd.h).
//
// The MEX-signature gateways reproduce the dispatch and validation of
// geom/fbct_os/fbct_geom_mex.c:75-326 and geom/cbct_os/cbct_geom_mex.c:83-600 on top
// of the device plan API; the plan API keeps geometry and the per-view table resident
// on the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ctsim_amd.h"
#include "internal.h"

using namespace ctsim;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? CTSIM_ENOMEM : CTSIM_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                           \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

void copy_err(char* errbuf, size_t errlen) {
    if (errbuf && errlen) {
        std::strncpy(errbuf, g_err.c_str(), errlen - 1);
        errbuf[errlen - 1] = '\0';
    }
}

bool valid_dtype(int dt) { return dt == CTSIM_F32 || dt == CTSIM_F16 || dt == CTSIM_BF16 || dt == CTSIM_F64; }

// Scalar geometry in compute type T, rounded the way the reference rounds it:
// w* are evaluated in double and cast (geom/cbct_os/cbct_proj.c:454-458), sdd_du and
// (dz/dv)*sdd in T.
template <class T>
Geom<T> make_geom(const ctsim_geom& g) {
    Geom<T> k;
    k.nx = g.nx; k.ny = g.ny; k.nz = g.nz; k.nu = g.nu; k.nv = g.nv; k.dettype = g.detector_type;
    k.dx = (T)g.dx; k.dy = (T)g.dy; k.dz = (T)g.dz; k.du = (T)g.du; k.dv = (T)g.dv;
    k.sad = (T)g.sad; k.sdd = (T)g.sdd;
    k.sdd_du = (T)g.sdd / (T)g.du;
    k.zk = (T)g.dz / (T)g.dv * (T)g.sdd;
    k.wx = (T)(-(g.nx - 1.0) / 2.0 + (double)g.ox);
    k.wy = (T)(-(g.ny - 1.0) / 2.0 + (double)g.oy);
    k.wu = (T)(-(g.nu - 1.0) / 2.0 + (double)g.ou);
    k.wv = (T)(-(g.nv - 1.0) / 2.0 + (double)g.ov);
    k.ou = (T)g.ou;
    k.zstride = g.nz;
    k.zoff = T(0);
    k.sched_proj = 0;
    k.sched_back = 1;  // super-tile column order: measured faster at 41 and 720 views
    return k;
}

template <class T>
std::vector<View<T>> make_views(const ctsim_geom& g, int nviews, const double* betas, const float* couchz) {
    std::vector<View<T>> v(nviews);
    const T dx = (T)g.dx, dy = (T)g.dy;
    for (int b = 0; b < nviews; ++b) {
        View<T>& e = v[b];
        e.s = (T)std::sin(betas[b]);
        e.c = (T)std::cos(betas[b]);
        e.dxs2 = dx * e.s / T(2);
        e.dxc2 = dx * e.c / T(2);
        e.dys2 = dy * e.s / T(2);
        e.ndyc2 = -dy * e.c / T(2);
        const T cz = couchz ? (T)couchz[b] : T(0);
        e.wz = (T)(-(g.nz - 1.0) / 2.0 + (double)(T)g.oz + (double)(cz / (T)g.dz));
        e.pad = T(0);
    }
    return v;
}

int check_geom(int ndim, const ctsim_geom& g) {
    if (ndim != 2 && ndim != 3) return fail(CTSIM_EINVAL, "Error: geometry must be 2D or 3D.");
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || g.nu <= 0 || g.nv <= 0)
        return fail(CTSIM_EINVAL, "Error: image and detector sizes must be positive.");
    if (ndim == 2 && (g.nz != 1 || g.nv != 1)) return fail(CTSIM_EINVAL, "Error: 2D geometry needs nz = nv = 1.");
    if (!(g.du != 0.f) || !(g.dx != 0.f) || !(g.dy != 0.f) || (ndim == 3 && (!(g.dz != 0.f) || !(g.dv != 0.f))))
        return fail(CTSIM_EINVAL, "Error: pixel and detector spacings must be nonzero.");
    if ((double)g.nx * g.ny * g.nz > 4.0e9 || (double)g.nu * g.nv > 2.0e9)
        return fail(CTSIM_EINVAL, "Error: problem size exceeds 32-bit indexing of this build.");
    if (ndim == 3 && g.nu >= 32768)
        return fail(CTSIM_EUNSUP, "Error: 3D detectors up to 32767 columns are supported (nu = %d).", g.nu);
    return CTSIM_OK;
}

enum OpKind { OP_PROJ_RD, OP_PROJ_DD, OP_PROJ_TF, OP_BACK_DD, OP_BACK_TF, OP_BACK_PD, OP_BACK_HPD, OP_BAD };

OpKind parse_op(const char* op) {
    if (!op) return OP_BAD;
    if (!std::strcmp(op, "proj,rd")) return OP_PROJ_RD;
    if (!std::strcmp(op, "proj,dd")) return OP_PROJ_DD;
    if (!std::strcmp(op, "proj,tf")) return OP_PROJ_TF;
    if (!std::strcmp(op, "back,dd")) return OP_BACK_DD;
    if (!std::strcmp(op, "back,tf")) return OP_BACK_TF;
    if (!std::strcmp(op, "back,pd")) return OP_BACK_PD;
    if (!std::strcmp(op, "back,hpd")) return OP_BACK_HPD;
    return OP_BAD;
}

// The ray table of "proj,rd", in float with the reference's operations and promotions
// (geom/cbct_os/cbct_proj.c:38-134, geom/fbct_os/fbct_proj.c:22-121): box corners and dx/2 as
// double expressions rounded to float, sin / cos of beta and atan in double, phi rounded
// to float and sinf / cosf / sqrtf of floats.  The reference's axis-parallel test
// `raydir < -1E-12 && raydir > 1E-12` is never true, so both slab tests always divide.
RdConst rd_const(const ctsim_geom& g, int ndim) {
    RdConst c{};
    c.wv = ndim == 3 ? (float)((-(g.nv - 1.0) / 2.0 + g.ov) * g.dv) : 0.f;
    c.cmin_x = (float)((-(g.nx - 1.0) / 2.0 + g.ox) * g.dx);
    c.cmin_y = (float)((-(g.ny - 1.0) / 2.0 + g.oy) * g.dy);
    c.delta = (float)(g.dx / 2.0);
    c.dx = g.dx;
    c.dy = g.dy;
    c.dz = ndim == 3 ? g.dz : 1.f;
    c.dv = ndim == 3 ? g.dv : 1.f;
    c.nx = g.nx;
    c.ny = g.ny;
    c.nz = ndim == 3 ? g.nz : 1;
    c.nu = g.nu;
    c.nv = ndim == 3 ? g.nv : 1;
    c.inv_dx = 1.0 / (double)c.dx;
    c.inv_dy = 1.0 / (double)c.dy;
    c.inv_dz = 1.0 / (double)c.dz;
    return c;
}

std::vector<RayRec> rd_table(const ctsim_geom& g, int ndim, int nviews, const double* betas, const float* couchz) {
    const float wu = (float)((-(g.nu - 1.0) / 2.0 + g.ou) * g.du);
    const float cmin_x = (float)((-(g.nx - 1.0) / 2.0 + g.ox) * g.dx);
    const float cmax_x = (float)((+(g.nx - 1.0) / 2.0 + g.ox) * g.dx);
    const float cmin_y = (float)((-(g.ny - 1.0) / 2.0 + g.oy) * g.dy);
    const float cmax_y = (float)((+(g.ny - 1.0) / 2.0 + g.oy) * g.dy);
    std::vector<RayRec> t((size_t)nviews * g.nu);
    for (int ib = 0; ib < nviews; ++ib) {
        const double beta = betas[ib];
        const float sin_a = (float)std::sin(beta), cos_a = (float)std::cos(beta);
        const float sx = -g.sad * sin_a, sy = -g.sad * cos_a;
        const float czmin =
            ndim == 3 ? (float)((-(g.nz - 1.0) / 2.0 + g.oz) * g.dz + (couchz ? couchz[ib] : 0.f)) : 0.f;
        for (int iu = 0; iu < g.nu; ++iu) {
            RayRec r{1.f, 0.f, 0.f, 0.f, 1.f, sx, sy, czmin};  // tn > tf: the ray misses the box
            const float u = iu * g.du + wu;
            float spd, phi;
            if (g.detector_type == 1) {
                spd = sqrtf(g.sdd * g.sdd + u * u);
                phi = (float)(beta + std::atan((double)(u / g.sdd)));
            } else {
                spd = g.sdd;
                phi = (float)(beta + (double)(u / g.sdd));
            }
            const float rdx = sinf(phi), rdy = cosf(phi);
            float tn = -1E10f, tf = 1E10f;
            bool hit = true;
            {
                float t1 = (cmin_x - sx) / rdx, t2 = (cmax_x - sx) / rdx;
                if (t1 > t2) std::swap(t1, t2);
                if (t1 > tn) tn = t1;
                if (t2 < tf) tf = t2;
                if (tn > tf || tf < 0) hit = false;
            }
            if (hit) {
                float t1 = (cmin_y - sy) / rdy, t2 = (cmax_y - sy) / rdy;
                if (t1 > t2) std::swap(t1, t2);
                if (t1 > tn) tn = t1;
                if (t2 < tf) tf = t2;
                if (tn > tf || tf < 0) hit = false;
            }
            if (hit) {
                r.tn = tn;
                r.tf = tf;
            }
            r.rdx = rdx;
            r.rdy = rdy;
            r.spd = spd;
            t[(size_t)ib * g.nu + iu] = r;
        }
    }
    return t;
}

}  // namespace

struct ctsim_plan {
    int ndim = 0;
    int compute = CTSIM_F32;
    int nviews = 0;
    int device = 0;
    ctsim_geom g{};
    Geom<float> gf{};
    Geom<double> gd{};
    View<float>* d_vf = nullptr;
    View<double>* d_vd = nullptr;
    uint8_t* d_map = nullptr;
    std::vector<double> betas;  // kept for the ray table, built on the first "proj,rd"
    std::vector<float> couchz;
    mutable RayRec* d_rays = nullptr;
    mutable std::mutex rays_mutex;
};

extern "C" {

const char* ctsim_last_error(void) { return g_err.c_str(); }
int ctsim_version(void) { return 100; }

// ---------------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------------
int ctsim_plan_create(ctsim_plan** out, int32_t ndim, const ctsim_geom* g, int32_t nviews, const double* betas,
                      const float* couchz, const uint8_t* map, int32_t compute) {
    if (!out || !g) return fail(CTSIM_EINVAL, "Error: null plan or geometry.");
    *out = nullptr;
    if (int rc = check_geom(ndim, *g)) return rc;
    if (nviews <= 0 || !betas) return fail(CTSIM_EINVAL, "Error: mx_betas must have noviews X 1 double array.");
    if (compute != CTSIM_F32 && compute != CTSIM_F64)
        return fail(CTSIM_EINVAL, "Error: compute precision must be CTSIM_F32 or CTSIM_F64.");
    ctsim_plan* p = new (std::nothrow) ctsim_plan;
    if (!p) return fail(CTSIM_ENOMEM, "Error: host allocation failed.");
    p->ndim = ndim;
    p->compute = compute;
    p->nviews = nviews;
    p->g = *g;
    p->betas.assign(betas, betas + nviews);
    if (couchz) p->couchz.assign(couchz, couchz + nviews);
    HIP_TRY(hipGetDevice(&p->device), "hipGetDevice");
    hipError_t e = reduction_scratch_init();
    if (e != hipSuccess) {
        delete p;
        return hip_fail(e, "reduction scratch");
    }
    if (compute == CTSIM_F32) {
        p->gf = make_geom<float>(*g);
        auto v = make_views<float>(*g, nviews, betas, couchz);
        e = hipMalloc((void**)&p->d_vf, v.size() * sizeof(View<float>));
        if (e == hipSuccess) e = hipMemcpy(p->d_vf, v.data(), v.size() * sizeof(View<float>), hipMemcpyHostToDevice);
    } else {
        p->gd = make_geom<double>(*g);
        auto v = make_views<double>(*g, nviews, betas, couchz);
        e = hipMalloc((void**)&p->d_vd, v.size() * sizeof(View<double>));
        if (e == hipSuccess) e = hipMemcpy(p->d_vd, v.data(), v.size() * sizeof(View<double>), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && map && ndim == 3) {
        const size_t nm = (size_t)g->nx * g->ny;
        e = hipMalloc((void**)&p->d_map, nm);
        if (e == hipSuccess) e = hipMemcpy(p->d_map, map, nm, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        ctsim_plan_destroy(p);
        return hip_fail(e, "plan upload");
    }
    *out = p;
    return CTSIM_OK;
}

int ctsim_plan_destroy(ctsim_plan* p) {
    if (!p) return CTSIM_OK;
    if (p->d_vf) (void)hipFree(p->d_vf);
    if (p->d_vd) (void)hipFree(p->d_vd);
    if (p->d_map) (void)hipFree(p->d_map);
    if (p->d_rays) (void)hipFree(p->d_rays);
    delete p;
    return CTSIM_OK;
}

static int check_views(const ctsim_plan* p, int view0, int vstep, int nsub) {
    if (nsub <= 0 || vstep <= 0 || view0 < 0 || (long long)view0 + (long long)(nsub - 1) * vstep >= p->nviews)
        return fail(CTSIM_EINVAL, "Error: view selection (%d, %d, %d) outside the plan's %d views.", view0, vstep,
                    nsub, p->nviews);
    return CTSIM_OK;
}

static int plan_proj_impl(const ctsim_plan* p, const char* op, int view0, int vstep, int nsub, const void* d_img,
                          int img_dt, void* d_sino, int sino_dt, void* stream, int exhaustive) {
    if (!p) return fail(CTSIM_EINVAL, "Error: null plan.");
    const OpKind k = parse_op(op);
    if (k != OP_PROJ_DD && k != OP_PROJ_TF && k != OP_PROJ_RD) return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    if (int rc = check_views(p, view0, vstep, nsub)) return rc;
    if (!d_img || !d_sino) return fail(CTSIM_EINVAL, "Error: null device buffer.");
    if (!valid_dtype(img_dt) || !valid_dtype(sino_dt)) return fail(CTSIM_EINVAL, "Error: bad dtype.");
    hipStream_t st = (hipStream_t)stream;
    if (k == OP_PROJ_RD) {
        // the ray march is the reference's float arithmetic; the plan's map is not used
        // (cbct_geom_mex runs "proj,rd" only in the 15-argument, map-free form)
        if (p->compute != CTSIM_F32) return fail(CTSIM_EUNSUP, "Error: proj,rd runs in FP32 only.");
        {
            std::lock_guard<std::mutex> lock(p->rays_mutex);
            if (!p->d_rays) {
                const std::vector<RayRec> t =
                    rd_table(p->g, p->ndim, p->nviews, p->betas.data(), p->couchz.empty() ? nullptr : p->couchz.data());
                RayRec* d = nullptr;
                HIP_TRY(hipMalloc((void**)&d, t.size() * sizeof(RayRec)), "ray table");
                hipError_t ec = hipMemcpy(d, t.data(), t.size() * sizeof(RayRec), hipMemcpyHostToDevice);
                if (ec != hipSuccess) {
                    (void)hipFree(d);
                    return hip_fail(ec, "ray table upload");
                }
                p->d_rays = d;
            }
        }
        const RdConst c = rd_const(p->g, p->ndim);
        hipError_t er = p->ndim == 3 ? cbct_rd_launch(p->d_rays, view0, vstep, nsub, c, d_img, img_dt, d_sino, sino_dt, st)
                                     : fbct_rd_launch(p->d_rays, view0, vstep, nsub, c, d_img, img_dt, d_sino, sino_dt, st);
        if (er == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype combination not supported for proj,rd.");
        HIP_TRY(er, "ray-driven projection launch");
        return CTSIM_OK;
    }
    const int foot = k == OP_PROJ_DD ? 0 : 1;
    hipError_t e;
    if (p->compute == CTSIM_F32) {
        if (p->ndim == 3)
            e = cbct_proj_launch<float>(p->gf, p->d_vf + view0, vstep, nsub, d_img, img_dt, p->d_map, d_sino, sino_dt,
                                        foot, exhaustive, st);
        else
            e = fbct_proj_launch<float>(p->gf, p->d_vf + view0, vstep, nsub, d_img, img_dt, d_sino, sino_dt, foot,
                                        exhaustive, st);
    } else {
        if (p->ndim == 3)
            e = cbct_proj_launch<double>(p->gd, p->d_vd + view0, vstep, nsub, d_img, img_dt, p->d_map, d_sino, sino_dt,
                                         foot, exhaustive, st);
        else
            e = fbct_proj_launch<double>(p->gd, p->d_vd + view0, vstep, nsub, d_img, img_dt, d_sino, sino_dt, foot,
                                         exhaustive, st);
    }
    if (e == hipErrorInvalidValue)
        return fail(CTSIM_EUNSUP, "Error: dtype combination (img %d, sino %d) not supported for compute %d.", img_dt,
                    sino_dt, p->compute);
    if (e == hipErrorNotSupported)
        return fail(CTSIM_EUNSUP, "Error: 3D forward projection supports nu <= 32767 and nz <= 8192 (nu = %d, nz = %d).",
                    p->g.nu, p->g.nz);
    HIP_TRY(e, "forward projection launch");
    return CTSIM_OK;
}

int ctsim_plan_proj(const ctsim_plan* p, const char* op, int32_t view0, int32_t vstep, int32_t nsub, const void* d_img,
                    int32_t img_dt, void* d_sino, int32_t sino_dt, void* stream) {
    return plan_proj_impl(p, op, view0, vstep, nsub, d_img, img_dt, d_sino, sino_dt, stream, 0);
}

int ctsim_plan_proj_exhaustive(const ctsim_plan* p, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                               const void* d_img, int32_t img_dt, void* d_sino, int32_t sino_dt, void* stream) {
    return plan_proj_impl(p, op, view0, vstep, nsub, d_img, img_dt, d_sino, sino_dt, stream, 1);
}

int ctsim_plan_back(const ctsim_plan* p, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                    const void* d_sino, int32_t sino_dt, void* d_img, int32_t img_dt, double scale, int32_t accumulate,
                    void* stream) {
    if (!p) return fail(CTSIM_EINVAL, "Error: null plan.");
    const OpKind k = parse_op(op);
    if (k != OP_BACK_DD && k != OP_BACK_TF && k != OP_BACK_PD && k != OP_BACK_HPD)
        return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    if (p->ndim == 2 && k == OP_BACK_HPD) return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    if (int rc = check_views(p, view0, vstep, nsub)) return rc;
    if (!d_img || !d_sino) return fail(CTSIM_EINVAL, "Error: null device buffer.");
    if (!valid_dtype(img_dt) || !valid_dtype(sino_dt)) return fail(CTSIM_EINVAL, "Error: bad dtype.");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    const int kind = k == OP_BACK_DD ? 0 : k == OP_BACK_TF ? 1 : 2;
    const int off_boundary = k == OP_BACK_PD ? 1 : 0;
    if (p->compute == CTSIM_F32) {
        if (p->ndim == 3) {
            if (kind == 2)
                e = cbct_back_pd_launch<float>(p->gf, p->d_vf + view0, vstep, nsub, d_sino, sino_dt, p->d_map, d_img,
                                               img_dt, off_boundary, (float)scale, accumulate, st);
            else
                e = cbct_back_launch<float>(p->gf, p->d_vf + view0, vstep, nsub, d_sino, sino_dt, p->d_map, d_img,
                                            img_dt, kind, (float)scale, accumulate, st);
        } else {
            if (accumulate) return fail(CTSIM_EUNSUP, "Error: 2D backprojection does not accumulate.");
            e = fbct_back_launch<float>(p->gf, p->d_vf + view0, vstep, nsub, d_sino, sino_dt, d_img, img_dt, kind,
                                        (float)scale, st);
        }
    } else {
        if (p->ndim == 3) {
            if (kind == 2)
                e = cbct_back_pd_launch<double>(p->gd, p->d_vd + view0, vstep, nsub, d_sino, sino_dt, p->d_map, d_img,
                                                img_dt, off_boundary, scale, accumulate, st);
            else
                e = cbct_back_launch<double>(p->gd, p->d_vd + view0, vstep, nsub, d_sino, sino_dt, p->d_map, d_img,
                                             img_dt, kind, scale, accumulate, st);
        } else {
            if (accumulate) return fail(CTSIM_EUNSUP, "Error: 2D backprojection does not accumulate.");
            e = fbct_back_launch<double>(p->gd, p->d_vd + view0, vstep, nsub, d_sino, sino_dt, d_img, img_dt, kind,
                                         scale, st);
        }
    }
    if (e == hipErrorInvalidValue)
        return fail(CTSIM_EUNSUP, "Error: dtype combination (sino %d, img %d) not supported for compute %d.", sino_dt,
                    img_dt, p->compute);
    if (e == hipErrorNotSupported)
        return fail(CTSIM_EUNSUP,
                    "Error: 3D backprojection needs 64 views of the detector under 4 GiB (nu = %d, nv = %d).",
                    p->g.nu, p->g.nv);
    HIP_TRY(e, "backprojection launch");
    return CTSIM_OK;
}

int ctsim_plan_set_schedule(ctsim_plan* p, int32_t proj_sched, int32_t back_sched) {
    if (!p) return fail(CTSIM_EINVAL, "Error: null plan.");
    if (((proj_sched >> 2) & 3) == 3 || (proj_sched >> 4) != 0 || back_sched < 0 || back_sched > 3)
        return fail(CTSIM_EINVAL, "Error: unknown schedule (proj %d, back %d).", proj_sched, back_sched);
    p->gf.sched_proj = p->gd.sched_proj = proj_sched;
    p->gf.sched_back = p->gd.sched_back = back_sched;
    return CTSIM_OK;
}

int ctsim_plan_get_schedule(const ctsim_plan* p, int32_t* proj_sched, int32_t* back_sched) {
    if (!p || !proj_sched || !back_sched) return fail(CTSIM_EINVAL, "Error: null argument.");
    *proj_sched = p->compute == CTSIM_F32 ? p->gf.sched_proj : p->gd.sched_proj;
    *back_sched = p->compute == CTSIM_F32 ? p->gf.sched_back : p->gd.sched_back;
    return CTSIM_OK;
}

int ctsim_plan_autotune(ctsim_plan* p, const char* op, int32_t view0, int32_t vstep, int32_t nsub, const void* d_in,
                        int32_t in_dt, void* d_out, int32_t out_dt, int32_t reps, int32_t* best, float* ms_out,
                        int32_t* sched_out, int32_t max_out, void* stream) {
    if (!p || !op || !d_in || !d_out || !best) return fail(CTSIM_EINVAL, "Error: null argument.");
    const OpKind k = parse_op(op);
    const bool proj = k == OP_PROJ_DD || k == OP_PROJ_TF;
    const bool pd = k == OP_BACK_PD || k == OP_BACK_HPD;
    const bool back = k == OP_BACK_DD || k == OP_BACK_TF || pd;
    if (p->ndim != 3 || (!proj && !back))
        return fail(CTSIM_EINVAL, "Error: the 3D autotuner takes proj,dd proj,tf back,dd back,tf back,pd back,hpd.");
    if (d_in == d_out) return fail(CTSIM_EINVAL, "Error: the tuning output must not alias the input.");
    if (reps < 1) reps = 1;
    // the forward schedules (form | split << 2): automatic, LDS ring with whole and with shared
    // items, register pipeline with whole and with shared items, the ring with two row chunks
    // per item (whole items); the backward column orders
    // (DD / SF) or the FDK backprojection's staged and direct forms (pd / hpd), each keeping
    // the other bit of the current schedule
    static const int kProj[6] = {0, 1 | (1 << 2), 1 | (2 << 2), 2 | (1 << 2), 2 | (2 << 2), 3 | (1 << 2)};
    const int keep_p = p->gf.sched_proj, keep_b = p->gf.sched_back;
    const int kBack[2] = {pd ? (keep_b & 1) : (keep_b & 2), pd ? (keep_b & 1) | 2 : (keep_b & 2) | 1};
    const int n = proj ? 6 : 2;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0), "event");
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(he, "event");
    }
    int rc = CTSIM_OK, pick = -1;
    float best_ms = 3.4e38f;
    for (int i = 0; i < n && rc == CTSIM_OK; ++i) {
        const int sp = proj ? kProj[i] : keep_p, sb = proj ? keep_b : kBack[i];
        rc = ctsim_plan_set_schedule(p, sp, sb);
        auto run = [&]() {
            return proj ? ctsim_plan_proj(p, op, view0, vstep, nsub, d_in, in_dt, d_out, out_dt, stream)
                        : ctsim_plan_back(p, op, view0, vstep, nsub, d_in, in_dt, d_out, out_dt, 1.0, 0, stream);
        };
        if (rc == CTSIM_OK) rc = run();  // warm-up
        if (rc) break;
        if ((he = hipEventRecord(e0, st)) != hipSuccess) { rc = hip_fail(he, "event"); break; }
        for (int r = 0; r < reps && rc == CTSIM_OK; ++r) rc = run();
        if (rc) break;
        float ms = 0.f;
        if ((he = hipEventRecord(e1, st)) != hipSuccess || (he = hipEventSynchronize(e1)) != hipSuccess ||
            (he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
            rc = hip_fail(he, "event timing");
            break;
        }
        ms /= reps;
        if (ms_out && i < max_out) ms_out[i] = ms;
        if (sched_out && i < max_out) sched_out[i] = proj ? sp : sb;
        if (ms < best_ms) {
            best_ms = ms;
            pick = i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != CTSIM_OK || pick < 0) {
        (void)ctsim_plan_set_schedule(p, keep_p, keep_b);
        return rc ? rc : fail(CTSIM_EINVAL, "Error: no schedule ran.");
    }
    *best = proj ? kProj[pick] : kBack[pick];
    return ctsim_plan_set_schedule(p, proj ? kProj[pick] : keep_p, proj ? keep_b : kBack[pick]);
}

int ctsim_plan_back_nesterov(const ctsim_plan* p, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                             const void* d_sino, int32_t sino_dt, double scale, const float* x, const float* z,
                             float* s, const float* precom, float* x_out, float* z_out, float* zcopy,
                             const float* zref, double beta, double delta, int32_t pfun, double t, double t_next,
                             const double* d_tseq, const int32_t* d_base, int32_t k, void* stream) {
    if (!p) return fail(CTSIM_EINVAL, "Error: null plan.");
    const OpKind kind = parse_op(op);
    if (kind != OP_BACK_DD && kind != OP_BACK_TF) return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    if (p->ndim != 3 || p->compute != CTSIM_F32 || pfun != 0)
        return fail(CTSIM_EUNSUP, "Error: the fused update is 3D, FP32, Huber penalty.");
    if (int rc = check_views(p, view0, vstep, nsub)) return rc;
    if (!d_sino || !x || !z || !s || !precom || !x_out || !z_out) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (x_out == x || z_out == z || x_out == z || z_out == x)
        return fail(CTSIM_EINVAL, "Error: outputs must not alias the stencil inputs.");
    if (!valid_dtype(sino_dt)) return fail(CTSIM_EINVAL, "Error: bad dtype.");
    NestArgs na;
    na.x = x; na.z = z; na.s = s; na.precom = precom; na.x_out = x_out; na.z_out = z_out;
    na.zcopy = zcopy; na.zref = zref; na.beta = beta; na.delta = (float)delta;
    na.t = t; na.t_next = t_next; na.tseq = d_tseq; na.base = (const int*)d_base; na.k = k;
    hipError_t e = cbct_back_nesterov_launch(p->gf, p->d_vf + view0, vstep, nsub, d_sino, sino_dt, p->d_map,
                                             kind == OP_BACK_DD ? 0 : 1, (float)scale, na, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: sinogram dtype %d not supported.", sino_dt);
    if (e == hipErrorNotSupported)
        return fail(CTSIM_EUNSUP, "Error: the fused update needs a detector of at most 64 rows and nz <= 512.");
    HIP_TRY(e, "fused backprojection and update");
    return CTSIM_OK;
}

int ctsim_xyz_to_zxy(const void* src, int32_t sdt, void* dst, int32_t ddt, int32_t nx, int32_t ny, int32_t nz,
                     void* stream) {
    if (!src || !dst || nx <= 0 || ny <= 0 || nz <= 0) return fail(CTSIM_EINVAL, "Error: bad layout arguments.");
    hipError_t e = xyz_zxy_launch(src, sdt, dst, ddt, nx * ny, nz, true, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype pair not supported.");
    HIP_TRY(e, "layout conversion");
    return CTSIM_OK;
}

int ctsim_zxy_to_xyz(const void* src, int32_t sdt, void* dst, int32_t ddt, int32_t nx, int32_t ny, int32_t nz,
                     void* stream) {
    if (!src || !dst || nx <= 0 || ny <= 0 || nz <= 0) return fail(CTSIM_EINVAL, "Error: bad layout arguments.");
    hipError_t e = xyz_zxy_launch(src, sdt, dst, ddt, nx * ny, nz, false, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype pair not supported.");
    HIP_TRY(e, "layout conversion");
    return CTSIM_OK;
}

// ---------------------------------------------------------------------------------
// MEX-signature gateways
// ---------------------------------------------------------------------------------
}  // extern "C"

namespace {

// device buffers reused across gateway calls (grown on demand, freed at exit)
struct Pool {
    void* buf[3] = {nullptr, nullptr, nullptr};
    size_t cap[3] = {0, 0, 0};
    int dev[3] = {-1, -1, -1};
    ~Pool() {
        for (int i = 0; i < 3; ++i)
            if (buf[i]) (void)hipFree(buf[i]);
    }
    hipError_t get(int i, size_t bytes, void** out) {
        int d = 0;
        hipError_t e = hipGetDevice(&d);
        if (e != hipSuccess) return e;
        if (cap[i] < bytes || dev[i] != d) {
            if (buf[i]) (void)hipFree(buf[i]);
            buf[i] = nullptr;
            cap[i] = 0;
            e = hipMalloc(&buf[i], bytes);
            if (e != hipSuccess) return e;
            cap[i] = bytes;
            dev[i] = d;
        }
        *out = buf[i];
        return hipSuccess;
    }
};
Pool g_pool;
std::mutex g_gateway_mutex;  // the MEX gateway is not re-entrant; neither is this one

struct PlanGuard {
    ctsim_plan* p = nullptr;
    ~PlanGuard() { ctsim_plan_destroy(p); }
};

int gateway(int ndim, const char* op, const ctsim_geom& g, int32_t noviews, const double* betas, const float* couchz,
            const uint8_t* map, const float* in, float* out) {
    const OpKind k = parse_op(op);
    if (k == OP_BAD || (ndim == 2 && k == OP_BACK_HPD)) return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    if (!in || !out) return fail(CTSIM_EINVAL, "Error: null data pointer.");
    if (noviews <= 0 || !betas) return fail(CTSIM_EINVAL, "Error: mx_betas must have noviews X 1 double array.");
    // cbct_geom_mex's 16-argument (map) family has no ray-driven projector
    // (geom/cbct_os/cbct_geom_mex.c:282-286): the call fails there
    if (ndim == 3 && k == OP_PROJ_RD && map) return fail(CTSIM_EINVAL, "Error: cbct_geom_mex() failed.");
    std::lock_guard<std::mutex> lock(g_gateway_mutex);
    PlanGuard pg;
    if (int rc = ctsim_plan_create(&pg.p, ndim, &g, noviews, betas, couchz, map, CTSIM_F32)) return rc;
    const bool proj = k == OP_PROJ_DD || k == OP_PROJ_TF || k == OP_PROJ_RD;
    const size_t nimg = (size_t)g.nx * g.ny * g.nz;
    const size_t nsino = (size_t)g.nu * g.nv * noviews;
    void *d_in = nullptr, *d_out = nullptr, *d_tmp = nullptr;
    HIP_TRY(g_pool.get(0, (proj ? nimg : nsino) * 4, &d_in), "device buffer");
    HIP_TRY(g_pool.get(1, (proj ? nsino : nimg) * 4, &d_out), "device buffer");
    if (ndim == 3) HIP_TRY(g_pool.get(2, nimg * 4, &d_tmp), "device buffer");
    hipStream_t st = nullptr;
    HIP_TRY(hipMemcpy(d_in, in, (proj ? nimg : nsino) * 4, hipMemcpyHostToDevice), "upload");
    int rc;
    if (proj) {
        const void* vol = d_in;
        if (ndim == 3) {
            HIP_TRY(xyz_zxy_launch(d_in, DT_F32, d_tmp, DT_F32, g.nx * g.ny, g.nz, true, st), "layout");
            vol = d_tmp;
        }
        rc = ctsim_plan_proj(pg.p, op, 0, 1, noviews, vol, CTSIM_F32, d_out, CTSIM_F32, st);
        if (rc) return rc;
        HIP_TRY(hipMemcpy(out, d_out, nsino * 4, hipMemcpyDeviceToHost), "download");
    } else {
        void* vol = ndim == 3 ? d_tmp : d_out;
        rc = ctsim_plan_back(pg.p, op, 0, 1, noviews, d_in, CTSIM_F32, vol, CTSIM_F32, 1.0, 0, st);
        if (rc) return rc;
        if (ndim == 3) HIP_TRY(xyz_zxy_launch(d_tmp, DT_F32, d_out, DT_F32, g.nx * g.ny, g.nz, false, st), "layout");
        HIP_TRY(hipMemcpy(out, d_out, nimg * 4, hipMemcpyDeviceToHost), "download");
    }
    return CTSIM_OK;
}

ctsim_geom geom3_from_mex(const int32_t* nxyz, const float* dxyz, const float* cxyz, const int32_t* nuv,
                          const float* duv, const float* cuv, float sad, float add, float sdd, int32_t dettype) {
    ctsim_geom g;
    g.nx = nxyz[0]; g.ny = nxyz[1]; g.nz = nxyz[2];
    g.dx = dxyz[0]; g.dy = dxyz[1]; g.dz = dxyz[2];
    g.ox = cxyz[0]; g.oy = cxyz[1]; g.oz = cxyz[2];
    g.nu = nuv[0]; g.nv = nuv[1];
    g.du = duv[0]; g.dv = duv[1];
    g.ou = cuv[0]; g.ov = cuv[1];
    g.sad = sad; g.add = add; g.sdd = sdd;
    g.detector_type = dettype;
    return g;
}

int cbct_gateway_restricted(const char* allowed0, const char* allowed1, const char* op, const int32_t* nxyz,
                            const float* dxyz, const float* cxyz, const int32_t* nuv, const float* duv,
                            const float* cuv, float sad, float add, float sdd, int32_t noviews, const double* betas,
                            const float* couchz, int32_t dettype, const float* in, const uint8_t* map, float* out,
                            char* errbuf, size_t errlen) {
    int rc;
    if (!op || (std::strcmp(op, allowed0) && std::strcmp(op, allowed1)))
        rc = fail(CTSIM_EINVAL, "Error: unknown geom type.");
    else if (!nxyz || !dxyz || !cxyz || !nuv || !duv || !cuv || !couchz || !map)
        rc = fail(CTSIM_EINVAL, "Error: incorrect number of arguments.");
    else
        rc = gateway(3, op, geom3_from_mex(nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, dettype), noviews, betas,
                     couchz, map, in, out);
    if (rc) copy_err(errbuf, errlen);
    return rc;
}

}  // namespace

extern "C" {

int ctsim_fbct_geom_mex(const char* op, const int32_t nxy[2], const float dxy[2], const float oxy[2], int32_t nu,
                        float du, float ou, float sad, float add, float sdd, int32_t noviews, const double* betas,
                        int32_t dettype, const float* in, float* out, char* errbuf, size_t errlen) {
    int rc;
    if (!nxy || !dxy || !oxy) {
        rc = fail(CTSIM_EINVAL, "Error: incorrect number of arguments.");
    } else {
        ctsim_geom g;
        g.nx = nxy[0]; g.ny = nxy[1]; g.nz = 1;
        g.dx = dxy[0]; g.dy = dxy[1]; g.dz = 1.f;
        g.ox = oxy[0]; g.oy = oxy[1]; g.oz = 0.f;
        g.nu = nu; g.nv = 1; g.du = du; g.dv = 1.f; g.ou = ou; g.ov = 0.f;
        g.sad = sad; g.add = add; g.sdd = sdd; g.detector_type = dettype;
        rc = gateway(2, op, g, noviews, betas, nullptr, nullptr, in, out);
    }
    if (rc) copy_err(errbuf, errlen);
    return rc;
}

int ctsim_cbct_geom_mex(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                        const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add, float sdd,
                        int32_t noviews, const double* betas, const float* couchz, int32_t dettype, const float* in,
                        const uint8_t* map, float* out, char* errbuf, size_t errlen) {
    int rc;
    if (!nxyz || !dxyz || !cxyz || !nuv || !duv || !cuv || !couchz)
        rc = fail(CTSIM_EINVAL, "Error: incorrect number of arguments.");
    else
        rc = gateway(3, op, geom3_from_mex(nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, dettype), noviews, betas,
                     couchz, map, in, out);
    if (rc) copy_err(errbuf, errlen);
    return rc;
}

int ctsim_cbct_dd_cuda(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                       const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add, float sdd,
                       int32_t noviews, const double* betas, const float* couchz, int32_t dettype, const float* in,
                       const uint8_t* map, float* out, char* errbuf, size_t errlen) {
    return cbct_gateway_restricted("proj,dd", "back,dd", op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews,
                                   betas, couchz, dettype, in, map, out, errbuf, errlen);
}

int ctsim_cbct_sf_cuda(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                       const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add, float sdd,
                       int32_t noviews, const double* betas, const float* couchz, int32_t dettype, const float* in,
                       const uint8_t* map, float* out, char* errbuf, size_t errlen) {
    return cbct_gateway_restricted("proj,tf", "back,tf", op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews,
                                   betas, couchz, dettype, in, map, out, errbuf, errlen);
}

int ctsim_cbct_back_pd_cuda(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                            const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add,
                            float sdd, int32_t noviews, const double* betas, const float* couchz, int32_t dettype,
                            const float* in, const uint8_t* map, float* out, char* errbuf, size_t errlen) {
    return cbct_gateway_restricted("back,pd", "back,hpd", op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews,
                                   betas, couchz, dettype, in, map, out, errbuf, errlen);
}

}  // extern "C"

namespace {

// device buffers of one PMat gateway call, freed on every return path
struct DevBufs {
    void* p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    ~DevBufs() {
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
};

float f32_at_least(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}
float f32_at_most(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}

int pmat_gateway(const char* op, const int32_t* nxyz, const float* dxyz, const float* cxyz, const int32_t* nuv,
                 float du, float dv, int32_t noviews, const float* pmat, const float* cameras, const float* sdds,
                 const float* cuvs, const float* in, const uint8_t* map, float* out) {
    if (!op) return fail(CTSIM_EINVAL, "The 1st argument is invalid!");
    int kind;
    const bool proj = !std::strncmp(op, "proj", 4);
    if (!std::strcmp(op, "proj,rd"))
        return fail(CTSIM_EINVAL, "Error: ray driven method is not supported. Please consider using distance driven "
                                  "method.");
    else if (!std::strcmp(op, "proj,dd") || !std::strcmp(op, "proj,tf")) kind = PMAT_PROJ_DD;
    else if (!std::strcmp(op, "back,dd") || !std::strcmp(op, "back,tf")) kind = PMAT_BACK_DD;
    else if (!std::strcmp(op, "back,pd")) kind = PMAT_BACK_PD;
    else if (!std::strcmp(op, "back,hpd")) kind = PMAT_BACK_HPD;
    else return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    if (!nxyz || !dxyz || !cxyz || !nuv) return fail(CTSIM_EINVAL, "Error: incorrect number of arguments.");
    const int nx = nxyz[0], ny = nxyz[1], nz = nxyz[2], nu = nuv[0], nv = nuv[1];
    if (nx <= 0 || ny <= 0 || nz <= 0 || nu <= 0 || nv <= 0 || noviews <= 0)
        return fail(CTSIM_EINVAL, "Error: sizes must be positive.");
    if (!in)
        return fail(CTSIM_EINVAL, proj ? "Error: mx_image must have [nx ny nz] float array."
                                       : "Error: mx_proj must have [nv nu noviews] float array.");
    if (!pmat) return fail(CTSIM_EINVAL, "Error: mx_pmat must have [3 4 noViews ] float array.");
    if (!cameras) return fail(CTSIM_EINVAL, "Error: mx_cameras must have noviews X 2 float array.");
    if (!cuvs) return fail(CTSIM_EINVAL, "Error: mx_cuvs must have noviews X 2 float array.");
    if (!sdds) return fail(CTSIM_EINVAL, "Error: mx_sdds must have noviews X 1 float array.");
    if (!map) return fail(CTSIM_EINVAL, "Error: mx_map must have [nx ny] bool array.");
    if (!out) return fail(CTSIM_EINVAL, "Error: null data pointer.");
    if (nz > 512) return fail(CTSIM_EUNSUP, "Error: the PMat operators take at most 512 slices (nz = %d).", nz);
    if ((kind == PMAT_BACK_PD || kind == PMAT_BACK_HPD) && (nu < 2 || nv < 2))
        return fail(CTSIM_EUNSUP, "Error: back,pd needs at least 2 x 2 detector cells.");
    PmatConst c;
    c.nx = nx; c.ny = ny; c.nz = nz; c.nu = nu; c.nv = nv; c.nviews = noviews;
    c.dx = dxyz[0]; c.dy = dxyz[1]; c.dz = dxyz[2]; c.du = du; c.dv = dv;
    // float wx = - ( nx - 1.0 ) / 2.0 + offset_x; wu, wv without offsets (cbct_proj_pmat.c:34-38)
    c.wx = (float)(-(nx - 1.0) / 2.0 + (double)cxyz[0]);
    c.wy = (float)(-(ny - 1.0) / 2.0 + (double)cxyz[1]);
    c.wz = (float)(-(nz - 1.0) / 2.0 + (double)cxyz[2]);
    c.wu = (float)(-(nu - 1.0) / 2.0);
    c.wv = (float)(-(nv - 1.0) / 2.0);
    c.lo = f32_at_least(1e-6);
    c.u_hi = f32_at_most(nu - 1 - 1e-6);
    c.v_hi = f32_at_most(nv - 1 - 1e-6);
    std::vector<PmatView> views((size_t)noviews);
    for (int b = 0; b < noviews; ++b) {
        PmatView& v = views[(size_t)b];
        std::memcpy(v.pm, pmat + 12 * (size_t)b, sizeof v.pm);
        v.cx = cameras[3 * (size_t)b];
        v.cy = cameras[3 * (size_t)b + 1];
        v.sdd = sdds[b];
        v.ou = cuvs[2 * (size_t)b];
    }
    std::lock_guard<std::mutex> lock(g_gateway_mutex);
    const size_t nimg = (size_t)nx * ny * nz, nsino = (size_t)noviews * nu * nv;
    DevBufs d;
    // views, map, input, output (MATLAB layout), ZXY volume, amplitude-scaled sinogram (back,dd)
    HIP_TRY(hipMalloc(&d.p[0], views.size() * sizeof(PmatView)), "device buffer");
    HIP_TRY(hipMalloc(&d.p[1], (size_t)nx * ny), "device buffer");
    HIP_TRY(hipMalloc(&d.p[2], (proj ? nimg : nsino) * 4), "device buffer");
    HIP_TRY(hipMalloc(&d.p[3], (proj ? nsino : nimg) * 4), "device buffer");
    HIP_TRY(hipMalloc(&d.p[4], nimg * 4), "device buffer");
    if (kind == PMAT_BACK_DD) HIP_TRY(hipMalloc(&d.p[5], nsino * 4), "device buffer");
    HIP_TRY(hipMemcpy(d.p[0], views.data(), views.size() * sizeof(PmatView), hipMemcpyHostToDevice), "upload");
    HIP_TRY(hipMemcpy(d.p[1], map, (size_t)nx * ny, hipMemcpyHostToDevice), "upload");
    HIP_TRY(hipMemcpy(d.p[2], in, (proj ? nimg : nsino) * 4, hipMemcpyHostToDevice), "upload");
    hipStream_t st = nullptr;
    const PmatView* dview = (const PmatView*)d.p[0];
    const uint8_t* dmap = (const uint8_t*)d.p[1];
    if (proj) {
        HIP_TRY(xyz_zxy_launch(d.p[2], DT_F32, d.p[4], DT_F32, nx * ny, nz, true, st), "layout");
        HIP_TRY(pmat_launch(kind, c, dview, dmap, (const float*)d.p[4], (float*)d.p[3], nullptr, st),
                "PMat projection");
        HIP_TRY(hipMemcpy(out, d.p[3], nsino * 4, hipMemcpyDeviceToHost), "download");
    } else {
        HIP_TRY(pmat_launch(kind, c, dview, dmap, (const float*)d.p[2], (float*)d.p[4], (float*)d.p[5], st),
                "PMat backprojection");
        HIP_TRY(xyz_zxy_launch(d.p[4], DT_F32, d.p[3], DT_F32, nx * ny, nz, false, st), "layout");
        HIP_TRY(hipMemcpy(out, d.p[3], nimg * 4, hipMemcpyDeviceToHost), "download");
    }
    return CTSIM_OK;
}

}  // namespace

extern "C" {

int ctsim_cbct_geom_pmat_mex(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                             const int32_t nuv[2], float du, float dv, int32_t noviews, const float* pmat,
                             const float* cameras, const float* sdds, const float* cuvs, const float* in,
                             const uint8_t* map, float* out, char* errbuf, size_t errlen) {
    const int rc = pmat_gateway(op, nxyz, dxyz, cxyz, nuv, du, dv, noviews, pmat, cameras, sdds, cuvs, in, map, out);
    if (rc) copy_err(errbuf, errlen);
    return rc;
}

// one entry point per (dimension, operator): the gateway with the op string fixed
#define CTSIM_PER_OP(NAME, NDIM, OPSTR)                                                                            \
    int NAME(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz, const uint8_t* map,   \
             const float* in, float* out, char* errbuf, size_t errlen) {                                           \
        int rc;                                                                                                    \
        if (!g) {                                                                                                  \
            rc = fail(CTSIM_EINVAL, "Error: null geometry.");                                                      \
        } else {                                                                                                   \
            ctsim_geom gg = *g;                                                                                    \
            if (NDIM == 2) {                                                                                       \
                gg.nz = 1; gg.nv = 1; gg.dz = 1.f; gg.dv = 1.f; gg.oz = 0.f; gg.ov = 0.f;                           \
            }                                                                                                      \
            rc = gateway(NDIM, OPSTR, gg, nviews, betas, NDIM == 3 ? couchz : nullptr, NDIM == 3 ? map : nullptr,  \
                         in, out);                                                                                 \
        }                                                                                                          \
        if (rc) copy_err(errbuf, errlen);                                                                          \
        return rc;                                                                                                 \
    }

CTSIM_PER_OP(ctsim_fbct_proj_rd, 2, "proj,rd")
CTSIM_PER_OP(ctsim_fbct_proj_dd, 2, "proj,dd")
CTSIM_PER_OP(ctsim_fbct_proj_tf, 2, "proj,tf")
CTSIM_PER_OP(ctsim_fbct_back_dd, 2, "back,dd")
CTSIM_PER_OP(ctsim_fbct_back_tf, 2, "back,tf")
CTSIM_PER_OP(ctsim_fbct_back_pd, 2, "back,pd")
CTSIM_PER_OP(ctsim_cbct_proj_rd, 3, "proj,rd")
CTSIM_PER_OP(ctsim_cbct_proj_dd, 3, "proj,dd")
CTSIM_PER_OP(ctsim_cbct_proj_tf, 3, "proj,tf")
CTSIM_PER_OP(ctsim_cbct_back_dd, 3, "back,dd")
CTSIM_PER_OP(ctsim_cbct_back_tf, 3, "back,tf")
CTSIM_PER_OP(ctsim_cbct_back_pd, 3, "back,pd")
CTSIM_PER_OP(ctsim_cbct_back_hpd, 3, "back,hpd")
#undef CTSIM_PER_OP

// ---------------------------------------------------------------------------------
// batched uneven 2D slices
// ---------------------------------------------------------------------------------
}  // extern "C"

struct ctsim_batch2 {
    int nslices = 0;
    int compute = CTSIM_F32;
    int nitems_proj = 0, nitems_back = 0;
    long long img_elems = 0, sino_elems = 0;
    void* d_slices_proj = nullptr;  // Slice2 table, item0 counted for the forward work list
    void* d_slices_back = nullptr;  // same slices, item0 counted for the back work list
    std::vector<void*> d_views;
    int proj_split = 1;  // waves per forward work item (strip split)
    int back_split = 4;  // waves per backprojection work item (view split)
};

namespace {
template <class T>
int batch2_build(ctsim_batch2* b, int n, const ctsim_geom* geoms, const int32_t* nviews, const double* const* betas) {
    std::vector<Slice2<T>> sl(n), sp(n), sb(n);
    long long img = 0, sino = 0;
    int ip = 0, ib = 0;
    for (int i = 0; i < n; ++i) {
        const ctsim_geom& g = geoms[i];
        if (int rc = check_geom(2, g)) return rc;
        if (nviews[i] <= 0 || !betas[i]) return fail(CTSIM_EINVAL, "Error: slice %d has no views.", i);
        auto v = make_views<T>(g, nviews[i], betas[i], nullptr);
        void* dv = nullptr;
        HIP_TRY(hipMalloc(&dv, v.size() * sizeof(View<T>)), "batch views");
        b->d_views.push_back(dv);
        HIP_TRY(hipMemcpy(dv, v.data(), v.size() * sizeof(View<T>), hipMemcpyHostToDevice), "batch views");
        Slice2<T> s{};
        s.g = make_geom<T>(g);
        s.views = (const View<T>*)dv;
        s.nviews = nviews[i];
        s.img_off = img;
        s.sino_off = sino;
        sl[i] = s;
        img += (long long)g.nx * g.ny;
        sino += (long long)g.nu * nviews[i];
    }
    // work lists with the slices of the longest items first (a forward item's strip grows with
    // the image, a back item's view loop with the view count), so that the launch ends on
    // short items; a slice's place in the list changes no output (each keeps its offsets)
    std::vector<int> po(n), bo(n);
    for (int i = 0; i < n; ++i) po[i] = bo[i] = i;
    std::stable_sort(po.begin(), po.end(), [&](int a, int c) {
        return std::max(geoms[a].nx, geoms[a].ny) > std::max(geoms[c].nx, geoms[c].ny);
    });
    std::stable_sort(bo.begin(), bo.end(), [&](int a, int c) { return nviews[a] > nviews[c]; });
    for (int j = 0; j < n; ++j) {
        const int i = po[j];
        sp[j] = sl[i];
        sp[j].item0 = ip;
        ip += nviews[i] * ((geoms[i].nu + 63) / 64);
    }
    for (int j = 0; j < n; ++j) {
        const int i = bo[j];
        sb[j] = sl[i];
        sb[j].item0 = ib;
        ib += geoms[i].ny * ((geoms[i].nx + 63) / 64);
    }
    b->nitems_proj = ip;
    b->nitems_back = ib;
    b->proj_split = ip < kFewItems2D ? 4 : 1;  // the single-slice forward's rule
    b->img_elems = img;
    b->sino_elems = sino;
    HIP_TRY(hipMalloc(&b->d_slices_proj, n * sizeof(Slice2<T>)), "batch table");
    HIP_TRY(hipMalloc(&b->d_slices_back, n * sizeof(Slice2<T>)), "batch table");
    HIP_TRY(hipMemcpy(b->d_slices_proj, sp.data(), n * sizeof(Slice2<T>), hipMemcpyHostToDevice), "batch table");
    HIP_TRY(hipMemcpy(b->d_slices_back, sb.data(), n * sizeof(Slice2<T>), hipMemcpyHostToDevice), "batch table");
    return CTSIM_OK;
}
}  // namespace

extern "C" {

int ctsim_batch2_create(ctsim_batch2** out, int32_t nslices, const ctsim_geom* geoms, const int32_t* nviews,
                        const double* const* betas, int32_t compute) {
    if (!out || nslices <= 0 || !geoms || !nviews || !betas) return fail(CTSIM_EINVAL, "Error: bad batch arguments.");
    *out = nullptr;
    if (compute != CTSIM_F32 && compute != CTSIM_F64) return fail(CTSIM_EINVAL, "Error: bad compute precision.");
    ctsim_batch2* b = new (std::nothrow) ctsim_batch2;
    if (!b) return fail(CTSIM_ENOMEM, "Error: host allocation failed.");
    b->nslices = nslices;
    b->compute = compute;
    const int rc = compute == CTSIM_F32 ? batch2_build<float>(b, nslices, geoms, nviews, betas)
                                        : batch2_build<double>(b, nslices, geoms, nviews, betas);
    if (rc) {
        ctsim_batch2_destroy(b);
        return rc;
    }
    *out = b;
    return CTSIM_OK;
}

int ctsim_batch2_destroy(ctsim_batch2* b) {
    if (!b) return CTSIM_OK;
    for (void* p : b->d_views) (void)hipFree(p);
    if (b->d_slices_proj) (void)hipFree(b->d_slices_proj);
    if (b->d_slices_back) (void)hipFree(b->d_slices_back);
    delete b;
    return CTSIM_OK;
}

int ctsim_batch2_sizes(const ctsim_batch2* b, int64_t* img_elems, int64_t* sino_elems) {
    if (!b) return fail(CTSIM_EINVAL, "Error: null batch.");
    if (img_elems) *img_elems = b->img_elems;
    if (sino_elems) *sino_elems = b->sino_elems;
    return CTSIM_OK;
}

int ctsim_batch2_proj(const ctsim_batch2* b, const char* op, const void* d_img, int32_t img_dt, void* d_sino,
                      int32_t sino_dt, void* stream) {
    if (!b || !d_img || !d_sino) return fail(CTSIM_EINVAL, "Error: null argument.");
    const OpKind k = parse_op(op);
    if (k != OP_PROJ_DD && k != OP_PROJ_TF) return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    const int kind = k == OP_PROJ_DD ? 0 : 1;
    hipError_t e = b->compute == CTSIM_F32
                       ? fbct_proj_batch_launch<float>(b->d_slices_proj, b->nslices, b->nitems_proj, d_img, img_dt,
                                                       d_sino, sino_dt, kind, b->proj_split, (hipStream_t)stream)
                       : fbct_proj_batch_launch<double>(b->d_slices_proj, b->nslices, b->nitems_proj, d_img, img_dt,
                                                        d_sino, sino_dt, kind, b->proj_split, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype combination not supported.");
    HIP_TRY(e, "batched forward projection");
    return CTSIM_OK;
}

int ctsim_batch2_back(const ctsim_batch2* b, const char* op, const void* d_sino, int32_t sino_dt, void* d_img,
                      int32_t img_dt, void* stream) {
    if (!b || !d_img || !d_sino) return fail(CTSIM_EINVAL, "Error: null argument.");
    const OpKind k = parse_op(op);
    if (k != OP_BACK_DD && k != OP_BACK_TF && k != OP_BACK_PD) return fail(CTSIM_EINVAL, "Error: unknown geom type.");
    const int kind = k == OP_BACK_DD ? 0 : k == OP_BACK_TF ? 1 : 2;
    hipError_t e = b->compute == CTSIM_F32
                       ? fbct_back_batch_launch<float>(b->d_slices_back, b->nslices, b->nitems_back, d_sino, sino_dt,
                                                       d_img, img_dt, kind, b->back_split, (hipStream_t)stream)
                       : fbct_back_batch_launch<double>(b->d_slices_back, b->nslices, b->nitems_back, d_sino, sino_dt,
                                                        d_img, img_dt, kind, b->back_split, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype combination not supported.");
    HIP_TRY(e, "batched backprojection");
    return CTSIM_OK;
}

static bool valid_split(int s) { return s == 1 || s == 2 || s == 4; }

int ctsim_batch2_set_schedule(ctsim_batch2* b, int32_t proj_split, int32_t back_split) {
    if (!b) return fail(CTSIM_EINVAL, "Error: null batch.");
    if (!valid_split(proj_split) || !valid_split(back_split))
        return fail(CTSIM_EINVAL, "Error: schedule splits must be 1, 2 or 4.");
    b->proj_split = proj_split;
    b->back_split = back_split;
    return CTSIM_OK;
}

int ctsim_batch2_get_schedule(const ctsim_batch2* b, int32_t* proj_split, int32_t* back_split) {
    if (!b) return fail(CTSIM_EINVAL, "Error: null batch.");
    if (proj_split) *proj_split = b->proj_split;
    if (back_split) *back_split = b->back_split;
    return CTSIM_OK;
}

int ctsim_batch2_autotune(ctsim_batch2* b, const char* proj_op, const char* back_op, const void* d_img,
                          int32_t img_dt, void* d_sino, int32_t sino_dt, void* d_img_out, int32_t reps, float* ms_out,
                          void* stream) {
    if (!b || !d_img || !d_sino || !d_img_out) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (d_img_out == d_img) return fail(CTSIM_EINVAL, "Error: the tuning output must not alias the input image.");
    if (reps < 1) reps = 1;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0), "event");
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(he, "event");
    }
    const int splits[3] = {1, 2, 4};
    int rc = CTSIM_OK;
    float best[2] = {1e30f, 1e30f};
    int pick[2] = {b->proj_split, b->back_split};
    const int keep[2] = {b->proj_split, b->back_split};
    for (int pass = 0; pass < 2 && rc == CTSIM_OK; ++pass) {
        for (int i = 0; i < 3 && rc == CTSIM_OK; ++i) {
            if (pass == 0) b->proj_split = splits[i];
            else b->back_split = splits[i];
            auto run = [&]() {
                return pass == 0 ? ctsim_batch2_proj(b, proj_op, d_img, img_dt, d_sino, sino_dt, stream)
                                 : ctsim_batch2_back(b, back_op, d_sino, sino_dt, d_img_out, img_dt, stream);
            };
            rc = run();  // warm-up, and the sinogram the back pass reads
            if (rc) break;
            if ((he = hipEventRecord(e0, st)) != hipSuccess) { rc = hip_fail(he, "event"); break; }
            for (int r = 0; r < reps && rc == CTSIM_OK; ++r) rc = run();
            if (rc) break;
            float ms = 0.f;
            if ((he = hipEventRecord(e1, st)) != hipSuccess || (he = hipEventSynchronize(e1)) != hipSuccess ||
                (he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
                rc = hip_fail(he, "event timing");
                break;
            }
            ms /= reps;
            if (ms < best[pass]) {
                best[pass] = ms;
                pick[pass] = splits[i];
            }
        }
        if (pass == 0) b->proj_split = pick[0];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        b->proj_split = keep[0];
        b->back_split = keep[1];
        return rc;
    }
    b->proj_split = pick[0];
    b->back_split = pick[1];
    if (ms_out) {
        ms_out[0] = best[0];
        ms_out[1] = best[1];
    }
    return CTSIM_OK;
}

// ---------------------------------------------------------------------------------
// memory, streams, events, graphs
// ---------------------------------------------------------------------------------
int ctsim_malloc(void** p, size_t bytes) {
    if (!p) return fail(CTSIM_EINVAL, "Error: null pointer.");
    *p = nullptr;
    if (bytes == 0) return CTSIM_OK;
    HIP_TRY(hipMalloc(p, bytes), "hipMalloc");
    return CTSIM_OK;
}
int ctsim_free(void* p) {
    if (p) HIP_TRY(hipFree(p), "hipFree");
    return CTSIM_OK;
}
int ctsim_memcpy_h2d(void* d, const void* h, size_t bytes, void* stream) {
    if (!bytes) return CTSIM_OK;
    HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "h2d copy");
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "h2d copy");
    return CTSIM_OK;
}
int ctsim_memcpy_d2h(void* h, const void* d, size_t bytes, void* stream) {
    if (!bytes) return CTSIM_OK;
    HIP_TRY(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "d2h copy");
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "d2h copy");
    return CTSIM_OK;
}
int ctsim_memcpy_d2d(void* d, const void* s, size_t bytes, void* stream) {
    if (!bytes) return CTSIM_OK;
    HIP_TRY(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "d2d copy");
    return CTSIM_OK;
}
int ctsim_fill(void* p, int32_t dt, size_t n, double value, void* stream) {
    if (!p && n) return fail(CTSIM_EINVAL, "Error: null pointer.");
    hipError_t e = fill_launch(p, dt, n, value, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: bad dtype.");
    HIP_TRY(e, "fill");
    return CTSIM_OK;
}
int ctsim_convert(const void* s, int32_t sdt, void* d, int32_t ddt, size_t n, void* stream) {
    hipError_t e = convert_launch(s, sdt, d, ddt, n, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype pair not supported.");
    HIP_TRY(e, "convert");
    return CTSIM_OK;
}
int ctsim_stream_create(void** s) {
    if (!s) return fail(CTSIM_EINVAL, "Error: null pointer.");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream create");
    *s = st;
    return CTSIM_OK;
}
int ctsim_stream_destroy(void* s) {
    if (s) HIP_TRY(hipStreamDestroy((hipStream_t)s), "stream destroy");
    return CTSIM_OK;
}
int ctsim_stream_synchronize(void* s) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)s), "stream synchronize");
    return CTSIM_OK;
}
int ctsim_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize(), "device synchronize");
    return CTSIM_OK;
}
int ctsim_event_create(void** ev) {
    if (!ev) return fail(CTSIM_EINVAL, "Error: null pointer.");
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreate(&e), "event create");
    *ev = e;
    return CTSIM_OK;
}
int ctsim_event_destroy(void* ev) {
    if (ev) HIP_TRY(hipEventDestroy((hipEvent_t)ev), "event destroy");
    return CTSIM_OK;
}
int ctsim_event_record(void* ev, void* s) {
    HIP_TRY(hipEventRecord((hipEvent_t)ev, (hipStream_t)s), "event record");
    return CTSIM_OK;
}
int ctsim_event_elapsed_ms(void* a, void* b, float* ms) {
    if (!ms) return fail(CTSIM_EINVAL, "Error: null pointer.");
    HIP_TRY(hipEventSynchronize((hipEvent_t)b), "event synchronize");
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "event elapsed");
    return CTSIM_OK;
}
int ctsim_graph_begin(void* s) {
    if (!s) return fail(CTSIM_EINVAL, "Error: graph capture needs a non-default stream.");
    HIP_TRY(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeThreadLocal), "begin capture");
    return CTSIM_OK;
}
int ctsim_graph_end(void* s, void** exec) {
    if (!exec) return fail(CTSIM_EINVAL, "Error: null pointer.");
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamEndCapture((hipStream_t)s, &graph), "end capture");
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(e, "graph instantiate");
    *exec = ge;
    return CTSIM_OK;
}
int ctsim_graph_launch(void* exec, void* s) {
    HIP_TRY(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)s), "graph launch");
    return CTSIM_OK;
}
int ctsim_graph_destroy(void* exec) {
    if (exec) HIP_TRY(hipGraphExecDestroy((hipGraphExec_t)exec), "graph destroy");
    return CTSIM_OK;
}
int ctsim_device_count(int32_t* n) {
    if (!n) return fail(CTSIM_EINVAL, "Error: null pointer.");
    int c = 0;
    HIP_TRY(hipGetDeviceCount(&c), "device count");
    *n = c;
    return CTSIM_OK;
}

// ---------------------------------------------------------------------------------
// PWLS building blocks
// ---------------------------------------------------------------------------------
int ctsim_pwls_residual(const void* p, int32_t pdt, const void* y, int32_t ydt, const void* w, int32_t wdt, void* r,
                        int32_t rdt, size_t n, double wscale, double* phi, void* stream) {
    if (!p || !y || !w || !r || !phi) return fail(CTSIM_EINVAL, "Error: null argument.");
    hipError_t e = residual_launch(p, pdt, y, ydt, w, wdt, r, rdt, n, wscale, phi, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype combination not supported.");
    HIP_TRY(e, "pwls residual");
    return CTSIM_OK;
}

int ctsim_pwls_residual_noalloc(const void* p, int32_t pdt, const void* y, int32_t ydt, const void* w, int32_t wdt,
                                void* r, int32_t rdt, size_t n, double wscale, double* phi, void* stream) {
    if (!p || !y || !w || !r || !phi) return fail(CTSIM_EINVAL, "Error: null argument.");
    hipError_t e = residual_launch_noalloc(p, pdt, y, ydt, w, wdt, r, rdt, n, wscale, phi, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype combination not supported.");
    HIP_TRY(e, "pwls residual (allocation-free)");
    return CTSIM_OK;
}

}  // extern "C"

template <class V>
static int penalty_impl(const V* x, int32_t nx, int32_t ny, int32_t nz, int32_t pfun, double delta, int32_t mode,
                        V* out, double* cost, void* stream) {
    if (!x || nx <= 0 || ny <= 0 || nz <= 0) return fail(CTSIM_EINVAL, "Error: bad volume.");
    if ((mode == 1 || mode == 2 || mode == 3) && !out) return fail(CTSIM_EINVAL, "Error: null output.");
    if (mode == 0 && !cost) return fail(CTSIM_EINVAL, "Error: null cost output.");
    if (pfun < 0 || pfun > 6) return fail(CTSIM_EINVAL, "Error: unknow penalty function");
    if (mode == 3 && pfun != 0) return fail(CTSIM_EINVAL, "Error: wrong penalty mode.");
    hipError_t e = penalty_launch(x, nx, ny, nz, pfun, delta, mode, out, cost, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: wrong penalty mode.");
    HIP_TRY(e, "penalty");
    return CTSIM_OK;
}

extern "C" {

int ctsim_penalty(const float* x, int32_t nx, int32_t ny, int32_t nz, int32_t pfun, double delta, int32_t mode,
                  float* out, double* cost, void* stream) {
    return penalty_impl<float>(x, nx, ny, nz, pfun, delta, mode, out, cost, stream);
}

int ctsim_penalty_f64(const double* x, int32_t nx, int32_t ny, int32_t nz, int32_t pfun, double delta, int32_t mode,
                      double* out, double* cost, void* stream) {
    return penalty_impl<double>(x, nx, ny, nz, pfun, delta, mode, out, cost, stream);
}

int ctsim_huber(const float* x, int32_t nx, int32_t ny, int32_t nz, double delta, int32_t mode, float* out,
                double* cost, void* stream) {
    return penalty_impl<float>(x, nx, ny, nz, 0, delta, mode, out, cost, stream);
}

int ctsim_pwls_nesterov_update(const float* x, const float* z, float* s, const float* g, const float* precom,
                               float* x_out, float* z_out, void* zcopy, int32_t zcopy_dtype, const float* zref,
                               int32_t nx, int32_t ny, int32_t nz, double beta, double delta, int32_t pfun, double t,
                               double t_next, void* stream) {
    if (!x || !z || !s || !g || !precom || !x_out || !z_out) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (x_out == x || z_out == z || x_out == z || z_out == x)
        return fail(CTSIM_EINVAL, "Error: outputs must not alias the stencil inputs.");
    hipError_t e = nesterov_launch(x, z, s, g, precom, x_out, z_out, zcopy, zcopy_dtype, zref, nx, ny, nz, beta,
                                   delta, pfun, t, t_next, nullptr, nullptr, 0, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: z copy dtype or penalty not supported.");
    HIP_TRY(e, "nesterov update");
    return CTSIM_OK;
}

int ctsim_pwls_nesterov_update_seq(const float* x, const float* z, float* s, const float* g, const float* precom,
                                   float* x_out, float* z_out, void* zcopy, int32_t zcopy_dtype, const float* zref,
                                   int32_t nx, int32_t ny, int32_t nz, double beta, double delta, int32_t pfun,
                                   const double* d_tseq, const int32_t* d_base, int32_t k, void* stream) {
    if (!x || !z || !s || !g || !precom || !x_out || !z_out || !d_tseq || !d_base)
        return fail(CTSIM_EINVAL, "Error: null argument.");
    if (x_out == x || z_out == z || x_out == z || z_out == x)
        return fail(CTSIM_EINVAL, "Error: outputs must not alias the stencil inputs.");
    hipError_t e = nesterov_launch(x, z, s, g, precom, x_out, z_out, zcopy, zcopy_dtype, zref, nx, ny, nz, beta,
                                   delta, pfun, 1.0, 1.0, d_tseq, (const int*)d_base, k, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: z copy dtype or penalty not supported.");
    HIP_TRY(e, "nesterov update");
    return CTSIM_OK;
}

int ctsim_add_i32(int32_t* d_counter, int32_t increment, void* stream) {
    if (!d_counter) return fail(CTSIM_EINVAL, "Error: null argument.");
    HIP_TRY(add_i32_launch((int*)d_counter, increment, (hipStream_t)stream), "counter add");
    return CTSIM_OK;
}

int ctsim_pwls_sqs_update(const float* x, float* x_out, const float* g, const float* precom, int32_t nx, int32_t ny,
                          int32_t nz, double beta, double delta, int32_t pfun, void* stream) {
    if (!x || !x_out || !g || !precom) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (x == x_out) return fail(CTSIM_EINVAL, "Error: output must not alias the input.");
    hipError_t e = sqs_launch(x, x_out, g, precom, nx, ny, nz, beta, delta, pfun, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: unknow penalty function");
    HIP_TRY(e, "sqs update");
    return CTSIM_OK;
}

int ctsim_pwls_lalm_sg(const float* d_dl, float* d_g, float* d_s, size_t n, double rho, void* stream) {
    if ((!d_dl || !d_g || !d_s) && n) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (n) HIP_TRY(lalm_sg_launch(d_dl, d_g, d_s, n, rho, (hipStream_t)stream), "lalm g/s update");
    return CTSIM_OK;
}

int ctsim_pwls_lalm_inner(const float* x, const float* xold, const float* s, const float* precom, float* x_out,
                          int32_t nx, int32_t ny, int32_t nz, double beta, double delta, int32_t pfun, double rho,
                          void* stream) {
    if (!x || !xold || !s || !precom || !x_out) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (x == x_out) return fail(CTSIM_EINVAL, "Error: output must not alias the input.");
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(CTSIM_EINVAL, "Error: bad volume.");
    hipError_t e = lalm_inner_launch(x, xold, s, precom, x_out, nx, ny, nz, beta, delta, pfun, rho, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: unknow penalty function");
    HIP_TRY(e, "lalm update");
    return CTSIM_OK;
}

int ctsim_nusqs_seed(const float* x, float* v, int32_t nx, int32_t ny, int32_t nz, double eps, void* stream) {
    if (!x || !v || nx <= 0 || ny <= 0 || nz <= 0) return fail(CTSIM_EINVAL, "Error: bad volume.");
    if (x == v) return fail(CTSIM_EINVAL, "Error: d_v must not alias d_x.");
    HIP_TRY(nusqs_seed_launch(x, v, nx, ny, nz, eps, (hipStream_t)stream), "nusqs seed");
    return CTSIM_OK;
}

int ctsim_div(const float* a, const float* b, float* out, size_t n, void* stream) {
    if (!a || !b || !out) return fail(CTSIM_EINVAL, "Error: null buffer.");
    if (n == 0) return CTSIM_OK;
    HIP_TRY(div_launch(a, b, out, n, (hipStream_t)stream), "div");
    return CTSIM_OK;
}

int ctsim_absdiff_floor(const float* x, const float* xref, float* out, size_t n, double eps, void* stream) {
    if (!x || !xref || !out) return fail(CTSIM_EINVAL, "Error: null buffer.");
    if (n == 0) return CTSIM_OK;
    HIP_TRY(absdiff_floor_launch(x, xref, out, n, eps, (hipStream_t)stream), "absdiff floor");
    return CTSIM_OK;
}

int ctsim_dynamic_range(const float* x, size_t n, double t, double eta, float* out, void* stream) {
    if (!x || !out || n == 0) return fail(CTSIM_EINVAL, "Error: bad input.");
    HIP_TRY(dynamic_range_launch(x, n, t, eta, out, (hipStream_t)stream), "dynamic range");
    return CTSIM_OK;
}

int ctsim_pwls_nusqs_update(const float* x, float* x_out, const float* g, const float* dl, const float* h, int32_t nx,
                            int32_t ny, int32_t nz, double beta, double delta, int32_t pfun, void* stream) {
    if (!x || !x_out || !g || !dl || !h || nx <= 0 || ny <= 0 || nz <= 0) return fail(CTSIM_EINVAL, "Error: bad volume.");
    if (x == x_out) return fail(CTSIM_EINVAL, "Error: d_x_out must not alias d_x.");
    if (pfun < 0 || pfun > 6) return fail(CTSIM_EINVAL, "Error: unknow penalty function");
    HIP_TRY(nusqs_launch(x, x_out, g, dl, h, nx, ny, nz, beta, delta, pfun, (hipStream_t)stream), "nusqs update");
    return CTSIM_OK;
}

int ctsim_sqdiff_sum(const float* a, const float* b, size_t n, double* out, void* stream) {
    if (!a || !b || !out) return fail(CTSIM_EINVAL, "Error: null argument.");
    HIP_TRY(sqdiff_launch(a, b, n, out, (hipStream_t)stream), "sqdiff");
    return CTSIM_OK;
}

int ctsim_huber_f64(const double* x, int32_t nx, int32_t ny, int32_t nz, double delta, int32_t mode, double* out,
                    double* cost, void* stream) {
    return penalty_impl<double>(x, nx, ny, nz, 0, delta, mode, out, cost, stream);
}

int ctsim_pwls_nesterov_update_f64(const double* x, const double* z, double* s, const double* g,
                                   const double* precom, double* x_out, double* z_out, int32_t nx, int32_t ny,
                                   int32_t nz, double beta, double delta, int32_t pfun, double t, double t_next,
                                   void* stream) {
    if (!x || !z || !s || !g || !precom || !x_out || !z_out) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(CTSIM_EINVAL, "Error: bad volume.");
    if (x_out == x || z_out == z || x_out == z || z_out == x)
        return fail(CTSIM_EINVAL, "Error: outputs must not alias the stencil inputs.");
    hipError_t e = nesterov_launch(x, z, s, g, precom, x_out, z_out, nx, ny, nz, beta, delta, pfun, t, t_next,
                                   (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: unknow penalty function");
    HIP_TRY(e, "nesterov update");
    return CTSIM_OK;
}

int ctsim_sqdiff_sum_f64(const double* a, const double* b, size_t n, double* out, void* stream) {
    if (!a || !b || !out) return fail(CTSIM_EINVAL, "Error: null argument.");
    HIP_TRY(sqdiff_launch(a, b, n, out, (hipStream_t)stream), "sqdiff");
    return CTSIM_OK;
}

int ctsim_mul(const void* a, int32_t adt, const void* b, int32_t bdt, void* o, int32_t odt, size_t n, void* stream) {
    if (!a || !b || !o) return fail(CTSIM_EINVAL, "Error: null argument.");
    hipError_t e = mul_launch(a, adt, b, bdt, o, odt, n, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype combination not supported.");
    HIP_TRY(e, "mul");
    return CTSIM_OK;
}

int ctsim_sub(const float* a, const float* b, void* o, int32_t odt, size_t n, void* stream) {
    if (!a || !b || !o) return fail(CTSIM_EINVAL, "Error: null buffer.");
    hipError_t e = sub_launch(a, b, o, odt, n, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype not supported.");
    HIP_TRY(e, "sub");
    return CTSIM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// FDK pre-processing
// ---------------------------------------------------------------------------------
struct ctsim_fdk {
    FdkPlanImpl* impl = nullptr;
    int nviews = 0, f64 = 0;
};

extern "C" {

int ctsim_fdk_create(ctsim_fdk** out, const ctsim_geom* g, int32_t nviews, const double* H, int32_t order,
                     const float* view_weights, int32_t compute) {
    if (!out || !g || !H) return fail(CTSIM_EINVAL, "Error: null FDK argument.");
    *out = nullptr;
    if (g->nu <= 0 || g->nv <= 0 || nviews <= 0) return fail(CTSIM_EINVAL, "Error: detector sizes must be positive.");
    if (order < g->nu || (order & 1)) return fail(CTSIM_EINVAL, "Error: filter length must be even and >= nu.");
    if (compute != CTSIM_F32 && compute != CTSIM_F64)
        return fail(CTSIM_EINVAL, "Error: compute precision must be CTSIM_F32 or CTSIM_F64.");
    FdkGeom fg{g->du, g->dv, g->ou, g->ov, g->sad, g->sdd, g->detector_type == 1 ? 1 : 0};
    ctsim_fdk* f = new (std::nothrow) ctsim_fdk;
    if (!f) return fail(CTSIM_ENOMEM, "Error: host allocation failed.");
    f->nviews = nviews;
    f->f64 = compute == CTSIM_F64;
    // workspace of at most 1 GiB: views are filtered in chunks that fit it
    hipError_t e = fdk_plan_create(&f->impl, fg, g->nu, g->nv, nviews, H, order, view_weights, f->f64, size_t(1) << 30);
    if (e != hipSuccess) {
        delete f;
        return hip_fail(e, "FDK plan (device buffers or hipFFT plans)");
    }
    *out = f;
    return CTSIM_OK;
}

// the same plan for a fan-beam sinogram [nviews][nu] (one detector row): the cosine weight
// of reconFBP2d.m:32 (flat) / :38 (arc)
int ctsim_fbp2d_create(ctsim_fdk** out, const ctsim_geom* g, int32_t nviews, const double* H, int32_t order,
                       const float* view_weights, int32_t compute) {
    if (!out || !g || !H) return fail(CTSIM_EINVAL, "Error: null FBP argument.");
    *out = nullptr;
    if (g->nu <= 0 || nviews <= 0) return fail(CTSIM_EINVAL, "Error: detector sizes must be positive.");
    if (order < g->nu || (order & 1)) return fail(CTSIM_EINVAL, "Error: filter length must be even and >= nu.");
    if (compute != CTSIM_F32 && compute != CTSIM_F64)
        return fail(CTSIM_EINVAL, "Error: compute precision must be CTSIM_F32 or CTSIM_F64.");
    FdkGeom fg{g->du, 1.f, g->ou, 0.f, g->sad, g->sdd, g->detector_type == 1 ? 2 : 0};
    ctsim_fdk* f = new (std::nothrow) ctsim_fdk;
    if (!f) return fail(CTSIM_ENOMEM, "Error: host allocation failed.");
    f->nviews = nviews;
    f->f64 = compute == CTSIM_F64;
    hipError_t e = fdk_plan_create(&f->impl, fg, g->nu, 1, nviews, H, order, view_weights, f->f64, size_t(1) << 30);
    if (e != hipSuccess) {
        delete f;
        return hip_fail(e, "FBP plan (device buffers or hipFFT plans)");
    }
    *out = f;
    return CTSIM_OK;
}

int ctsim_fdk_destroy(ctsim_fdk* f) {
    if (!f) return CTSIM_OK;
    fdk_plan_destroy(f->impl);
    delete f;
    return CTSIM_OK;
}

int ctsim_fdk_filter(const ctsim_fdk* f, int32_t view0, int32_t nsub, const void* d_sino, int32_t sino_dtype,
                     void* d_out, int32_t out_dtype, void* stream) {
    if (!f || !d_sino || !d_out) return fail(CTSIM_EINVAL, "Error: null FDK argument.");
    if (nsub <= 0 || view0 < 0 || (long long)view0 + nsub > f->nviews)
        return fail(CTSIM_EINVAL, "Error: view range outside the FDK plan.");
    if (d_sino == d_out) return fail(CTSIM_EINVAL, "Error: the filter does not work in place.");
    hipError_t e = fdk_filter_launch(f->impl, view0, nsub, d_sino, sino_dtype, d_out, out_dtype, (hipStream_t)stream);
    if (e == hipErrorInvalidValue)
        return fail(CTSIM_EUNSUP, "Error: dtype combination (sino %d, out %d) not supported for this plan.", sino_dtype,
                    out_dtype);
    HIP_TRY(e, "FDK filter");
    return CTSIM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// data side: counts, log attenuation, PWLS weights
// ---------------------------------------------------------------------------------
extern "C" {

int ctsim_poisson_counts(const void* d_p, int32_t p_dtype, size_t n, double i0, uint64_t seed, float* d_counts,
                         void* stream) {
    if ((!d_p || !d_counts) && n) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (!(i0 > 0)) return fail(CTSIM_EINVAL, "Error: the source intensity must be positive.");
    if (n == 0) return CTSIM_OK;
    hipError_t e = poisson_launch(d_p, p_dtype, n, i0, (unsigned long long)seed, d_counts, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: line integrals must be F32 or F64.");
    HIP_TRY(e, "poisson counts");
    return CTSIM_OK;
}

int ctsim_log_attenuation(const float* d_counts, size_t n, double i0, double low_count, void* d_y, int32_t y_dtype,
                          void* stream) {
    if ((!d_counts || !d_y) && n) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (!(i0 > 0) || !(low_count > 0)) return fail(CTSIM_EINVAL, "Error: intensities must be positive.");
    if (n == 0) return CTSIM_OK;
    hipError_t e = neglog_launch(d_counts, n, i0, low_count, d_y, y_dtype, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: bad dtype.");
    HIP_TRY(e, "log attenuation");
    return CTSIM_OK;
}

int ctsim_log_attenuation_ex(const float* d_counts, size_t n, int32_t ncell, double source_intensity,
                             double low_count, int32_t compound, const float* d_flat_field,
                             const float* d_tube_profile, void* d_y, int32_t y_dtype, void* stream) {
    if ((!d_counts || !d_y) && n) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (!(source_intensity > 0) || !(low_count > 0)) return fail(CTSIM_EINVAL, "Error: intensities must be positive.");
    if (ncell <= 0 || n % (size_t)ncell) return fail(CTSIM_EINVAL, "Error: n must be a multiple of the cells per view.");
    if (n == 0) return CTSIM_OK;
    hipError_t e = neglog_ex_launch(d_counts, n, ncell, source_intensity, low_count, compound != 0, d_flat_field,
                                    d_tube_profile, d_y, y_dtype, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: bad dtype.");
    HIP_TRY(e, "log attenuation");
    return CTSIM_OK;
}

int ctsim_water_correction(const void* d_in, int32_t in_dtype, void* d_out, int32_t out_dtype, int32_t nv, int32_t nu,
                           int32_t nviews, const float* d_coeffs, int32_t order, int32_t per_channel, void* stream) {
    if (!d_in || !d_out || !d_coeffs) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (nv <= 0 || nu <= 0 || nviews < 0) return fail(CTSIM_EINVAL, "Error: bad sinogram size.");
    if (order < 1 || order > 16) return fail(CTSIM_EINVAL, "Error: polynomial order must be 1 .. 16.");
    const size_t n = (size_t)nv * nu * nviews;
    if (n == 0) return CTSIM_OK;
    hipError_t e = water_launch(d_in, in_dtype, d_out, out_dtype, n, nv, nu, d_coeffs, order, per_channel ? 1 : 0,
                                (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EUNSUP, "Error: dtype pair not supported.");
    HIP_TRY(e, "water correction");
    return CTSIM_OK;
}

int ctsim_pwls_weights(const float* d_counts, size_t n, double r, double v, void* d_w, int32_t w_dtype, void* stream) {
    if ((!d_counts || !d_w) && n) return fail(CTSIM_EINVAL, "Error: null argument.");
    if (n == 0) return CTSIM_OK;
    hipError_t e = weights_launch(d_counts, n, r, v, d_w, w_dtype, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(CTSIM_EINVAL, "Error: bad dtype.");
    HIP_TRY(e, "pwls weights");
    return CTSIM_OK;
}

}  // extern "C"
