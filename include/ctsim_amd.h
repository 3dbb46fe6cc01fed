/*
 * ctsim_amd.h - C ABI of the MI355X (gfx950) projector / backprojector library
 * libctsim_hip.so.  Plain C: pointers, sizes and status codes only.
 *
 * Two layers:
 *
 *  1. MEX-signature gateways.  Host pointers, MATLAB column-major layout, the exact
 *     argument lists of CTSim's MEX functions.  A mexFunction shim (see
 *     ctsim_amd/csrc/mex/) marshals mxArrays into these calls, so geom/shell's
 *     forwardProjectMex / backProjectMex, and through them reconFDK, reconFBP and the
 *     statsrecon/pwls ordered-subset loops, run unchanged.
 *
 *  2. A device-resident plan API for loops that keep volumes and sinograms in HBM
 *     (ordered-subset PWLS, FDK, benchmarks): geometry and the per-view table are
 *     uploaded once; operators take device pointers, storage dtypes and a stream.
 *
 * Every function returns CTSIM_OK (0) or a nonzero status; ctsim_last_error()
 * returns the message of the calling thread's last failure.  Gateways additionally
 * copy the message into (errbuf, errlen) when errbuf is not NULL.
 *
 * Units and conventions follow the reference: lengths in mm, offsets in cell units,
 * projections scaled by dx/10 (attenuation in 1/cm), detector_type 1 = flat panel,
 * anything else = arc.
 */
#ifndef CTSIM_AMD_H
#define CTSIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CTSIM_OK 0
#define CTSIM_EINVAL 1   /* bad argument: op string, size, class, count */
#define CTSIM_EDEVICE 2  /* HIP runtime or launch failure */
#define CTSIM_ENOMEM 3   /* device allocation failed */
#define CTSIM_EUNSUP 4   /* valid request this build does not implement */

/* storage dtypes for the plan API */
#define CTSIM_F32 0
#define CTSIM_F16 1
#define CTSIM_BF16 2
#define CTSIM_F64 3

const char* ctsim_last_error(void);
int ctsim_version(void); /* 100 * major + minor */

/* ------------------------------------------------------------------------------
 * 1. MEX-signature gateways (host pointers, MATLAB layout, caller-allocated output
 *    that the callee overwrites completely, inputs never modified).
 * ---------------------------------------------------------------------------- */

/* Replaces fbct_geom_mex (geom/fbct_os/fbct_geom_mex.c:250-326), the 14-argument 2D
 * fan-beam gateway called by geom/shell/forwardProjectMex.m:52-55 and
 * geom/shell/backProjectMex.m:145-148.
 * op: "proj,rd" "proj,dd" "proj,tf" "back,dd" "back,tf" "back,pd".
 * in:  image [nx ny] for proj, sinogram [nu noviews] for back (float32).
 * out: sinogram [nu noviews] for proj, image [nx ny] for back (float32). */
int ctsim_fbct_geom_mex(const char* op, const int32_t nxy[2], const float dxy[2], const float oxy[2], int32_t nu,
                        float du, float ou, float sad, float add, float sdd, int32_t noviews, const double* betas,
                        int32_t detector_type, const float* in, float* out, char* errbuf, size_t errlen);

/* Replaces cbct_geom_mex (geom/cbct_os/cbct_geom_mex.c:516-600), the 3D cone-beam
 * gateway called by geom/shell/forwardProjectMex.m:82-85 and
 * geom/shell/backProjectMex.m:181-184.  map == NULL is the 15-argument (no map)
 * family, otherwise the 16-argument family with logical map [nx ny] (one byte per
 * column).  op: "proj,rd" (15-argument family only; with a map the call fails as the
 * reference's does), "proj,dd" "proj,tf" "back,dd" "back,tf" "back,pd" "back,hpd".
 * in:  image [nx ny nz] for proj, sinogram [nv nu noviews] for back (float32).
 * out: sinogram [nv nu noviews] for proj, image [nx ny nz] for back (float32). */
int ctsim_cbct_geom_mex(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                        const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add,
                        float sdd, int32_t noviews, const double* betas, const float* couchz,
                        int32_t detector_type, const float* in, const uint8_t* map, float* out, char* errbuf,
                        size_t errlen);

/* Same arguments as ctsim_cbct_geom_mex, restricted to the ops of the reference's
 * CUDA gateways, which take the 16-argument list only:
 *   cbct_dd_cuda      (geom/dd_cuda/cbct_dd_cuda.cu:21-299)        "proj,dd" "back,dd"
 *   cbct_sf_cuda      (geom/sf_cuda/cbct_sf_cuda.cu)               "proj,tf" "back,tf"
 *   cbct_back_pd_cuda (geom/fdk_cuda/cbct_back_pd_cuda.cu:21-193)  "back,pd" "back,hpd" */
int ctsim_cbct_dd_cuda(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                       const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add, float sdd,
                       int32_t noviews, const double* betas, const float* couchz, int32_t detector_type,
                       const float* in, const uint8_t* map, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_sf_cuda(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                       const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add, float sdd,
                       int32_t noviews, const double* betas, const float* couchz, int32_t detector_type,
                       const float* in, const uint8_t* map, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_back_pd_cuda(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                            const int32_t nuv[2], const float duv[2], const float cuv[2], float sad, float add,
                            float sdd, int32_t noviews, const double* betas, const float* couchz,
                            int32_t detector_type, const float* in, const uint8_t* map, float* out, char* errbuf,
                            size_t errlen);

/* Projection-matrix (PMat / C-arm) cone beam: geom/pcbct/cbct_geom_pmat_mex.c:260-309
 * (13-argument gateway: op, nxyz, dxyz, cxyz, nuv, duv, noviews, pmat, cameras, sdds,
 * cuvs, data, map).  pmat [3 4 noviews] single (column-major, u/v/w rows of each view),
 * cameras [3 noviews] (x, y used), sdds [noviews], cuvs [2 noviews] (the u offset is the
 * one used, for both detector coordinates, as in cbct_proj_pmat.c:197-205), map [nx ny]
 * required.  du, dv are the detector pitches: the reference reads duv[1] and duv[2]
 * (cbct_geom_pmat_mex.c:49-50); the MEX shim passes those for a three-element duv and
 * duv[0], duv[1] for a two-element one.  ops proj,dd | proj,tf (runs dd, as the
 * reference does after its warning) | back,dd | back,tf (dd) | back,pd | back,hpd;
 * proj,rd is refused with the reference's message.  FP32, at most 512 slices.  proj,dd
 * sums voxel footprints in on-chip patches and adds them to the sinogram with FP32
 * atomics, so its last bits can vary from run to run; the backprojectors gather (one
 * writer per voxel). */
int ctsim_cbct_geom_pmat_mex(const char* op, const int32_t nxyz[3], const float dxyz[3], const float cxyz[3],
                             const int32_t nuv[2], float du, float dv, int32_t noviews, const float* pmat,
                             const float* cameras, const float* sdds, const float* cuvs, const float* in,
                             const uint8_t* map, float* out, char* errbuf, size_t errlen);

/* Geometry of one operator; mirrors the fields of CTSim's geom struct that the
 * gateways receive (reconSize, reconSpacing, reconOffset, detSize, detSpacing,
 * detOffset, SAD, ADD, SDD, flatPanel; parameters/loadProjectionGeometryCT.m).
 * 2D plans use nz = nv = 1. */
typedef struct ctsim_geom {
    int32_t nx, ny, nz;
    float dx, dy, dz;
    float ox, oy, oz;
    int32_t nu, nv;
    float du, dv;
    float ou, ov;
    float sad, add, sdd;
    int32_t detector_type;
} ctsim_geom;

/* One entry point per (dimension, operator), the gateways above with the op string
 * fixed and the geometry in a struct: the reference's per-operator workers
 * fbct_{proj,back}_{dd,tf} and fbct_back_pd (geom/fbct_os/fbct_proj.c:155-466,
 * geom/fbct_os/fbct_back.c:82-399), cbct_proj1_{dd,tf} and cbct_back1_{dd,tf,pd}
 * (geom/cbct_os/cbct_proj.c:427-642,912-1169; geom/cbct_os/cbct_back.c:135-262,
 * 479-692,945-1197), as selected by the gateways' op-string dispatch
 * (geom/fbct_os/fbct_geom_mex.c:250-326, geom/cbct_os/cbct_geom_mex.c:516-600).
 * Same layouts and ownership as the gateways: host pointers in MATLAB order, out
 * allocated by the caller and overwritten completely, in never modified.  2D entry
 * points ignore nz, nv, dz, dv, oz, ov, couchz and map; 3D ones take couchz [nviews]
 * (NULL = zeros) and map [nx ny] (NULL = all columns). */
int ctsim_fbct_proj_dd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
/* Ray-driven forward projection, op "proj,rd" (geom/fbct_os/fbct_proj.c:3-153,
 * geom/cbct_os/cbct_proj.c:10-198): rays marched in steps of dx/2 with bilinear / trilinear
 * interpolation, in the reference's float arithmetic.  The 3D form takes no map: with a
 * map it fails ("Error: cbct_geom_mex() failed."), as the reference's 16-argument
 * gateway does (geom/cbct_os/cbct_geom_mex.c:282-286). */
int ctsim_fbct_proj_rd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_proj_rd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_fbct_proj_tf(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_fbct_back_dd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_fbct_back_tf(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_fbct_back_pd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_proj_dd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_proj_tf(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_back_dd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_back_tf(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_back_pd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                       const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);
int ctsim_cbct_back_hpd(const ctsim_geom* g, int32_t nviews, const double* betas, const float* couchz,
                        const uint8_t* map, const float* in, float* out, char* errbuf, size_t errlen);

/* ------------------------------------------------------------------------------
 * 2. Device-resident plan API.
 * ---------------------------------------------------------------------------- */

typedef struct ctsim_plan ctsim_plan;

/* ndim 2 or 3.  betas[nviews] (radians, double), couchz[nviews] (mm, may be NULL =
 * zeros), map [nx ny] host bytes (may be NULL = all columns active).  compute is
 * CTSIM_F32 (FP32 footprints and accumulation) or CTSIM_F64 (everything in double). */
int ctsim_plan_create(ctsim_plan** plan, int32_t ndim, const ctsim_geom* g, int32_t nviews, const double* betas,
                      const float* couchz, const uint8_t* map, int32_t compute);
int ctsim_plan_destroy(ctsim_plan* plan);
/* Launch schedules of a 3D plan (no counterpart in the reference, whose CUDA path uses
 * fixed 16 x 16 blocks, geom/dd_cuda/cbct_dd_proj.c:441-442).  proj_sched = form | split << 2:
 * form 0 automatic, 1 LDS-ring forward kernel (FP32 volumes of at most 256 slices, FP16
 * volumes of an even number of at most 256), 2 register-pipeline forward kernel, 3 the ring
 * kernel with two 64-row chunks per item and eight-cell tiles (FP32 volumes of at most 256
 * slices, one wave per item; others take the automatic form); split 0
 * automatic (a block's waves share a work item below 16384 items), 1 one wave per item,
 * 2 shared items.  back_sched bit 0 (DD / SF backprojection): 0 four consecutive image
 * columns per block, 1 2 x 2 column patches in 32 x 32 super-tiles (both with a contiguous
 * range per XCD); bit 1 (FDK pd / hpd, FP32 plans): 0 the LDS-staged form (detectors of at
 * most 1024 rows; taller ones take the direct form), 1 the direct form.  A new plan has
 * (0, 1).  Every schedule is deterministic; the forward forms and splits, and the two FDK
 * forms, add terms in different orders, so their results agree to FP32 rounding, not bit
 * for bit. */
int ctsim_plan_set_schedule(ctsim_plan* plan, int32_t proj_sched, int32_t back_sched);
int ctsim_plan_get_schedule(const ctsim_plan* plan, int32_t* proj_sched, int32_t* back_sched);

/* The project's autotuner for 3D plans: times every forward schedule (op proj,dd or
 * proj,tf: automatic, ring whole, ring shared, register whole, register shared, ring with two
 * row chunks), every
 * backprojector column order (back,dd or back,tf) or both FDK forms (back,pd or back,hpd)
 * on the caller's view selection and
 * buffers with HIP events on the stream (one warm-up, then reps timed launches each),
 * keeps the fastest in the plan and returns it in *best.  d_out is overwritten; it must not
 * alias d_in.  ms_out / sched_out (may be NULL) receive each candidate's mean time and
 * schedule code, at most max_out of them. */
int ctsim_plan_autotune(ctsim_plan* plan, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                        const void* d_in, int32_t in_dtype, void* d_out, int32_t out_dtype, int32_t reps,
                        int32_t* best, float* ms_out, int32_t* sched_out, int32_t max_out, void* stream);


/* Forward projection of the views view0, view0 + vstep, ... (nsub of them), the
 * ordered-subset selection of geom/shell/forwardProjectMex.m:35-39.
 * op "proj,dd", "proj,tf" or "proj,rd" (FP32 plans only; the plan's map is not used: the
 * ray table is built on the plan's first "proj,rd").  d_img: 3D volume in device ZXY order (z fastest,
 * vol[(iy*nx+ix)*nz+iz]) or 2D image [nx ny]; d_sino: [nsub][nu][nv] (v fastest),
 * every element written. */
int ctsim_plan_proj(const ctsim_plan* plan, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                    const void* d_img, int32_t img_dtype, void* d_sino, int32_t sino_dtype, void* stream);

/* Backprojection of the same view selection.  op "back,dd" "back,tf" "back,pd"
 * "back,hpd".  out = scale * A^T sino (+ out when accumulate != 0); scale carries the
 * ordered-subset factor M of geom/shell/backProjectMex.m:191-193. */
int ctsim_plan_back(const ctsim_plan* plan, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                    const void* d_sino, int32_t sino_dtype, void* d_img, int32_t img_dtype, double scale,
                    int32_t accumulate, void* stream);

/* Backprojection of a view selection fused with the Nesterov-SQS subset update of
 * statsrecon/pwls/reconPwlsSeNesterovSqs.m:67-87 (Huber penalty, penalty/huberPenalty.m):
 * the gradient scale * A^T sino of each voxel column goes straight into the update of its
 * voxels, as ctsim_pwls_nesterov_update (or _seq, when d_tseq is not null) computes it from
 * a gradient volume, without storing that volume.  3D FP32 plans, op "back,dd" or
 * "back,tf", pfun 0 (Huber), detectors of at most 64 rows; CTSIM_EUNSUP otherwise (the
 * caller then runs ctsim_plan_back and the update).  zcopy (FP32) receives z' - zref (z'
 * when zref is null) unless null.  No counterpart in the reference, where the
 * backprojection and the update are separate MATLAB statements. */
int ctsim_plan_back_nesterov(const ctsim_plan* plan, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                             const void* d_sino, int32_t sino_dtype, double scale, const float* x, const float* z,
                             float* s, const float* precom, float* x_out, float* z_out, float* zcopy,
                             const float* zref, double beta, double delta, int32_t pfun, double t, double t_next,
                             const double* d_tseq, const int32_t* d_base, int32_t k, void* stream);

/* Exhaustive-candidate variant of ctsim_plan_proj used by tests: every image column
 * is considered for every detector tile (no strip pruning); results must be
 * identical to ctsim_plan_proj. */
int ctsim_plan_proj_exhaustive(const ctsim_plan* plan, const char* op, int32_t view0, int32_t vstep, int32_t nsub,
                               const void* d_img, int32_t img_dtype, void* d_sino, int32_t sino_dtype,
                               void* stream);

/* Volume layout changes between MATLAB order [nx ny nz] and device ZXY order, with
 * dtype conversion (src and dst dtypes may differ; F64 only to F64). */
int ctsim_xyz_to_zxy(const void* d_src, int32_t src_dtype, void* d_dst, int32_t dst_dtype, int32_t nx, int32_t ny,
                     int32_t nz, void* stream);
int ctsim_zxy_to_xyz(const void* d_src, int32_t src_dtype, void* d_dst, int32_t dst_dtype, int32_t nx, int32_t ny,
                     int32_t nz, void* stream);

/* ------------------------------------------------------------------------------
 * Batched uneven 2D slices: nslices independent fan-beam slices, each with its own
 * geometry and views, projected (or backprojected) in one launch.  Images and
 * sinograms of all slices are packed back to back in d_img / d_sino in slice order
 * ([nx ny] and [nu nviews] each).
 * ---------------------------------------------------------------------------- */
typedef struct ctsim_batch2 ctsim_batch2;
int ctsim_batch2_create(ctsim_batch2** batch, int32_t nslices, const ctsim_geom* geoms, const int32_t* nviews,
                        const double* const* betas, int32_t compute);
int ctsim_batch2_destroy(ctsim_batch2* batch);
/* element counts of the packed image / sinogram buffers */
int ctsim_batch2_sizes(const ctsim_batch2* batch, int64_t* img_elems, int64_t* sino_elems);
int ctsim_batch2_proj(const ctsim_batch2* batch, const char* op, const void* d_img, int32_t img_dtype, void* d_sino,
                      int32_t sino_dtype, void* stream);
int ctsim_batch2_back(const ctsim_batch2* batch, const char* op, const void* d_sino, int32_t sino_dtype,
                      void* d_img, int32_t img_dtype, void* stream);
/* Launch schedule of the batch: each forward work item (slice, view, 64-cell tile)
 * splits its pixel strip over proj_split wavefronts, each backprojection item (slice,
 * row, 64-pixel block) splits its views over back_split wavefronts (1, 2 or 4; shares
 * are added in a fixed order).  Default 1 and 4. */
int ctsim_batch2_set_schedule(ctsim_batch2* batch, int32_t proj_split, int32_t back_split);
int ctsim_batch2_get_schedule(const ctsim_batch2* batch, int32_t* proj_split, int32_t* back_split);
/* Autotuner: times every schedule of the forward pass (d_img -> d_sino) and then of the
 * backprojection (d_sino -> d_img_out, which must not alias d_img) with HIP events,
 * reps launches each after a warm-up, keeps the fastest of each in the batch and
 * returns their times in ms_out[2] (may be NULL).  Overwrites d_sino and d_img_out. */
int ctsim_batch2_autotune(ctsim_batch2* batch, const char* proj_op, const char* back_op, const void* d_img,
                          int32_t img_dtype, void* d_sino, int32_t sino_dtype, void* d_img_out, int32_t reps,
                          float* ms_out, void* stream);

/* ------------------------------------------------------------------------------
 * Device memory, streams, events and graphs (thin HIP wrappers so that host code in
 * any language can drive the library through this ABI alone).
 * ---------------------------------------------------------------------------- */
int ctsim_malloc(void** d_ptr, size_t bytes);
int ctsim_free(void* d_ptr);
int ctsim_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, void* stream);
int ctsim_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, void* stream);
int ctsim_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream);
/* fills with a kernel (never hipMemsetAsync), so it is safe inside graph capture */
int ctsim_fill(void* d_ptr, int32_t dtype, size_t count, double value, void* stream);
int ctsim_convert(const void* d_src, int32_t src_dtype, void* d_dst, int32_t dst_dtype, size_t count, void* stream);
int ctsim_stream_create(void** stream);
int ctsim_stream_destroy(void* stream);
int ctsim_stream_synchronize(void* stream);
int ctsim_device_synchronize(void);
int ctsim_event_create(void** event);
int ctsim_event_destroy(void* event);
int ctsim_event_record(void* event, void* stream);
int ctsim_event_elapsed_ms(void* start, void* stop, float* ms);
int ctsim_graph_begin(void* stream);
int ctsim_graph_end(void* stream, void** graph_exec);
int ctsim_graph_launch(void* graph_exec, void* stream);
int ctsim_graph_destroy(void* graph_exec);
int ctsim_device_count(int32_t* count);

/* ------------------------------------------------------------------------------
 * PWLS ordered-subset building blocks (statsrecon/pwls/reconPwlsSeNesterovSqs.m:59-137,
 * statsrecon/pwls/reconPwlsSeSqs.m:51-110, penalty/huberPenalty.m:101-161).  Volumes
 * are device ZXY float32 (nx, ny, nz) unless stated; sinogram-side arrays carry a
 * dtype.  Reductions accumulate in FP64 and write one double to d_out.
 * ---------------------------------------------------------------------------- */

/* r = w .* (p - y) written as dtype r_dtype, and phi += wscale * sum(w .* (p - y).^2) / 2
 * in FP64 at d_phi (statsrecon/pwls/reconPwlsSeNesterovSqs.m:65,95-96).  wscale lets
 * half-precision storage keep w / max(w): the true weights are wscale * w.  FP32
 * arithmetic for F32/F16/BF16 operands; FP64 arithmetic when all four are F64 (F64
 * mixed with other dtypes is rejected). */
int ctsim_pwls_residual(const void* d_p, int32_t p_dtype, const void* d_y, int32_t y_dtype, const void* d_w,
                        int32_t w_dtype, void* d_r, int32_t r_dtype, size_t n, double wscale, double* d_phi,
                        void* stream);

/* ctsim_pwls_residual without any device allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as ctsim_pwls_residual.  The reduction's scratch (1024 FP64 partial
 * sums and a counter, 8448 bytes) is a fixed device array, one per GPU, held until the
 * process ends; the first call on a device finds its address and clears the counter on
 * the caller's stream.  The reduction grid never exceeds the array's 1024 partial sums,
 * so every n is accepted.  No hipMalloc, hipFree or allocating library call is made.
 * Calls that overlap on one GPU (other streams, other threads) share the array and
 * race: serialise them. */
int ctsim_pwls_residual_noalloc(const void* d_p, int32_t p_dtype, const void* d_y, int32_t y_dtype,
                                const void* d_w, int32_t w_dtype, void* d_r, int32_t r_dtype, size_t n,
                                double wscale, double* d_phi, void* stream);

/* Roughness penalties of a ZXY volume (penalty/loadPenaltyOperator.m:22-53): pfun
 * CTSIM_PEN_HUBER (huberPenalty.m:101-161), CTSIM_PEN_QUAD (quadPenalty.m; T is the
 * constant 4 + 4/sqrt2 (+1 in 3D) of the reference), CTSIM_PEN_HYPER (hyperbolaPenalty.m,
 * psi = delta^2 (sqrt(1 + (d/delta)^2) - 1)), CTSIM_PEN_ANISO (anisotropicPenalty.m:79-142,
 * c = exp(-d^2/delta^2), S adds d c), all over the neighbours +-x, +-y weight 1, xy
 * diagonals 1/sqrt2, +-z weight 1/2; CTSIM_PEN_ISOHUBER (huberIsoPenalty.m:73-104, axis
 * neighbours, weight 1); CTSIM_PEN_TV and CTSIM_PEN_ISOTV (totalVariationPenalty.m:53-92,
 * totalVariationIsoPenalty.m:53-92: c = 1/sqrt(sum d^2 + 1e-8) on the grid cropped by one,
 * x/y differences, plus z for the isotropic form of a 3D volume; delta unused).
 * mode 1 gradient S(x), mode 2 curvature T(x), written to d_out; mode 0 cost R(x)
 * accumulated in FP64 into d_cost; mode 3 (CTSIM_PEN_HUBER only) the curvature H of
 * huberPenalty.m:163-185, sum_k w_k [|x_i - x_k| <= delta] (x_i + x_k) / x_i, used by NU-SQS. */
#define CTSIM_PEN_HUBER 0
#define CTSIM_PEN_QUAD 1
#define CTSIM_PEN_HYPER 2
#define CTSIM_PEN_ANISO 3
#define CTSIM_PEN_ISOHUBER 4
#define CTSIM_PEN_TV 5
#define CTSIM_PEN_ISOTV 6
int ctsim_penalty(const float* d_x, int32_t nx, int32_t ny, int32_t nz, int32_t pfun, double delta, int32_t mode,
                  float* d_out, double* d_cost, void* stream);
int ctsim_penalty_f64(const double* d_x, int32_t nx, int32_t ny, int32_t nz, int32_t pfun, double delta, int32_t mode,
                      double* d_out, double* d_cost, void* stream);
/* ctsim_penalty with pfun = CTSIM_PEN_HUBER */
int ctsim_huber(const float* d_x, int32_t nx, int32_t ny, int32_t nz, double delta, int32_t mode, float* d_out,
                double* d_cost, void* stream);

/* One Nesterov-SQS subset update (reconPwlsSeNesterovSqs.m:67-87), fused:
 *   grad = g + beta*S(z);  du = grad ./ (precom + beta*T(x));  x' = z - du;
 *   s = s - t*du;  z' = (1 - 1/t_next)*x' + (1/t_next)*s;  clamp x', s, z' to >= 0
 *   and NaN/Inf -> 0.  S(z) and T(x) are stencils evaluated in-kernel, so x' and z'
 *   go to separate buffers (d_x_out, d_z_out must not alias d_x or d_z); s is
 *   updated in place.  t and t_next are the momentum scalars of the current and next
 *   step.  The denominator and the step are formed in FP64.  d_zcopy (may be NULL)
 *   receives z' again in zcopy_dtype, minus d_zref when d_zref is not NULL: the input of
 *   the next forward projection, e.g. z' - x0 when the stored data are y - A x0 (A is
 *   linear, so A(z') - y = A(z' - x0) - (y - A x0)), or an FP16 copy of z'. */
int ctsim_pwls_nesterov_update(const float* d_x, const float* d_z, float* d_s, const float* d_g,
                               const float* d_precom, float* d_x_out, float* d_z_out, void* d_zcopy,
                               int32_t zcopy_dtype, const float* d_zref, int32_t nx, int32_t ny, int32_t nz,
                               double beta, double delta, int32_t pfun, double t, double t_next, void* stream);

/* Same update with the momentum scalars read from device memory:
 * t = d_tseq[*d_base + k], t_next = d_tseq[*d_base + k + 1].  Lets one captured graph
 * per iteration be replayed while the sequence advances (ctsim_add_i32 moves the
 * base at the end of the graph). */
int ctsim_pwls_nesterov_update_seq(const float* d_x, const float* d_z, float* d_s, const float* d_g,
                                   const float* d_precom, float* d_x_out, float* d_z_out, void* d_zcopy,
                                   int32_t zcopy_dtype, const float* d_zref, int32_t nx, int32_t ny, int32_t nz,
                                   double beta, double delta, int32_t pfun, const double* d_tseq,
                                   const int32_t* d_base, int32_t k, void* stream);

/* *d_counter += increment, as a kernel (graph-capturable). */
int ctsim_add_i32(int32_t* d_counter, int32_t increment, void* stream);

/* One plain SQS subset update (reconPwlsSeSqs.m:57-66): x = max(x - (g + beta S(x)) ./
 * (precom + beta T(x)), 0) with NaN/Inf -> 0; the denominator in FP64.  Writes
 * d_xnew (must not alias d_x). */
int ctsim_pwls_sqs_update(const float* d_x, float* d_xnew, const float* d_g, const float* d_precom, int32_t nx,
                          int32_t ny, int32_t nz, double beta, double delta, int32_t pfun, void* stream);

/* LALM-OS (statsrecon/pwls/reconPwlsLALMOs.m:66-90).  After the subset gradient
 * dL = Atos(Os(w) .* (Aos(x) - Os(y))):  g = rho/(rho+1) dL + g/(rho+1),
 * s = rho dL + (1 - rho) g with NaN/Inf -> 0 (g and s updated in place). */
int ctsim_pwls_lalm_sg(const float* d_dl, float* d_g, float* d_s, size_t n, double rho, void* stream);
/* One of the inner steps: x_out = max(x - (beta S(x) + rho precom (x - xold) + s) /
 * (beta T(x) + rho precom), 0), NaN/Inf -> 0, denominator in FP64.  x_out must not alias x. */
int ctsim_pwls_lalm_inner(const float* d_x, const float* d_xold, const float* d_s, const float* d_precom,
                          float* d_x_out, int32_t nx, int32_t ny, int32_t nz, double beta, double delta, int32_t pfun,
                          double rho, void* stream);

/* NU-SQS (statsrecon/pwls/reconPwlsSeNusqs.m:47-90).
 * ctsim_nusqs_seed: v = max(2 (|G1*x| + |G2*x|) + |x|, eps), G1 = [1 2 1; 0 0 0; -1 -2 1],
 *   G2 = [1 0 -1; 2 0 -2; 1 0 -1] correlated per z slice with replicated edges
 *   (filter/imageFilter2D.m; :47-51).  d_v must not alias d_x.
 * ctsim_div: out = a ./ b (dL = At(w .* A(v)) ./ v, :54 and :75).
 * ctsim_absdiff_floor: out = max(|x - xref|, eps) (:72-73).
 * ctsim_dynamic_range: statsrecon/pwls/dynamicRangeAdjustment.m, out = max(F(x)^t, eta)
 *   with F the piecewise-linear empirical CDF through min, (min+mean)/2, mean,
 *   (mean+max)/2, max; x must be non-negative (the NU-SQS input is floored at eps).
 *   Uses a small per-device statistics block owned by the library.
 * ctsim_pwls_nusqs_update: x_out = max(x - (g + beta S(x)) ./ (dl + beta h), 0), NaN/Inf
 *   -> 0, the denominator in FP64 (:81-90), with h = H(v) (ctsim_penalty mode 3); d_x_out
 *   must not alias d_x. */
int ctsim_nusqs_seed(const float* d_x, float* d_v, int32_t nx, int32_t ny, int32_t nz, double eps, void* stream);
int ctsim_div(const float* d_a, const float* d_b, float* d_out, size_t n, void* stream);
int ctsim_absdiff_floor(const float* d_x, const float* d_xref, float* d_out, size_t n, double eps, void* stream);
int ctsim_dynamic_range(const float* d_x, size_t n, double t, double eta, float* d_out, void* stream);
int ctsim_pwls_nusqs_update(const float* d_x, float* d_x_out, const float* d_g, const float* d_dl, const float* d_h,
                            int32_t nx, int32_t ny, int32_t nz, double beta, double delta, int32_t pfun, void* stream);

/* sum((a - b).^2) in FP64 into d_out (RMSD numerator, reconPwlsSeNesterovSqs.m:103). */
int ctsim_sqdiff_sum(const float* d_a, const float* d_b, size_t n, double* d_out, void* stream);

/* FP64 forms of the Huber map / cost, the Nesterov-SQS update and the squared-difference
 * sum, for a fully double-precision loop (accuracy reference); same semantics as the
 * float32 entry points above, without the z copy. */
int ctsim_huber_f64(const double* d_x, int32_t nx, int32_t ny, int32_t nz, double delta, int32_t mode, double* d_out,
                    double* d_cost, void* stream);
int ctsim_pwls_nesterov_update_f64(const double* d_x, const double* d_z, double* d_s, const double* d_g,
                                   const double* d_precom, double* d_x_out, double* d_z_out, int32_t nx, int32_t ny,
                                   int32_t nz, double beta, double delta, int32_t pfun, double t, double t_next,
                                   void* stream);
int ctsim_sqdiff_sum_f64(const double* d_a, const double* d_b, size_t n, double* d_out, void* stream);

/* out = a .* b elementwise with dtypes (precom = At(w .* A(ones)), reconPwlsSeNesterovSqs.m:42-43) */
int ctsim_mul(const void* d_a, int32_t a_dtype, const void* d_b, int32_t b_dtype, void* d_out, int32_t out_dtype,
              size_t n, void* stream);
/* out = a - b for FP32 a and b, written as out_dtype (F32, F16 or BF16); the mixed PWLS
 * form stores y - A x0 this way. */
int ctsim_sub(const float* d_a, const float* d_b, void* d_out, int32_t out_dtype, size_t n, void* stream);

/* ------------------------------------------------------------------------------
 * FDK pre-processing on the device (fbp/reconFDK.m:33-76): every detector row of every
 * view is multiplied by the cosine weight (flat panel: SAD sqrt(1 + u^2/SDD^2) /
 * sqrt(u^2 + v^2 + SDD^2), reconFDK.m:39; arc: SAD/SDD cos(u / sqrt(SDD^2 + v^2)),
 * reconFDK.m:44) and by view_weights[view][u] when given (the short-scan weight of
 * shortScanParkerWeight.m), zero padded to `order` cells and filtered as
 * fbp/filter/filterFreq.m:42-51 does: real(ifft(fft(row) .* H)), truncated to nu.
 * H is the full length-`order` frequency response (designFilter2 output times du);
 * only H[0 .. order/2] is used (the response is real and even).  compute CTSIM_F32 or
 * CTSIM_F64 (FFT and weights in that precision).
 * ---------------------------------------------------------------------------- */
typedef struct ctsim_fdk ctsim_fdk;
int ctsim_fdk_create(ctsim_fdk** fdk, const ctsim_geom* g, int32_t nviews, const double* H, int32_t order,
                     const float* view_weights, int32_t compute);
/* The same pre-processing for a fan-beam sinogram [nu nviews] (fbp/reconFBP2d.m:31-55):
 * cosine weight SAD / sqrt(u^2 + SDD^2) (flat, :32) or SAD/SDD cos(u/SDD) (arc, :38), the
 * optional short-scan weight view_weights[view][u] (shortScanSliverWeight.m), filterFreq.
 * Filter with ctsim_fdk_filter (the sinogram is [view][u] with one detector row);
 * g->nv, dv, ov, nz... are ignored. */
int ctsim_fbp2d_create(ctsim_fdk** fdk, const ctsim_geom* g, int32_t nviews, const double* H, int32_t order,
                       const float* view_weights, int32_t compute);
int ctsim_fdk_destroy(ctsim_fdk* fdk);
/* views view0 .. view0+nsub-1: d_sino and d_out are device [nsub][nu][nv] (v fastest,
 * MATLAB's [nv nu nviews] order); d_out in out_dtype (F32, F16, BF16; F64 for F64 plans).
 * The sinogram dtype is F32/F16/BF16 for F32 plans and F64 for F64 plans. */
int ctsim_fdk_filter(const ctsim_fdk* fdk, int32_t view0, int32_t nsub, const void* d_sino, int32_t sino_dtype,
                     void* d_out, int32_t out_dtype, void* stream);

/* ------------------------------------------------------------------------------
 * Data side of the reconstruction loop on the device.
 * ---------------------------------------------------------------------------- */
/* d_counts[i] ~ Poisson(i0 * exp(-p[i])), p in F32 or F64.  Philox-4x32-10 with key =
 * seed and counter = (i, attempt): the draws depend on (seed, i) only.  The reference
 * draws counts with MATLAB's poissrnd before computeSimplePoissonSinogramAttunation.m;
 * the two generators differ, the distribution is the same. */
int ctsim_poisson_counts(const void* d_p, int32_t p_dtype, size_t n, double i0, uint64_t seed, float* d_counts,
                         void* stream);
/* y = -log(max(counts, low_count) / i0) in y_dtype (computeSimplePoissonSinogramAttunation.m:14-16,
 * monochromatic source, low_count = dummyLowPhotonCount = 10 there). */
int ctsim_log_attenuation(const float* d_counts, size_t n, double i0, double low_count, void* d_y, int32_t y_dtype,
                          void* stream);
/* The raw-data log of sinogram/log/processCTRawData.m:14-33 in full.  Element i belongs to
 * view i / ncell and detector cell i % ncell (ncell = nu * nv); counts are first divided by
 * d_tube_profile[view] when given.  compound = 0 (computeSimplePoissonSinogramAttunation.m):
 * y = -log((max(c, low_count) / source_intensity) / d_flat_field[cell]); compound = 1
 * (computeCompoundPoissonSinogramAttunation.m): y = -log(max((c / source_intensity) /
 * d_flat_field[cell], low_count / source_intensity)).  d_flat_field (the bowtie's
 * flatFieldRatio, [nv nu]) and d_tube_profile ([nviews]) may be NULL.  source_intensity is
 * the spectrum's expected unattenuated count (ctsim_amd.sim.source_intensity). */
int ctsim_log_attenuation_ex(const float* d_counts, size_t n, int32_t ncell, double source_intensity,
                             double low_count, int32_t compound, const float* d_flat_field,
                             const float* d_tube_profile, void* d_y, int32_t y_dtype, void* stream);
/* Water beam-hardening correction of a sinogram of line integrals
 * (artifact/beamhardening/beamHardeningWarterCorrection.m:40-69, the last step of
 * sinogram/log/processCTRawData.m:40-43): out = sum_{k=1..order} c_k(u) in^k in single
 * precision, per detector channel u (the column of a [nv nu nviews] sinogram; nv = 1 for
 * a 2D [nu nviews] one) when per_channel is nonzero, else one coefficient set.
 * d_coeffs: device floats, [channel][order] with the order fastest (order values when
 * per_channel is 0); the coefficients come from the host fit of
 * beamHardeningCorrectionPolynomialCoefficients.m (ctsim_amd.beamhardening).  F32, F16 or
 * BF16 in and out; in place allowed. */
int ctsim_water_correction(const void* d_in, int32_t in_dtype, void* d_out, int32_t out_dtype, int32_t nv, int32_t nu,
                           int32_t nviews, const float* d_coeffs, int32_t order, int32_t per_channel, void* stream);
/* w = (s - r)^2 / (s + v) in single precision, 0 where s <= r (statsrecon/pwls/computeWeightsPwls.m:8-9). */
int ctsim_pwls_weights(const float* d_counts, size_t n, double r, double v, void* d_w, int32_t w_dtype, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CTSIM_AMD_H */
