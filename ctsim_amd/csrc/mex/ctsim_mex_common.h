/*
 * ctsim_mex_common.h - marshalling shared by the mexFunction shims that put
 * libctsim_hip.so behind CTSim's MEX names (fbct_geom_mex, cbct_geom_mex,
 * cbct_dd_cuda, cbct_sf_cuda, cbct_back_pd_cuda).
 *
 * Built only where MATLAB's mex.h exists, e.g.
 *     mex -I../../../include cbct_geom_mex.c -L../.. -lctsim_hip
 * The shim checks argument counts and classes, converts numeric scalars and size
 * vectors of any real class (the reference reads them through unchecked casts,
 * geom/dd_cuda/ct_def.h:19-22, so MATLAB callers pass int32, single or double),
 * allocates the single-precision output MATLAB owns, and reports every library
 * failure through mexErrMsgTxt with the library's message.
 */
#ifndef CTSIM_MEX_COMMON_H
#define CTSIM_MEX_COMMON_H

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "ctsim_amd.h"
#include "mex.h"

#define CTSIM_MEX_ERRLEN 512

/* element k of a real numeric array as double, any class */
static double ctsim_mx_get(const mxArray* a, mwSize k) {
    const void* p = mxGetData(a);
    switch (mxGetClassID(a)) {
        case mxDOUBLE_CLASS: return ((const double*)p)[k];
        case mxSINGLE_CLASS: return ((const float*)p)[k];
        case mxINT8_CLASS: return ((const int8_t*)p)[k];
        case mxUINT8_CLASS: return ((const uint8_t*)p)[k];
        case mxINT16_CLASS: return ((const int16_t*)p)[k];
        case mxUINT16_CLASS: return ((const uint16_t*)p)[k];
        case mxINT32_CLASS: return ((const int32_t*)p)[k];
        case mxUINT32_CLASS: return ((const uint32_t*)p)[k];
        case mxINT64_CLASS: return (double)((const int64_t*)p)[k];
        case mxUINT64_CLASS: return (double)((const uint64_t*)p)[k];
        default: mexErrMsgTxt("Error: numeric argument expected."); return 0.0;
    }
}

static void ctsim_mx_need(const mxArray* a, mwSize n, const char* what) {
    if (!mxIsNumeric(a) || mxIsComplex(a) || mxGetNumberOfElements(a) < n) {
        char msg[160];
        snprintf(msg, sizeof msg, "Error: %s must be a real numeric array of %d element(s).", what, (int)n);
        mexErrMsgTxt(msg);
    }
}

static void ctsim_mx_i32(const mxArray* a, int32_t* out, mwSize n, const char* what) {
    ctsim_mx_need(a, n, what);
    for (mwSize k = 0; k < n; ++k) out[k] = (int32_t)ctsim_mx_get(a, k);
}

static void ctsim_mx_f32(const mxArray* a, float* out, mwSize n, const char* what) {
    ctsim_mx_need(a, n, what);
    for (mwSize k = 0; k < n; ++k) out[k] = (float)ctsim_mx_get(a, k);
}

/* betas: double, noviews x 1 (geom/cbct_os/cbct_geom_mex.c checks class and length) */
static const double* ctsim_mx_betas(const mxArray* a, int32_t noviews) {
    if (!mxIsDouble(a) || mxIsComplex(a) || (int64_t)mxGetNumberOfElements(a) != noviews)
        mexErrMsgTxt("Error: mx_betas must have noviews X 1 double array.");
    return (const double*)mxGetData(a);
}

/* couchz: single, noviews elements (the reference's message says double:
 * geom/cbct_os/cbct_geom_mex.c:143-144) */
static const float* ctsim_mx_couchz(const mxArray* a, int32_t noviews) {
    if (!mxIsSingle(a) || (int64_t)mxGetNumberOfElements(a) != noviews)
        mexErrMsgTxt("Error: mx_couchz must have noviews X 1 double array.");
    return (const float*)mxGetData(a);
}

/* map: logical (or uint8) [nx ny] */
static const uint8_t* ctsim_mx_map(const mxArray* a, int32_t nx, int32_t ny) {
    if (!(mxIsLogical(a) || mxGetClassID(a) == mxUINT8_CLASS) || (int64_t)mxGetNumberOfElements(a) != (int64_t)nx * ny)
        mexErrMsgTxt("Error: mx_map must have [nx ny] bool array.");
    return (const uint8_t*)mxGetData(a);
}

/* data: single with exactly n elements (trailing singleton dimensions are dropped by
 * MATLAB); msg is the reference gateway's message for it */
static const float* ctsim_mx_data(const mxArray* a, int64_t n, const char* msg) {
    if (!mxIsSingle(a) || mxIsComplex(a) || (int64_t)mxGetNumberOfElements(a) != n) mexErrMsgTxt(msg);
    return (const float*)mxGetData(a);
}

static float* ctsim_mx_out(mxArray** plhs, mwSize d0, mwSize d1, mwSize d2) {
    mwSize dims[3] = {d0, d1, d2};
    plhs[0] = mxCreateNumericArray(d2 > 1 ? 3 : 2, dims, mxSINGLE_CLASS, mxREAL);
    if (!plhs[0]) mexErrMsgTxt("Error: cannot allocate output.");
    return (float*)mxGetData(plhs[0]);
}

static void ctsim_mx_op(const mxArray* a, char* op, size_t len) {
    if (!mxIsChar(a) || mxGetString(a, op, (mwSize)len)) mexErrMsgTxt("The 1st argument is invalid!");
}

static int ctsim_mx_is_proj(const char* op) { return strncmp(op, "proj", 4) == 0; }

/* 3D gateways (15 or 16 arguments):
 * op, nxyz, dxyz, cxyz, nuv, duv, cuv, SAD, ADD, SDD, noviews, betas, couchz,
 * detector_type, data [, map] */
typedef int (*ctsim_cbct_fn)(const char*, const int32_t*, const float*, const float*, const int32_t*, const float*,
                             const float*, float, float, float, int32_t, const double*, const float*, int32_t,
                             const float*, const uint8_t*, float*, char*, size_t);

static void ctsim_mex_cbct(ctsim_cbct_fn fn, int require_map, int nlhs, mxArray* plhs[], int nrhs,
                           const mxArray* prhs[]) {
    char op[32], err[CTSIM_MEX_ERRLEN];
    int32_t nxyz[3], nuv[2], noviews, dettype;
    float dxyz[3], cxyz[3], duv[2], cuv[2], sad, add, sdd;
    if (nrhs < 1) mexErrMsgTxt("Error: incorrect number of arguments.");
    ctsim_mx_op(prhs[0], op, sizeof op);
    if (!((nrhs == 16) || (!require_map && nrhs == 15)) || nlhs != 1)
        mexErrMsgTxt("Error: incorrect number of arguments.");
    ctsim_mx_i32(prhs[1], nxyz, 3, "nxyz");
    ctsim_mx_f32(prhs[2], dxyz, 3, "dxyz");
    ctsim_mx_f32(prhs[3], cxyz, 3, "cxyz");
    ctsim_mx_i32(prhs[4], nuv, 2, "nuv");
    ctsim_mx_f32(prhs[5], duv, 2, "duv");
    ctsim_mx_f32(prhs[6], cuv, 2, "cuv");
    ctsim_mx_f32(prhs[7], &sad, 1, "SAD");
    ctsim_mx_f32(prhs[8], &add, 1, "ADD");
    ctsim_mx_f32(prhs[9], &sdd, 1, "SDD");
    ctsim_mx_i32(prhs[10], &noviews, 1, "noviews");
    const double* betas = ctsim_mx_betas(prhs[11], noviews);
    const float* couchz = ctsim_mx_couchz(prhs[12], noviews);
    ctsim_mx_i32(prhs[13], &dettype, 1, "detector_type");
    const int proj = ctsim_mx_is_proj(op);
    const int64_t nimg = (int64_t)nxyz[0] * nxyz[1] * nxyz[2];
    const int64_t nsino = (int64_t)nuv[0] * nuv[1] * noviews;
    const float* in = ctsim_mx_data(prhs[14], proj ? nimg : nsino,
                                    proj ? "Error: mx_image must have [nx ny nz] float array."
                                         : "Error: mx_proj must have [nv nu noviews] float array.");
    const uint8_t* map = nrhs == 16 ? ctsim_mx_map(prhs[15], nxyz[0], nxyz[1]) : NULL;
    float* out = proj ? ctsim_mx_out(plhs, nuv[1], nuv[0], noviews) : ctsim_mx_out(plhs, nxyz[0], nxyz[1], nxyz[2]);
    if (fn(op, nxyz, dxyz, cxyz, nuv, duv, cuv, sad, add, sdd, noviews, betas, couchz, dettype, in, map, out, err,
           sizeof err))
        mexErrMsgTxt(err);
}

#endif /* CTSIM_MEX_COMMON_H */
