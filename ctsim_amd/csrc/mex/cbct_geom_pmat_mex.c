/*
 * cbct_geom_pmat_mex - mexFunction answering to CTSim's projection-matrix gateway name,
 * replacing geom/pcbct/cbct_geom_pmat_mex.c:260-309 with libctsim_hip.so.  13 arguments:
 *   out = cbct_geom_pmat_mex(op, nxyz, dxyz, cxyz, nuv, duv, noviews, pmat, cameras, sdds,
 *                            cuvs, data, map)
 * ops: proj,dd | proj,tf | back,dd | back,tf | back,pd | back,hpd.  The reference reads
 * the detector pitches as duv[1], duv[2] (cbct_geom_pmat_mex.c:49-50): a three-element
 * duv is read that way, a two-element one as duv[0], duv[1].
 */
#include "ctsim_mex_common.h"

static const float* ctsim_mx_single(const mxArray* a, int64_t n, const char* msg) {
    if (!mxIsSingle(a) || mxIsComplex(a) || (int64_t)mxGetNumberOfElements(a) != n) mexErrMsgTxt(msg);
    return (const float*)mxGetData(a);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    char op[32], err[CTSIM_MEX_ERRLEN];
    int32_t nxyz[3], nuv[2], noviews;
    float dxyz[3], cxyz[3], duv[3];
    if (nrhs < 1) mexErrMsgTxt("Error: incorrect number of arguments.");
    ctsim_mx_op(prhs[0], op, sizeof op);
    if (nrhs != 13 || nlhs != 1) mexErrMsgTxt("Error: incorrect number of arguments.");
    /* the op string is checked before the arrays (cbct_geom_pmat_mex.c:271-300) */
    static const char* const ops[] = {"proj,rd", "proj,dd", "proj,tf", "back,pd", "back,hpd", "back,dd", "back,tf"};
    int known = 0;
    for (unsigned k = 0; k < sizeof ops / sizeof ops[0]; ++k) known |= !strcmp(op, ops[k]);
    if (!known) mexErrMsgTxt("Error: unknown geom type.");
    ctsim_mx_i32(prhs[1], nxyz, 3, "nxyz");
    ctsim_mx_f32(prhs[2], dxyz, 3, "dxyz");
    ctsim_mx_f32(prhs[3], cxyz, 3, "cxyz");
    ctsim_mx_i32(prhs[4], nuv, 2, "nuv");
    ctsim_mx_need(prhs[5], 2, "duv");
    const int three = mxGetNumberOfElements(prhs[5]) >= 3;
    ctsim_mx_f32(prhs[5], duv, three ? 3 : 2, "duv");
    ctsim_mx_i32(prhs[6], &noviews, 1, "noviews");
    const int64_t nv = noviews;
    const float* pmat = ctsim_mx_single(prhs[7], 12 * nv, "Error: mx_pmat must have [3 4 noViews ] float array.");
    const float* cams = ctsim_mx_single(prhs[8], 3 * nv, "Error: mx_cameras must have noviews X 2 float array.");
    const float* sdds = ctsim_mx_single(prhs[9], nv, "Error: mx_sdds must have noviews X 1 float array.");
    const float* cuvs = ctsim_mx_single(prhs[10], 2 * nv, "Error: mx_cuvs must have noviews X 2 float array.");
    const int proj = ctsim_mx_is_proj(op);
    const int64_t nimg = (int64_t)nxyz[0] * nxyz[1] * nxyz[2];
    const int64_t nsino = (int64_t)nuv[0] * nuv[1] * noviews;
    const float* in = ctsim_mx_data(prhs[11], proj ? nimg : nsino,
                                    proj ? "Error: mx_image must have [nx ny nz] float array."
                                         : "Error: mx_proj must have [nv nu noviews] float array.");
    if (!mxIsLogical(prhs[12]) || (int64_t)mxGetNumberOfElements(prhs[12]) != (int64_t)nxyz[0] * nxyz[1])
        mexErrMsgTxt("Error: mx_map must have [nx ny] bool array.");
    const uint8_t* map = (const uint8_t*)mxGetData(prhs[12]);
    float* out = proj ? ctsim_mx_out(plhs, nuv[1], nuv[0], noviews) : ctsim_mx_out(plhs, nxyz[0], nxyz[1], nxyz[2]);
    if (ctsim_cbct_geom_pmat_mex(op, nxyz, dxyz, cxyz, nuv, three ? duv[1] : duv[0], three ? duv[2] : duv[1],
                                 noviews, pmat, cams, sdds, cuvs, in, map, out, err, sizeof err))
        mexErrMsgTxt(err);
}
