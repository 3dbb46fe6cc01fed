/*
 * fbct_geom_mex - mexFunction answering to the 2D fan-beam gateway name that
 * geom/shell/forwardProjectMex.m:52-55 and geom/shell/backProjectMex.m:145-148 call,
 * replacing geom/fbct_os/fbct_geom_mex.c:250-326 with libctsim_hip.so.
 *   sino = fbct_geom_mex(op, nxy, dxy, oxy, nu, du, ou, SAD, ADD, SDD, noviews, betas,
 *                        detector_type, img)      op "proj,dd" | "proj,tf"
 *   img  = fbct_geom_mex(op, ..., sino)           op "back,dd" | "back,tf" | "back,pd"
 */
#include "ctsim_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    char op[32], err[CTSIM_MEX_ERRLEN];
    int32_t nxy[2], nu, noviews, dettype;
    float dxy[2], oxy[2], du, ou, sad, add, sdd;
    if (nrhs < 1) mexErrMsgTxt("Error: incorrect number of arguments.");
    ctsim_mx_op(prhs[0], op, sizeof op);
    if (nrhs != 14 || nlhs != 1) mexErrMsgTxt("Error: incorrect number of arguments.");
    ctsim_mx_i32(prhs[1], nxy, 2, "nxy");
    ctsim_mx_f32(prhs[2], dxy, 2, "dxy");
    ctsim_mx_f32(prhs[3], oxy, 2, "oxy");
    ctsim_mx_i32(prhs[4], &nu, 1, "nu");
    ctsim_mx_f32(prhs[5], &du, 1, "du");
    ctsim_mx_f32(prhs[6], &ou, 1, "ou");
    ctsim_mx_f32(prhs[7], &sad, 1, "SAD");
    ctsim_mx_f32(prhs[8], &add, 1, "ADD");
    ctsim_mx_f32(prhs[9], &sdd, 1, "SDD");
    ctsim_mx_i32(prhs[10], &noviews, 1, "noviews");
    const double* betas = ctsim_mx_betas(prhs[11], noviews);
    ctsim_mx_i32(prhs[12], &dettype, 1, "detector_type");
    const int proj = ctsim_mx_is_proj(op);
    const int64_t nimg = (int64_t)nxy[0] * nxy[1];
    const int64_t nsino = (int64_t)nu * noviews;
    const float* in = ctsim_mx_data(prhs[13], proj ? nimg : nsino,
                                    proj ? "Error: mx_image must have [nx ny] float array."
                                         : "Error: mx_proj must have [nu noviews] float array.");
    float* out = proj ? ctsim_mx_out(plhs, nu, noviews, 1) : ctsim_mx_out(plhs, nxy[0], nxy[1], 1);
    if (ctsim_fbct_geom_mex(op, nxy, dxy, oxy, nu, du, ou, sad, add, sdd, noviews, betas, dettype, in, out, err,
                            sizeof err))
        mexErrMsgTxt(err);
}
