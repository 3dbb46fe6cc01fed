/*
 * cbct_back_pd_cuda - mexFunction answering to CTSim's cbct_back_pd_cuda name, replacing
 * geom/fdk_cuda/cbct_back_pd_cuda.cu:21-193 (16 arguments)
 * with libctsim_hip.so.  ops: back,pd | back,hpd.
 *   out = cbct_back_pd_cuda(op, nxyz, dxyz, cxyz, nuv, duv, cuv, SAD, ADD, SDD, noviews, betas,
 *               couchz, detector_type, data, map)
 */
#include "ctsim_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    ctsim_mex_cbct(ctsim_cbct_back_pd_cuda, 1, nlhs, plhs, nrhs, prhs);
}
