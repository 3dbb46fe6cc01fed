/*
 * cbct_dd_cuda - mexFunction answering to CTSim's cbct_dd_cuda name, replacing
 * geom/dd_cuda/cbct_dd_cuda.cu:21-299 (16 arguments)
 * with libctsim_hip.so.  ops: proj,dd | back,dd.
 *   out = cbct_dd_cuda(op, nxyz, dxyz, cxyz, nuv, duv, cuv, SAD, ADD, SDD, noviews, betas,
 *               couchz, detector_type, data, map)
 */
#include "ctsim_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    ctsim_mex_cbct(ctsim_cbct_dd_cuda, 1, nlhs, plhs, nrhs, prhs);
}
