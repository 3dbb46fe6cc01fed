/*
 * cbct_sf_cuda - mexFunction answering to CTSim's cbct_sf_cuda name, replacing
 * geom/sf_cuda/cbct_sf_cuda.cu with libctsim_hip.so.  16 arguments.
 * ops: proj,tf | back,tf.
 *   out = cbct_sf_cuda(op, nxyz, dxyz, cxyz, nuv, duv, cuv, SAD, ADD, SDD, noviews, betas,
 *                      couchz, detector_type, data, map)
 */
#include "ctsim_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    ctsim_mex_cbct(ctsim_cbct_sf_cuda, 1, nlhs, plhs, nrhs, prhs);
}
