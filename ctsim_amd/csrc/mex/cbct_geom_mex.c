/*
 * cbct_geom_mex - mexFunction answering to CTSim's cbct_geom_mex name, replacing
 * geom/cbct_os/cbct_geom_mex.c:516-600 with libctsim_hip.so.  15 or 16 arguments
 * (forwardProjectMex.m:82-85 and backProjectMex.m:181-184 send 16, with the map).
 * ops: proj,dd | proj,tf | proj,rd | back,dd | back,tf | back,pd | back,hpd.
 *   out = cbct_geom_mex(op, nxyz, dxyz, cxyz, nuv, duv, cuv, SAD, ADD, SDD, noviews, betas,
 *                       couchz, detector_type, data [, map])
 */
#include "ctsim_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    ctsim_mex_cbct(ctsim_cbct_geom_mex, 0, nlhs, plhs, nrhs, prhs);
}
