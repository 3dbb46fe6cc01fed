/*
 * mex.h - a minimal stand-in for MATLAB's MEX header, declaring only the mx/mex calls
 * that the shims in ctsim_amd/csrc/mex/ use, with MATLAB's documented signatures
 * (large-array-dims API: mwSize is size_t).  It lets the tests compile the shims with
 * gcc -Wall -Werror and drive mexFunction with hand-made arrays (mexstub.c), which
 * catches type and signature drift between the shims and include/ctsim_amd.h.
 * It is test infrastructure, not a MATLAB build.
 */
#ifndef CTSIM_TEST_MEX_H
#define CTSIM_TEST_MEX_H

#include <stdbool.h>
#include <stddef.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;

typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxCELL_CLASS,
    mxSTRUCT_CLASS,
    mxLOGICAL_CLASS,
    mxCHAR_CLASS,
    mxVOID_CLASS,
    mxDOUBLE_CLASS,
    mxSINGLE_CLASS,
    mxINT8_CLASS,
    mxUINT8_CLASS,
    mxINT16_CLASS,
    mxUINT16_CLASS,
    mxINT32_CLASS,
    mxUINT32_CLASS,
    mxINT64_CLASS,
    mxUINT64_CLASS,
    mxFUNCTION_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

void* mxGetData(const mxArray* pm);
mxClassID mxGetClassID(const mxArray* pm);
size_t mxGetNumberOfElements(const mxArray* pm);
bool mxIsNumeric(const mxArray* pm);
bool mxIsComplex(const mxArray* pm);
bool mxIsDouble(const mxArray* pm);
bool mxIsSingle(const mxArray* pm);
bool mxIsLogical(const mxArray* pm);
bool mxIsChar(const mxArray* pm);
int mxGetString(const mxArray* pm, char* str, mwSize strlen);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity flag);
void mexErrMsgTxt(const char* errormsg);

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#endif
