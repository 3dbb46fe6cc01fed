/*
 * mexstub.c - the array model behind tests/mexstub/mex.h.  An mxArray is a class id,
 * a complexity flag, up to four dimensions and a data pointer the test owns (inputs)
 * or the stub allocates (outputs of mxCreateNumericArray, zero-filled as MATLAB's
 * are).  mexErrMsgTxt records the message and longjmps back to stub_call, as MATLAB
 * unwinds a MEX call on error.
 */
#include <setjmp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    int complex_flag;
    mwSize ndim;
    mwSize dims[4];
    void* data;
    int owned;
};

static jmp_buf g_jmp;
static char g_msg[1024];

static size_t elem_size(mxClassID c) {
    switch (c) {
        case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
        case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
        case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
        case mxINT8_CLASS: case mxUINT8_CLASS: case mxLOGICAL_CLASS: return 1;
        default: return 0;
    }
}

/* an array over caller-owned data (class id as in mxClassID; char arrays hold uint16 code units) */
mxArray* stub_wrap(int cls, int complex_flag, size_t n, void* data) {
    mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
    a->cls = (mxClassID)cls;
    a->complex_flag = complex_flag;
    a->ndim = 2;
    a->dims[0] = n;
    a->dims[1] = 1;
    a->data = data;
    return a;
}

void stub_free(mxArray* a) {
    if (!a) return;
    if (a->owned) free(a->data);
    free(a);
}

size_t stub_ndim(const mxArray* a) { return a->ndim; }
size_t stub_dim(const mxArray* a, int k) { return a->dims[k]; }

/* calls a mexFunction; returns 0, or 1 with the mexErrMsgTxt message in msg */
int stub_call(void (*fn)(int, mxArray**, int, const mxArray**), int nlhs, mxArray** plhs, int nrhs,
              const mxArray** prhs, char* msg, size_t len) {
    g_msg[0] = 0;
    if (setjmp(g_jmp) == 0) {
        fn(nlhs, plhs, nrhs, prhs);
        return 0;
    }
    strncpy(msg, g_msg, len - 1);
    msg[len - 1] = 0;
    return 1;
}

void* mxGetData(const mxArray* pm) { return pm->data; }
mxClassID mxGetClassID(const mxArray* pm) { return pm->cls; }

size_t mxGetNumberOfElements(const mxArray* pm) {
    size_t n = 1;
    for (mwSize k = 0; k < pm->ndim; ++k) n *= pm->dims[k];
    return n;
}

bool mxIsNumeric(const mxArray* pm) { return pm->cls >= mxDOUBLE_CLASS && pm->cls <= mxUINT64_CLASS; }
bool mxIsComplex(const mxArray* pm) { return pm->complex_flag != 0; }
bool mxIsDouble(const mxArray* pm) { return pm->cls == mxDOUBLE_CLASS; }
bool mxIsSingle(const mxArray* pm) { return pm->cls == mxSINGLE_CLASS; }
bool mxIsLogical(const mxArray* pm) { return pm->cls == mxLOGICAL_CLASS; }
bool mxIsChar(const mxArray* pm) { return pm->cls == mxCHAR_CLASS; }

int mxGetString(const mxArray* pm, char* str, mwSize strlen) {
    if (pm->cls != mxCHAR_CLASS || strlen == 0) return 1;
    const size_t n = mxGetNumberOfElements(pm);
    const uint16_t* c = (const uint16_t*)pm->data;
    size_t k = 0;
    for (; k < n && k + 1 < strlen; ++k) str[k] = (char)c[k];
    str[k] = 0;
    return k < n ? 1 : 0; /* MATLAB returns 1 when the buffer truncates the string */
}

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity flag) {
    if (ndim > 4) return NULL;
    mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
    a->cls = classid;
    a->complex_flag = flag == mxCOMPLEX;
    a->ndim = ndim;
    size_t n = 1;
    for (mwSize k = 0; k < ndim; ++k) {
        a->dims[k] = dims[k];
        n *= dims[k];
    }
    a->data = calloc(n ? n : 1, elem_size(classid));
    a->owned = 1;
    return a;
}

void mexErrMsgTxt(const char* errormsg) {
    strncpy(g_msg, errormsg, sizeof g_msg - 1);
    g_msg[sizeof g_msg - 1] = 0;
    longjmp(g_jmp, 1);
}
