/* Implementation of the stand-in MATLAB API (tests only).  stub_call runs a
 * gateway and turns mexErrMsgIdAndTxt into a non-zero return with the
 * message; stub_array / stub_data / stub_dims / stub_free let ctypes build
 * and read arrays. */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    size_t ndim, dims[8], n, esz;
    void* data;
};

static jmp_buf g_jmp;
static char g_msg[1024];

static size_t class_size(mxClassID c) {
    switch (c) {
        case mxINT8_CLASS: case mxUINT8_CLASS: return 1;
        case mxSINGLE_CLASS: case mxINT32_CLASS: return 4;
        default: return 0;
    }
}

void* mxGetData(const mxArray* a) { return a->data; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->n; }
int mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
int mxIsInt32(const mxArray* a) { return a->cls == mxINT32_CLASS; }
int mxIsUint8(const mxArray* a) { return a->cls == mxUINT8_CLASS; }
int mxIsInt8(const mxArray* a) { return a->cls == mxINT8_CLASS; }

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c) {
    (void)c;
    mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
    a->cls = cls;
    a->ndim = ndim;
    a->n = 1;
    for (size_t i = 0; i < ndim && i < 8; ++i) {
        a->dims[i] = dims[i];
        a->n *= dims[i];
    }
    a->esz = class_size(cls);
    a->data = calloc(a->n ? a->n : 1, a->esz ? a->esz : 1);  /* MATLAB zero-fills */
    return a;
}

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
    va_list ap;
    int k = snprintf(g_msg, sizeof g_msg, "%s: ", id);
    va_start(ap, fmt);
    vsnprintf(g_msg + k, sizeof g_msg - (size_t)k, fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}

/* --- test hooks (called through ctypes) --- */
mxArray* stub_array(int cls, int ndim, const size_t* dims, const void* src) {
    mxArray* a = mxCreateNumericArray((mwSize)ndim, dims, (mxClassID)cls, mxREAL);
    if (src) memcpy(a->data, src, a->n * a->esz);
    return a;
}
void stub_free(mxArray* a) {
    if (a) {
        free(a->data);
        free(a);
    }
}
void* stub_data(const mxArray* a) { return a->data; }
size_t stub_dims(const mxArray* a, size_t* dims) {
    for (size_t i = 0; i < a->ndim; ++i) dims[i] = a->dims[i];
    return a->ndim;
}
int stub_call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* msg, size_t cap) {
    g_msg[0] = 0;
    if (setjmp(g_jmp)) {
        snprintf(msg, cap, "%s", g_msg);
        return 1;
    }
    mexFunction(nlhs, plhs, nrhs, prhs);
    msg[0] = 0;
    return 0;
}
