/* Minimal stand-in for MATLAB's mex.h/matrix.h: only what the ddfan MEX
 * gateways use, so they compile and run in tests without MATLAB.  Arrays own
 * their data; mexErrMsgIdAndTxt returns control to stub_call. */
#ifndef DDFAN_TEST_MEX_H
#define DDFAN_TEST_MEX_H
#include <stddef.h>

typedef size_t mwSize;
typedef enum { mxUNKNOWN_CLASS = 0, mxINT8_CLASS = 8, mxUINT8_CLASS = 9, mxSINGLE_CLASS = 7,
               mxINT32_CLASS = 12 } mxClassID;
typedef enum { mxREAL = 0 } mxComplexity;
typedef struct mxArray_tag mxArray;

void* mxGetData(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
int mxIsSingle(const mxArray* a);
int mxIsInt32(const mxArray* a);
int mxIsUint8(const mxArray* a);
int mxIsInt8(const mxArray* a);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
#endif
