/* Minimal stand-in for MATLAB's mex.h (R2018a interleaved-complex API), for
 * building and running MEX gateways on machines without MATLAB. TEST
 * INFRASTRUCTURE ONLY: it implements the subset of the mx / mex functions the
 * gateways in this repository call, with MATLAB's documented semantics
 * (column-major storage, mwIndex = size_t sparse indices, mexErrMsgIdAndTxt
 * does not return). */
#ifndef SHIM_MEX_H
#define SHIM_MEX_H

#include <stddef.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef struct { double real, imag; } mxComplexDouble;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

bool mxIsSparse(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsEmpty(const mxArray* a);
bool mxIsCell(const mxArray* a);
bool mxIsStruct(const mxArray* a);
bool mxIsChar(const mxArray* a);
bool mxIsLogical(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
mwIndex* mxGetJc(const mxArray* a);
mwIndex* mxGetIr(const mxArray* a);
double* mxGetDoubles(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
double mxGetScalar(const mxArray* a);
mxArray* mxGetCell(const mxArray* a, mwIndex i);
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name);
int mxGetString(const mxArray* a, char* buf, mwSize len);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
void mxDestroyArray(mxArray* a);
int mexAtExit(void (*fn)(void));
#if defined(__cplusplus)
[[noreturn]]
#endif
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#ifdef __cplusplus
}
#endif

#endif
