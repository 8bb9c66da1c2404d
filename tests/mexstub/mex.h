/* Declarations of the subset of the Octave/MATLAB MEX API that the drop-in
 * shims under slope_stability_octave_amd/mex/ use, for a syntax and type
 * check of those sources with gcc where Octave is not installed.  Test
 * infrastructure only: nothing links against it. */
#ifndef SSR_TEST_MEX_STUB_H
#define SSR_TEST_MEX_STUB_H
#include <stddef.h>
#include <stdbool.h>

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateString(const char *s);
mxArray *mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char **names);
double *mxGetPr(const mxArray *a);
void *mxGetData(const mxArray *a);
mwSize mxGetM(const mxArray *a);
mwSize mxGetN(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double mxGetScalar(const mxArray *a);
bool mxIsInt64(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsComplex(const mxArray *a);
bool mxIsChar(const mxArray *a);
bool mxIsNumeric(const mxArray *a);
bool mxIsStruct(const mxArray *a);
bool mxIsEmpty(const mxArray *a);
bool mxIsLogical(const mxArray *a);
bool mxIsLogicalScalarTrue(const mxArray *a);
bool mxIsSparse(const mxArray *a);
mwIndex *mxGetIr(const mxArray *a);
mwIndex *mxGetJc(const mxArray *a);
mxArray *mxGetField(const mxArray *a, mwIndex i, const char *name);
void mxSetField(mxArray *a, mwIndex i, const char *name, mxArray *v);
char *mxArrayToString(const mxArray *a);
void *mxMalloc(size_t n);
void mxFree(void *p);
void mexErrMsgTxt(const char *msg);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexPrintf(const char *fmt, ...);
#endif
