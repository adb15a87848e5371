/*
 * mex.h -- a minimal stand-in for MATLAB's MEX API, owned by this project, so that the MEX
 * gateways in this directory (the files a MATLAB driver would build with `mex`) compile and run
 * without MATLAB.  It declares the subset of the API the gateways use, with MATLAB's names and
 * meanings: double (real, column-major), char, 1x1 struct, cell and uint64 arrays; mexErrMsgTxt
 * does not return (here it jumps back to mexshim_call, as MATLAB unwinds to its prompt).
 * Against MATLAB the gateways are built with MATLAB's own mex.h instead; nothing here is MATLAB's.
 */
#ifndef BBPD_MEX_SHIM_H
#define BBPD_MEX_SHIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;

typedef enum { mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxCHAR_CLASS, mxDOUBLE_CLASS = 6,
               mxUINT64_CLASS = 13 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

double *mxGetPr(const mxArray *a);
void *mxGetData(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double mxGetScalar(const mxArray *a);
int mxIsChar(const mxArray *a);
int mxIsStruct(const mxArray *a);
int mxIsCell(const mxArray *a);
int mxIsDouble(const mxArray *a);
mxClassID mxGetClassID(const mxArray *a);
mxArray *mxGetField(const mxArray *a, mwIndex i, const char *name);
mxArray *mxGetCell(const mxArray *a, mwIndex i);
char *mxArrayToString(const mxArray *a); /* free with mxFree */
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID c, mxComplexity cx);
void mxDestroyArray(mxArray *a);
void mxFree(void *p);
void mexErrMsgTxt(const char *msg)
#if defined(__GNUC__)
    __attribute__((noreturn))
#endif
    ;
int mexPrintf(const char *fmt, ...);

/* --- the harness side (not MATLAB API): wrap host buffers, call a gateway ------------------ */
mxArray *mexshim_double(size_t m, size_t n, double *pr); /* wraps pr (not copied, not freed) */
mxArray *mexshim_string(const char *s);
mxArray *mexshim_struct(int nfields, const char **names, mxArray **values); /* takes the values */
mxArray *mexshim_cell(int n, mxArray **values);                             /* takes the values */
typedef void (*mexshim_gateway)(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
/* runs fn; 0 on return, 1 after mexErrMsgTxt (its text in err) */
int mexshim_call(mexshim_gateway fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *err,
                 int errlen);

#ifdef __cplusplus
}
#endif
#endif
