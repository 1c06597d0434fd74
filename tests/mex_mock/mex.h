/*
 * mex.h -- a minimal stand-in for MATLAB's MEX API, for testing this
 * repository's own gateway (lowrankmatrixdecompositioncodes_amd/mex/) without
 * MATLAB.  Only the calls that gateway makes are provided; an mxArray is a
 * real double column-major matrix.  mexErrMsgIdAndTxt records the message and
 * jumps back to the test driver, as MATLAB unwinds to the prompt.
 */
#ifndef RSVD_TEST_MEX_H
#define RSVD_TEST_MEX_H

#include <stdbool.h>
#include <stddef.h>

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
double *mxGetPr(const mxArray *a);
double mxGetScalar(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsComplex(const mxArray *a);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);

/* test side */
mxArray *mock_matrix(size_t m, size_t n, const double *data);
mxArray *mock_scalar(double v);
void mock_free(mxArray *a);
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
/* Runs mexFunction; returns 0, or 1 after mexErrMsgIdAndTxt (message in *err). */
int mock_call(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[], const char **err);

#endif
