/* mex_mock.c -- the stand-in MEX runtime of mex.h (test infrastructure). */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    size_t m, n;
    double *d;
};

static jmp_buf g_jmp;
static char g_err[1024];

size_t mxGetM(const mxArray *a) { return a->m; }
size_t mxGetN(const mxArray *a) { return a->n; }
double *mxGetPr(const mxArray *a) { return a->d; }
double mxGetScalar(const mxArray *a) { return (a->m > 0 && a->n > 0) ? a->d[0] : 0.0; }
bool mxIsDouble(const mxArray *a) { (void)a; return true; }
bool mxIsComplex(const mxArray *a) { (void)a; return false; }

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
    (void)c;
    mxArray *a = (mxArray *)malloc(sizeof(mxArray));
    a->m = m;
    a->n = n;
    a->d = (double *)calloc(m * n + 1, sizeof(double));
    return a;
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    int k = snprintf(g_err, sizeof g_err, "%s: ", id);
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err + k, sizeof g_err - (size_t)k, fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}

mxArray *mock_matrix(size_t m, size_t n, const double *data) {
    mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
    if (data) memcpy(a->d, data, sizeof(double) * m * n);
    return a;
}

mxArray *mock_scalar(double v) { return mock_matrix(1, 1, &v); }

void mock_free(mxArray *a) {
    if (!a) return;
    free(a->d);
    free(a);
}

int mock_call(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[], const char **err) {
    g_err[0] = 0;
    if (setjmp(g_jmp)) {
        *err = g_err;
        return 1;
    }
    mexFunction(nlhs, plhs, nrhs, prhs);
    *err = NULL;
    return 0;
}
