/*
 * rsvd_fixed_rank_mex.c -- MATLAB gateway to librsvdpack's fixed-rank
 * randomized SVD.
 *
 *   [U, S, V] = rsvd_fixed_rank_mex(A, k, p, vnum, q, s)
 *
 * The calling sequence and outputs of the reference gateway
 * matlab_code/mex_code_multi_core_mkl/low_rank_svd_rand_decomp_fixed_rank_mkl_mex.c:16-118
 * (A real double m x n; U m x k, S k x k diagonal, V n x k; vnum 2 returns the
 * values ascending, as low_rank_svd_rand_decomp_fixed_rank does).  MATLAB's
 * column-major A is handed to the library without a copy; the library copies
 * it to HBM, runs the decomposition on the GPU and returns host factors, which
 * are copied into new MATLAB arrays.  A library failure (no GPU, bad sizes) is
 * raised as a MATLAB error carrying rsvd_last_error().
 *
 * Build (MATLAB's mex on a host with ROCm):
 *   mex -I$REPO/include rsvd_fixed_rank_mex.c \
 *       $REPO/lowrankmatrixdecompositioncodes_amd/librsvdpack.so
 */
#include <string.h>

#include "mex.h"
#include "rsvdpack.h"
#include "rsvdpack_ext.h"

static int int_arg(const mxArray *a, const char *name) {
    if (!mxIsDouble(a) || mxIsComplex(a) || mxGetM(a) * mxGetN(a) != 1)
        mexErrMsgIdAndTxt("rsvdpack:arg", "%s must be a real double scalar", name);
    double v = mxGetScalar(a);
    if (v != (double)(int)v)
        mexErrMsgIdAndTxt("rsvdpack:arg", "%s must be an integer", name);
    return (int)v;
}

static mxArray *to_matlab(const mat *M) {
    mxArray *out = mxCreateDoubleMatrix((mwSize)M->nrows, (mwSize)M->ncols, mxREAL);
    memcpy(mxGetPr(out), M->d, sizeof(double) * (size_t)M->nrows * (size_t)M->ncols);
    return out;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs != 6)
        mexErrMsgIdAndTxt("rsvdpack:nrhs", "six inputs required: matrix A, target rank k, "
                          "oversampling p, algorithm version vnum, power scheme power q, "
                          "power scheme reorthonormalization amount s");
    if (nlhs > 3) mexErrMsgIdAndTxt("rsvdpack:nlhs", "at most three outputs: U, S, V");
    const mxArray *A = prhs[0];
    if (!mxIsDouble(A) || mxIsComplex(A))
        mexErrMsgIdAndTxt("rsvdpack:arg", "A must be a real double matrix");
    const int m = (int)mxGetM(A), n = (int)mxGetN(A);
    const int k = int_arg(prhs[1], "k"), p = int_arg(prhs[2], "p");
    const int vnum = int_arg(prhs[3], "vnum"), q = int_arg(prhs[4], "q"), s = int_arg(prhs[5], "s");
    const int mn = m < n ? m : n;
    if (k + p > mn)
        mexErrMsgIdAndTxt("rsvdpack:arg", "(k+p) = %d, but must be at most min(m,n) = %d", k + p,
                          mn);
    mat M = {m, n, mxGetPr(A)};  /* read only in the library */
    mat *U = NULL, *S = NULL, *V = NULL;
    int frank = 0;
    low_rank_svd_rand_decomp_fixed_rank(&M, k, p, vnum, q, s, &frank, &U, &S, &V);
    if (!U || !S || !V)
        mexErrMsgIdAndTxt("rsvdpack:failed", "low_rank_svd_rand_decomp_fixed_rank: %s",
                          rsvd_last_error());
    plhs[0] = to_matlab(U);
    if (nlhs > 1) plhs[1] = to_matlab(S);
    if (nlhs > 2) plhs[2] = to_matlab(V);
    matrix_delete(U);
    matrix_delete(S);
    matrix_delete(V);
}
