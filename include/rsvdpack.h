/*
 * rsvdpack.h -- drop-in C ABI of the RSVDPACK randomized SVD / blocked QB path,
 * implemented by librsvdpack.so on AMD Instinct MI355X (gfx950).
 *
 * Every declaration below keeps the reference signature and argument meaning so
 * that an existing caller (driver, MEX gateway) recompiles and links against
 * librsvdpack.so unchanged.  Matrices stay host-side `mat`/`vec` structs
 * (column-major, calloc'ed `double` payload); the library copies A to HBM once
 * per call, runs the whole sketch / power iteration / orthonormalisation /
 * small-SVD pipeline resident on the GPU, and returns freshly matrix_new()'ed
 * outputs that the caller frees with matrix_delete().
 *
 * Errors: the reference API returns void and reports nothing.  Here a failed
 * call leaves every output pointer NULL (or the output untouched for the L1
 * routines), never exits, and records a message readable through
 * rsvd_last_error() declared in rsvdpack_ext.h.
 *
 * Reference citations are relative to the RSVDPACK source tree
 * (multi_core_mkl_code/ = "MKL/", mvf = matrix_vector_functions_intel_mkl,
 * rra = rank_revealing_algorithms_intel_mkl).
 */
#ifndef RSVDPACK_H
#define RSVDPACK_H

#include <stdio.h>
#include <sys/time.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef min
#define min(x,y) (((x) < (y)) ? (x) : (y))
#endif
#ifndef max
#define max(x,y) (((x) > (y)) ? (x) : (y))
#endif

/* MKL/mvf.h:18-21 -- column-major, element (i,j) at d[j*nrows + i]. */
typedef struct {
    int nrows, ncols;
    double * d;
} mat;

/* MKL/mvf.h:24-27 */
typedef struct {
    int nrows;
    double * d;
} vec;

/* ---------------- L1: storage (host) ---------------- */

/* MKL/mvf.h:32, mvf.c:8-16 -- zero-filled (calloc) payload. */
mat * matrix_new(int nrows, int ncols);
/* MKL/mvf.h:35, mvf.c:20-27 */
vec * vector_new(int nrows);
/* MKL/mvf.h:37, mvf.c:30-35 */
void matrix_delete(mat *M);
/* MKL/mvf.h:39, mvf.c:38-43 */
void vector_delete(vec *v);
/* MKL/mvf.h:43-54, mvf.c:47-65 */
void matrix_set_element(mat *M, int row_num, int col_num, double val);
double matrix_get_element(mat *M, int row_num, int col_num);
void vector_set_element(vec *v, int row_num, double val);
double vector_get_element(vec *v, int row_num);

/* MKL/mvf.h:57-60, mvf.c:77-133 -- int32 m, int32 n, then m*n doubles in
 * row-major element order.  Bulk read/written here (one fread per row). */
mat * matrix_load_from_binary_file(char *fname);
void matrix_write_to_binary_file(mat *M, char *fname);

/* MKL/mvf.h:64-68, mvf.c:137-160 */
void matrix_print(mat * M);
void vector_print(vec * v);

/* MKL/mvf.h:96 mvf.c:216-226; :104 :246-253; :112 :274-283 */
void matrix_copy(mat *D, mat *S);
void matrix_build_transpose(mat *Mt, mat *M);
void matrix_sub(mat *A, mat *B);

/* MKL/mvf.h:173-185, mvf.c:647-753 -- column / row gather and scatter. */
void matrix_get_selected_columns(mat *M, int *inds, mat *Mc);
void matrix_set_selected_columns(mat *M, int *inds, mat *Mc);
void matrix_get_selected_rows(mat *M, int *inds, mat *Mr);
void matrix_set_selected_rows(mat *M, int *inds, mat *Mr);

/* MKL/mvf.h:191 mvf.c:757-768; :199 :490-499; :209 :519-528 */
void matrix_copy_symmetric(mat *S, mat *M);
void initialize_diagonal_matrix(mat *D, vec *data);
void invert_diagonal_matrix(mat *Dinv, mat *D);

/* MKL/mvf.h:287-296, mvf.c:1078-1128 -- keep first / last k columns or rows. */
void resize_matrix_by_columns(mat **M, int k);
void resize_matrix_by_columns_from_end(mat **M, int k);
void resize_matrix_by_rows(mat **M, int k);
void resize_matrix_by_rows_from_end(mat **M, int k);

/* MKL/mvf.h:360, mvf.c:1534-1539 */
double get_seconds_frac(struct timeval start_timeval, struct timeval end_timeval);

/* ---------------- L1: compute (GPU) ---------------- */

/* MKL/mvf.h:133, mvf.c:458-486.  Fills M with N(0,1) samples.  The reference
 * uses MKL VSL MCG31 + Gaussian ICDF in float, seeded from time(NULL); here a
 * counter-based Philox4x64-10 stream + inverse normal CDF (AS241), rounded to
 * float and widened to double, generated on the GPU.  The seed is time(NULL)
 * unless fixed with rsvd_set_seed(). */
void initialize_random_matrix(mat *M);

/* MKL/mvf.h:137-145, mvf.c:538-561 -- C = A*B, C = A^T*B, C = A*B^T. */
void matrix_matrix_mult(mat *A, mat *B, mat *C);
void matrix_transpose_matrix_mult(mat *A, mat *B, mat *C);
void matrix_matrix_transpose_mult(mat *A, mat *B, mat *C);

/* MKL/mvf.h:119 mvf.c:360-372; :313 :391-405 */
double get_matrix_frobenius_norm(mat *M);
double get_percent_error_between_two_mats(mat *A, mat *B);

/* MKL/mvf.h:328 mvf.c:1206-1209 -- eigenpairs of the symmetric S (upper
 * triangle used); eigenvalues ascending into evals, eigenvectors into S. */
void compute_evals_and_evecs_of_symm_matrix(mat *S, vec *evals);
/* MKL/mvf.h:333 mvf.c:1214-1245 -- thin [Q,R] = qr(M,0). */
void compact_QR_factorization(mat *M, mat *Q, mat *R);
/* MKL/mvf.h:338 mvf.c:1251-1263 -- thin Q of qr(M,0). */
void QR_factorization_getQ(mat *M, mat *Q);
/* mvf.c:1270-1284 -- M = U*S*Vt, singular values descending on diag(S). */
void singular_value_decomposition(mat *M, mat *U, mat *S, mat *Vt);
/* MKL/mvf.h:348 mvf.c:1304-1318 -- P = U*S*V^T. */
void form_svd_product_matrix(mat *U, mat *S, mat *V, mat *P);

/* ---------------- L2: decompositions (GPU-resident pipelines) ---------------- */

/* MKL/rra.h:5, rra.c:73-234.  Fixed-rank randomized SVD with l = k+p samples,
 * q-1 power iterations re-orthonormalised every s half-steps, and either the
 * QR-of-B^T method (vnum == 1 or vnum > 2: singular values descending, first k
 * kept) or the eigendecomposition-of-BB^T method (vnum == 2: ascending, last k
 * kept).  Outputs U (m x k), S (k x k diagonal), V (n x k).  frank is not
 * written, as in the reference. */
void low_rank_svd_rand_decomp_fixed_rank(mat *M, int k, int p, int vnum, int q, int s,
    int *frank, mat **U, mat **S, mat **V);

/* MKL/rra.h:7-8, rra.c:239-381.  Blocked randomized SVD: randQB_pb_new, then
 * the vnum branch on Q.  Rank mode when k > 0 (frank = k); tolerance mode
 * (absolute Frobenius TOL on the deflated A) only when nstep = (k+p)/kstep
 * evaluates to 0, exactly as the reference's integer arithmetic does. */
void low_rank_svd_blockrand_decomp_fixed_rank_or_prec(mat *M, int k, int p, double TOL,
    int vnum, int kstep, int q, int s, int *frank, mat **U, mat **S, mat **V);

/* MKL/rra.h:57, rra.c:1425-1569.  Blocked QB, rank mode, re-orthogonalisation
 * against the previous blocks at every step > 0.  q is the power-scheme count. */
void randQB_pb(mat *M, int kstep, int nstep, int q, int s, mat **Q, mat **B);

/* MKL/rra.h:60, rra.c:1576-1801.  Blocked QB in rank mode (nstep > 0) or
 * absolute-TOL mode (nstep <= 0); re-orthogonalisation on even steps > 0. */
void randQB_pb_new(mat *M, int kstep, int nstep, double TOL, int q, int s, int *frank,
    mat **Q, mat **B);

/* MKL/rra.h:95, rra.c:2337-2347 and MKL/rra.h:101, rra.c:2388-2398: print
 * norm(M), norm(P) and 100*norm(M-P)/norm(M) for P = U*S*V^T or P = Q*B. */
void use_low_rank_svd_for_approximation(mat *M, mat *U, mat *S, mat *V);
void use_QB_decomp_for_approximation(mat *M, mat *Q, mat *B);

/* ---------------- L2: interpolative and CUR decompositions ---------------- */
/* The index vectors hold 0-based column (row) indices as doubles, as the
 * reference's vec does. */

/* MKL/rra.h:41, rra.c:924-975 -- column-pivoted QR, M(:, I) = Q R with every
 * column free to pivot (LAPACK dgeqp3 rules); k = min(m, n): Q m x k, R k x n
 * for m <= n (else k x k), I the n pivots. */
void pivotedQR_mkl(mat *M, mat **Q, mat **R, vec **I);
/* MKL/rra.h:71, rra.c:1969-2027 -- column ID from the blocked QB:
 * randQB_pb_new with nstep = (k + p) / kstep (TOL mode when k <= 0), the
 * column-pivoted QR of B, rank k or (TOL mode) round(frank / (frank + p +
 * 1e-6) * frank); I: the n pivots, T = Rk1^-1 Rk2 (frank x (n - frank)), so
 * M(:, I) ~ M(:, I(1:frank)) [I T]. */
void id_blockrand_decomp_fixed_rank_or_prec(mat *M, int k, int p, double TOL, int kstep, int q,
    int s, int *frank, vec **I, mat **T);
/* MKL/rra.h:81, rra.c:2086-2112 -- the column ID, then the full ID of
 * M(:, Icol(1:frank))^T: Irow (m row indices), S (frank x (m - frank)). */
void id_two_sided_blockrand_decomp_fixed_rank_or_prec(mat *M, int k, int p, double TOL,
    int kstep, int q, int s, int *frank, vec **Icol, vec **Irow, mat **T, mat **S);
/* MKL/rra.h:91, rra.c:2262-2332 -- C = M(:, Icol(1:frank)), R = M(Irow(1:frank), :),
 * U (frank x frank) solving (R R^T) U^T = R V for the column ID's
 * interpolation matrix V; M ~ C U R.  The reference solves with dgesv; here
 * R R^T (symmetric positive definite) is factored by Cholesky. */
void cur_blockrand_decomp_fixed_rank_or_prec(mat *M, int k, int p, double TOL, int kstep, int q,
    int s, int *frank, mat **C, mat **U, mat **R);
/* MKL/mvf.h:351, mvf.c:1322-1336 -- P = C U R. */
void form_cur_product_matrix(mat *C, mat *U, mat *R, mat *P);
/* MKL/rra.h:104,107,110; rra.c:2402-2573 -- print norm(M), norm of the
 * approximation and the percent error, as use_QB_decomp_for_approximation. */
void use_id_decomp_for_approximation(mat *M, mat *T, vec *I, int k);
void use_id_two_sided_decomp_for_approximation(mat *M, mat *T, mat *S, vec *Icol, vec *Irow,
    int k);
void use_cur_decomp_for_approximation(mat *M, mat *C, mat *U, mat *R);

#ifdef __cplusplus
}
#endif

#endif /* RSVDPACK_H */
