/* Host-side calls of the RSVDPACK API with no GPU visible, for the sanitized
 * build: the L1 matrix/vector helpers (MKL/mvf.c), the binary file round trip,
 * the resizes, and every decomposition's failure path (no device: outputs
 * NULL, the reason in rsvd_last_error).  Exit status 0 when every check
 * holds; a sanitizer finding aborts the process. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rsvdpack.h"
#include "rsvdpack_ext.h"

static int failures = 0;
#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            printf("check failed at line %d: %s\n", __LINE__, #cond); \
            ++failures;                                               \
        }                                                             \
    } while (0)

static mat *filled(int m, int n, double base) {
    mat *M = matrix_new(m, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) matrix_set_element(M, i, j, base + i + 100.0 * j);
    return M;
}

int main(int argc, char **argv) {
    const char *tmp = argc > 1 ? argv[1] : ".";
    char path[4096];
    /* L1 helpers */
    mat *A = filled(7, 5, 0.5), *B = filled(7, 5, 0.25), *C = matrix_new(7, 5);
    matrix_copy(C, A);
    for (int j = 0; j < 5; ++j)
        for (int i = 0; i < 7; ++i) CHECK(matrix_get_element(C, i, j) == matrix_get_element(A, i, j));
    matrix_sub(C, B);
    CHECK(matrix_get_element(C, 3, 2) == 0.25);
    mat *T = matrix_new(5, 7);
    matrix_build_transpose(T, A);
    CHECK(matrix_get_element(T, 4, 6) == matrix_get_element(A, 6, 4));
    int cols[3] = {4, 0, 2}, rows[2] = {6, 1};
    mat *Ac = matrix_new(7, 3), *Ar = matrix_new(2, 5);
    matrix_get_selected_columns(A, cols, Ac);
    CHECK(matrix_get_element(Ac, 5, 0) == matrix_get_element(A, 5, 4));
    matrix_set_selected_columns(C, cols, Ac);
    CHECK(matrix_get_element(C, 2, 2) == matrix_get_element(A, 2, 2));
    matrix_get_selected_rows(A, rows, Ar);
    CHECK(matrix_get_element(Ar, 0, 3) == matrix_get_element(A, 6, 3));
    matrix_set_selected_rows(C, rows, Ar);
    CHECK(matrix_get_element(C, 1, 4) == matrix_get_element(A, 1, 4));
    mat *Sq = filled(5, 5, 1.0), *Sy = matrix_new(5, 5);
    matrix_copy_symmetric(Sy, Sq);
    CHECK(matrix_get_element(Sy, 1, 3) == matrix_get_element(Sq, 1, 3));
    CHECK(matrix_get_element(Sy, 3, 1) == 0.0);
    vec *d = vector_new(5);
    for (int i = 0; i < 5; ++i) vector_set_element(d, i, 2.0 + i);
    mat *D = matrix_new(5, 5), *Di = matrix_new(5, 5);
    initialize_diagonal_matrix(D, d);
    invert_diagonal_matrix(Di, D);
    CHECK(fabs(matrix_get_element(Di, 3, 3) - 0.2) < 1e-15);
    CHECK(vector_get_element(d, 4) == 6.0);
    /* resizes (allocate, copy, free the old buffer) */
    mat *R = filled(6, 8, 0.0);
    resize_matrix_by_columns(&R, 3);
    CHECK(R->ncols == 3 && matrix_get_element(R, 5, 2) == 5.0 + 200.0);
    resize_matrix_by_rows_from_end(&R, 2);
    CHECK(R->nrows == 2 && matrix_get_element(R, 0, 0) == 4.0);
    mat *R2 = filled(6, 8, 0.0);
    resize_matrix_by_columns_from_end(&R2, 2);
    CHECK(R2->ncols == 2 && matrix_get_element(R2, 0, 0) == 600.0);
    resize_matrix_by_rows(&R2, 4);
    CHECK(R2->nrows == 4 && matrix_get_element(R2, 3, 1) == 703.0);
    /* binary file round trip */
    snprintf(path, sizeof path, "%s/rt.bin", tmp);
    matrix_write_to_binary_file(A, path);
    mat *L = matrix_load_from_binary_file(path);
    CHECK(L && L->nrows == 7 && L->ncols == 5);
    if (L) {
        for (int j = 0; j < 5; ++j)
            for (int i = 0; i < 7; ++i) CHECK(matrix_get_element(L, i, j) == matrix_get_element(A, i, j));
        matrix_delete(L);
    }
    snprintf(path, sizeof path, "%s/missing.bin", tmp);
    CHECK(matrix_load_from_binary_file(path) == NULL && strlen(rsvd_last_error()) > 0);
    /* decompositions with no device: NULL outputs, an error message */
    mat *M = filled(40, 30, 1.0), *U = NULL, *S = NULL, *V = NULL, *Q = NULL, *Bq = NULL;
    int frank = -7;
    low_rank_svd_rand_decomp_fixed_rank(M, 5, 2, 2, 2, 1, &frank, &U, &S, &V);
    CHECK(!U && !S && !V && frank == -7 && strlen(rsvd_last_error()) > 0);
    low_rank_svd_blockrand_decomp_fixed_rank_or_prec(M, 5, 2, 0.0, 2, 4, 1, 1, &frank, &U, &S, &V);
    CHECK(!U && !S && !V && strlen(rsvd_last_error()) > 0);
    randQB_pb(M, 4, 2, 1, 1, &Q, &Bq);
    CHECK(!Q && !Bq && strlen(rsvd_last_error()) > 0);
    randQB_pb_new(M, 4, 2, 0.0, 1, 1, &frank, &Q, &Bq);
    CHECK(!Q && !Bq && strlen(rsvd_last_error()) > 0);
    mat *P = matrix_new(40, 30);
    matrix_matrix_mult(M, Sq, P);  /* shape mismatch (30 x 30 needed): error, P untouched */
    CHECK(strlen(rsvd_last_error()) > 0);
    mat *Mats[] = {A, B, C, T, Ac, Ar, Sq, Sy, D, Di, R, R2, M, P};
    for (size_t i = 0; i < sizeof Mats / sizeof Mats[0]; ++i) matrix_delete(Mats[i]);
    vector_delete(d);
    printf(failures ? "FAILED %d\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
