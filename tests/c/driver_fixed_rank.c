/* A caller written against the RSVDPACK API only (the shape of
 * multi_core_mkl_code/driver_multi_core_mkl1.c): build M, call
 * low_rank_svd_rand_decomp_fixed_rank, form U*S*V^T, print the percent error.
 * It links librsvdpack.so unchanged; rsvd_set_seed / rsvd_last_error are the
 * only extension calls. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "rsvdpack.h"
#include "rsvdpack_ext.h"

int main(int argc, char **argv) {
    int m = 600, n = 400, k = 40, p = 10, q = 2, s = 1, vnum = 2, frank = -7;
    int i, j;
    mat *M, *U, *S, *V, *P;
    if (argc > 1) vnum = atoi(argv[1]);
    M = matrix_new(m, n);
    /* M(i,j) = sum_r exp(-r/5) * cos((i+1)(r+1)) * sin((j+1)(r+1)), r < 60 */
    for (i = 0; i < m; i++)
        for (j = 0; j < n; j++) {
            double v = 0;
            int r;
            for (r = 0; r < 60; r++) v += exp(-r / 5.0) * cos(0.01 * (i + 1) * (r + 1)) * sin(0.013 * (j + 1) * (r + 1));
            matrix_set_element(M, i, j, v);
        }
    rsvd_set_seed(12345);
    low_rank_svd_rand_decomp_fixed_rank(M, k, p, vnum, q, s, &frank, &U, &S, &V);
    if (!U || !S || !V) {
        printf("FAILED: %s\n", rsvd_last_error());
        return 1;
    }
    P = matrix_new(m, n);
    form_svd_product_matrix(U, S, V, P);
    printf("normM = %f ; normS = %f\n", get_matrix_frobenius_norm(M), get_matrix_frobenius_norm(S));
    printf("percent_error between M and U S V^T = %.10f\n", get_percent_error_between_two_mats(M, P));
    printf("frank untouched = %d\n", frank);
    printf("S(0,0) = %.15e S(k-1,k-1) = %.15e\n", matrix_get_element(S, 0, 0), matrix_get_element(S, k - 1, k - 1));
    matrix_delete(M); matrix_delete(U); matrix_delete(S); matrix_delete(V); matrix_delete(P);
    return 0;
}
