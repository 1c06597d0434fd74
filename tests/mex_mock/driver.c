/*
 * driver.c -- calls the gateway the way MATLAB would:
 *   driver A.bin k p vnum q s seed out_prefix [nrhs]
 * A.bin in the reference's binary layout (int32 rows, cols, row-major
 * doubles); U, S, V written as out_prefix{U,S,V}.bin in the same layout.  A
 * gateway error prints "ERR <message>" and exits 3.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "rsvdpack.h"
#include "rsvdpack_ext.h"

int main(int argc, char **argv) {
    if (argc < 9) {
        fprintf(stderr, "usage: driver A.bin k p vnum q s seed out_prefix [nrhs]\n");
        return 2;
    }
    mat *A = matrix_load_from_binary_file(argv[1]);
    if (!A) {
        printf("ERR load: %s\n", rsvd_last_error());
        return 2;
    }
    rsvd_set_seed(strtoull(argv[7], NULL, 10));
    const mxArray *in[6];
    in[0] = mock_matrix((size_t)A->nrows, (size_t)A->ncols, A->d);
    for (int i = 1; i < 6; ++i) in[i] = mock_scalar(atof(argv[1 + i]));
    const int nrhs = argc > 9 ? atoi(argv[9]) : 6;
    mxArray *out[3] = {NULL, NULL, NULL};
    const char *err = NULL;
    if (mock_call(3, out, nrhs, in, &err)) {
        printf("ERR %s\n", err);
        for (int i = 0; i < 6; ++i) mock_free((mxArray *)in[i]);
        matrix_delete(A);
        return 3;
    }
    const char *names[3] = {"U", "S", "V"};
    for (int i = 0; i < 3; ++i) {
        mat M = {(int)mxGetM(out[i]), (int)mxGetN(out[i]), mxGetPr(out[i])};
        char path[4096];
        snprintf(path, sizeof path, "%s%s.bin", argv[8], names[i]);
        matrix_write_to_binary_file(&M, path);
        mock_free(out[i]);
    }
    for (int i = 0; i < 6; ++i) mock_free((mxArray *)in[i]);
    matrix_delete(A);
    printf("OK\n");
    return 0;
}
