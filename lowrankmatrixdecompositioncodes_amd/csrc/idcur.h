// idcur.h -- interpolative and CUR decompositions on the GPU (idcur.hip).
#pragma once

#include <vector>

#include "pipeline.h"

namespace rsvd {

// A column ID: I (n 0-based pivots), T = Rk1^-1 Rk2 (frank x (n - frank), ld
// frank) on the device.
struct IdOut {
  std::vector<int> I;
  DBuf<double> T;
  int frank = 0;
  int64_t n = 0;
};

// A CUR decomposition: C = A(:, Icol(1:f)) (m x f), U (f x f), R = A(Irow(1:f),
// :) (f x n), column-major with leading dimensions m, f, f.
struct CurOut {
  DBuf<double> C, U, R;
  int frank = 0;
};

// Column-pivoted Householder QR of A (m x n, in place): R in the upper
// trapezoid of the first min(m, n) rows, reflectors below (v_i = 1 implied),
// tau (min(m, n)), perm (n, 0-based: column j of the factored matrix is
// column perm[j] of the input).  Pivot choice and norm downdating as LAPACK
// dgeqp3 / dlaqp2 with every column free.
int pivoted_qr_f64(Context* c, int64_t m, int64_t n, double* A, int64_t lda, int* perm,
                   double* tau);
// B (k x nrhs) <- R^-1 B for an upper triangular R (k x k): blocked back
// substitution with inverted 64 x 64 diagonal blocks and GEMM updates.
int trsm_upper_f64(Context* c, int64_t k, int64_t nrhs, const double* R, int64_t ldr, double* B,
                   int64_t ldb);

int id_blockrand_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                     double TOL, int kstep, int q, int s, const OmegaSrc& om, IdOut* out);
// col: the column ID of A; row: the full ID of A(:, Icol(1:frank))^T, whose
// pivots are the row indices of A (m of them).
int id_two_sided_blockrand_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda,
                               int k, int p, double TOL, int kstep, int q, int s,
                               const OmegaSrc& om, IdOut* col, IdOut* row);
int cur_blockrand_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                      double TOL, int kstep, int q, int s, const OmegaSrc& om, CurOut* out);

// A(i, j) = 0 below the diagonal (rows x cols).
int zero_lower_f64(Context* c, int64_t rows, int64_t cols, double* A, int64_t lda);
// out(:, j) = A(:, idx[j]) and out(r, :) = A(idx[r], :) (device index arrays).
int gather_cols_f64(Context* c, int64_t m, int64_t cnt, const double* A, int64_t lda,
                    const int* idx, double* out, int64_t ldo);
int gather_rows_f64(Context* c, int64_t cnt, int64_t n, const double* A, int64_t lda,
                    const int* idx, double* out, int64_t ldo);
// V = [I_k; T^T](inv, :) (n x k), inv the inverse permutation of an ID's pivots.
int id_interp_matrix_f64(Context* c, int64_t n, int64_t k, const int* inv, const double* T,
                         int64_t ldt, double* V, int64_t ldv);

}  // namespace rsvd
