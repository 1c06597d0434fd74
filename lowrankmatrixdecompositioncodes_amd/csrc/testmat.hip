// testmat.hip -- synthetic test matrices generated in HBM.
//
// A = U diag(sigma) V^T (+ noise_std * N(0,1)), with U (m x r) and V (n x r)
// Haar-distributed: orthonormalised seeded Gaussian matrices (the R factor of
// the CholeskyQR / Householder orthonormalisation has a positive diagonal, so
// Q is Haar).  This is the device counterpart of the reference's MATLAB
// generators (make_matrix_binary.m:11-41: qr(randn) factors and a prescribed
// spectrum), sized for matrices that only fit in HBM.
#include <vector>

#include "pipeline.h"

namespace rsvd {

int lowrank_test_matrix(Context* c, int64_t m, int64_t n, int64_t r, const double* sigma_host,
                        double noise_std, uint64_t seed, double* A64, int64_t lda64, float* A32,
                        int64_t lda32) {
  if (r <= 0 || r > m || r > n) return set_error("test matrix: need 0 < r <= min(m,n)");
  DBuf<double> U, V, sig, A;
  RSVD_TRY(U.alloc((size_t)(m * r)));
  RSVD_TRY(V.alloc((size_t)(n * r)));
  RSVD_TRY(sig.alloc((size_t)r));
  RSVD_TRY(gaussian_f64(c, seed, m, r, 0, m, U.p, m));
  RSVD_TRY(gaussian_f64(c, seed + 1, n, r, 0, n, V.p, n));
  RSVD_TRY(orth_f64(c, m, r, U.p, m, nullptr, nullptr));
  RSVD_TRY(orth_f64(c, n, r, V.p, n, nullptr, nullptr));
  RSVD_HIP(hipMemcpyAsync(sig.p, sigma_host, sizeof(double) * r, hipMemcpyHostToDevice, c->stream));
  RSVD_TRY(scale_cols_f64(c, m, r, U.p, m, sig.p, 0));
  double* dst = A64;
  int64_t ldd = lda64;
  if (!dst) {
    RSVD_TRY(A.alloc((size_t)(m * n)));
    dst = A.p;
    ldd = m;
  }
  RSVD_TRY(gemm_f64(c, 'N', 'T', m, n, r, 1.0, U.p, m, V.p, n, 0.0, dst, ldd));
  if (noise_std != 0.0) RSVD_TRY(add_gaussian_f64(c, seed + 2, m, n, noise_std, dst, ldd));
  if (A32) {
    RSVD_TRY(f64_to_f32(c, m, n, dst, ldd, A32, lda32));
    // the double copy holds exactly the float values, so both precisions see
    // the same matrix
    if (A64) RSVD_TRY(f32_to_f64(c, m, n, A32, lda32, A64, lda64));
  }
  RSVD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace rsvd
