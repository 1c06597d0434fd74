// pipeline_mp.hip -- mixed-precision fixed-rank rSVD on a float32 A in HBM.
//
// Same control flow as low_rank_svd_rand_decomp_fixed_rank (MKL/rra.c:73-234)
// with the precision of each step chosen by its sensitivity:
//   * sketch and power-iteration passes over A (Y = A Omega, Z = A^T Y,
//     Y = A Z): bf16 matrix cores, fp32 accumulation (prec = MP_BF16),
//     split-bf16 (prec = MP_BF16X3, fp32-level accuracy), or fp8 e4m3 with
//     per-128-column scales of A (prec = MP_FP8);
//   * orthonormalisation: CholeskyQR with the Gram in fp32 (split-bf16
//     products) and the Cholesky / triangular inverse in FP64; one pass inside
//     the power iteration (only the range matters there), two passes (CQR2)
//     for the final Q;
//   * projection B^T = A^T Q: split-bf16 (fp32-level accuracy), because every
//     singular value is read off B;
//   * BB^T, its eigendecomposition (two-sided block Jacobi) and sqrt: FP64;
//   * U = Q Uhat and V = B^T Uhat Sigma^-1: split-bf16 (fp32-level accuracy).
// B^T (n x l) is formed instead of B (l x n): the A pass then has the same
// A^T-times-skinny shape as the power iteration.
#include <cmath>

#include "pipeline.h"
#include "pipeline_mp.h"

namespace rsvd {

// Y (m x n, fp32) <- orthonormal basis of its range.  passes = 1 (CholeskyQR)
// or 2 (CholeskyQR2).  precise = 1 uses the fp32 matrix cores for the Gram
// and Y R^-1 products, else split-bf16.  Falls back to FP64 Householder on a
// Cholesky breakdown.
static int orth_f32(Context* c, int64_t m, int64_t n, float* Y, int64_t ldy, int passes,
                    int precise, int* fallback) {
  DBuf<float> G32, W, Rinv32;
  DBuf<double> G64, R64, Rinv64, diag0;
  DBuf<int> fail;
  RSVD_TRY(G32.alloc((size_t)(n * n)));
  RSVD_TRY(W.alloc((size_t)(m * n)));
  RSVD_TRY(Rinv32.alloc((size_t)(n * n)));
  RSVD_TRY(G64.alloc((size_t)(n * n)));
  RSVD_TRY(R64.alloc((size_t)(n * n)));
  RSVD_TRY(Rinv64.alloc((size_t)(n * n)));
  RSVD_TRY(fail.alloc(1));
  float* src = Y;
  int64_t lds = ldy;
  for (int pass = 0; pass < passes; ++pass) {
    if (precise)
      RSVD_TRY(gemm_f32(c, 'T', 'N', n, n, m, 1.0f, src, lds, src, lds, 0.0f, G32.p, n));
    else
      RSVD_TRY(gemm_bf16(c, "gram_bf16x3", 'T', 'N', n, n, m, 1.0f, src, lds, src, lds, 0.0f,
                         G32.p, n, 3));
    RSVD_TRY(f32_to_f64(c, n, n, G32.p, n, G64.p, n));
    int ok = 1;
    RSVD_TRY(chol_inv_upper_f64(c, n, G64.p, R64.p, Rinv64.p, &ok));
    if (!ok) {
      // FP64 Householder on the current basis (same range as Y)
      if (fallback) *fallback += 1;
      DBuf<double> Y64;
      RSVD_TRY(Y64.alloc((size_t)(m * n)));
      RSVD_TRY(f32_to_f64(c, m, n, src, lds, Y64.p, m));
      RSVD_TRY(householder_qr_f64(c, m, n, Y64.p, m, nullptr));
      RSVD_TRY(f64_to_f32(c, m, n, Y64.p, m, Y, ldy));
      return 0;
    }
    RSVD_TRY(f64_to_f32(c, n, n, Rinv64.p, n, Rinv32.p, n));
    float* dst = (src == Y) ? W.p : Y;
    const int64_t ldd = (src == Y) ? m : ldy;
    if (precise)
      RSVD_TRY(gemm_f32(c, 'N', 'N', m, n, n, 1.0f, src, lds, Rinv32.p, n, 0.0f, dst, ldd));
    else
      RSVD_TRY(gemm_bf16(c, "orth_bf16x3", 'N', 'N', m, n, n, 1.0f, src, lds, Rinv32.p, n, 0.0f,
                         dst, ldd, 3));
    src = dst;
    lds = ldd;
  }
  if (src != Y) RSVD_TRY(copy2d_f32(c, m, n, src, lds, Y, ldy));
  return 0;
}

int fixed_rank_mp(Context* c, const float* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                  int vnum, int q, int s, uint64_t seed, const float* dOmega, int prec, float* U,
                  int64_t ldu, double* S, float* V, int64_t ldv, MpInfo* info) {
  const int64_t l = (int64_t)k + p;
  if (k <= 0 || p < 0) return set_error("fixed_rank_mp: need k > 0 and p >= 0");
  if (l > (m < n ? m : n)) return set_error("fixed_rank_mp: k+p exceeds min(m,n)");
  if (s <= 0) return set_error("fixed_rank_mp: s must be positive");
  if (prec != MP_BF16 && prec != MP_BF16X3 && prec != MP_FP8)
    return set_error("fixed_rank_mp: unknown precision %d", prec);
  const int split = (prec == MP_BF16X3) ? 3 : 1;
  const char* tag = (prec == MP_FP8) ? "apass_fp8" : (split == 3) ? "apass_bf16x3" : "apass_bf16";
  // fp8: the per-128-column scales of A, once per call
  DBuf<float> ascale;
  if (prec == MP_FP8) {
    RSVD_TRY(ascale.alloc((size_t)((n + 127) / 128)));
    RSVD_TRY(fp8_column_scales(c, m, n, A, lda, ascale.p));
  }
  // one sketch / power pass: C = op(A) S
  auto pass = [&](bool ta, int64_t M, int64_t K, const float* S, int64_t lds_, float* C,
                  int64_t ldc) -> int {
    if (prec == MP_FP8)
      return apass_fp8(c, tag, ta, M, l, K, 1.0f, A, lda, ascale.p, S, lds_, 0.0f, C, ldc);
    return gemm_bf16(c, tag, ta ? 'T' : 'N', 'N', M, l, K, 1.0f, A, lda, S, lds_, 0.0f, C, ldc,
                     split);
  };
  DBuf<float> Om, Y, Z;
  const float* Omp = dOmega;
  if (!Omp) {
    RSVD_TRY(Om.alloc((size_t)(n * l)));
    RSVD_TRY(gaussian_f32(c, seed, n, l, 0, n, Om.p, n));           // rra.c:90-91
    Omp = Om.p;
  }
  RSVD_TRY(Y.alloc((size_t)(m * l)));
  RSVD_TRY(Z.alloc((size_t)(n * l)));
  int fb = 0;
  RSVD_TRY(pass(false, m, n, Omp, n, Y.p, m));
  for (int j = 1; j < q; ++j) {                                        // rra.c:101-126
    if ((2 * j - 2) % s == 0) RSVD_TRY(orth_f32(c, m, l, Y.p, m, 1, 0, &fb));
    RSVD_TRY(pass(true, n, m, Y.p, m, Z.p, n));
    if ((2 * j - 1) % s == 0) RSVD_TRY(orth_f32(c, n, l, Z.p, n, 1, 0, &fb));
    RSVD_TRY(pass(false, m, n, Z.p, n, Y.p, m));
  }
  RSVD_TRY(orth_f32(c, m, l, Y.p, m, 2, 1, &fb));                      // rra.c:129-130
  // B^T = A^T Q (n x l), split-bf16
  DBuf<float> Bt;
  RSVD_TRY(Bt.alloc((size_t)(n * l)));
  RSVD_TRY(gemm_bf16(c, "proj_bf16x3", 'T', 'N', n, l, m, 1.0f, A, lda, Y.p, m, 0.0f, Bt.p, n, 3));
  DBuf<double> Bt64, G, Uh, lam, sig;
  RSVD_TRY(Bt64.alloc((size_t)(n * l)));
  RSVD_TRY(G.alloc((size_t)(l * l)));
  RSVD_TRY(Uh.alloc((size_t)(l * l)));
  RSVD_TRY(lam.alloc((size_t)l));
  RSVD_TRY(sig.alloc((size_t)l));
  RSVD_TRY(f32_to_f64(c, n, l, Bt.p, n, Bt64.p, n));
  int sw = 0;
  DBuf<float> X32;
  RSVD_TRY(X32.alloc((size_t)(l * l)));
  if (vnum == 1 || vnum > 2) {
    // [Qhat, Rhat] = qr(B^T) in FP64, Rhat = Uhat Sigma Vhat^T;
    // U = Q Vhat, V = Qhat Uhat; descending, first k (rra.c:133-174)
    DBuf<double> R, Vh;
    RSVD_TRY(R.alloc((size_t)(l * l)));
    RSVD_TRY(Vh.alloc((size_t)(l * l)));
    int f1 = 0;
    RSVD_TRY(orth_f64(c, n, l, Bt64.p, n, R.p, &f1));
    fb += f1;
    RSVD_TRY(jacobi_f64(c, l, l, R.p, l, Vh.p, l, 60, &sw));
    RSVD_TRY(colnorms_f64(c, l, l, R.p, l, sig.p, 1));
    RSVD_TRY(sort_pairs_f64(c, l, sig.p, 1, R.p, l, l, Vh.p, l, l));
    RSVD_TRY(f64_to_f32(c, l, k, Vh.p, l, X32.p, l));
    RSVD_TRY(gemm_bf16(c, "uv_bf16x3", 'N', 'N', m, k, l, 1.0f, Y.p, m, X32.p, l, 0.0f, U, ldu, 3));
    DBuf<double> V64;
    RSVD_TRY(V64.alloc((size_t)(n * k)));
    RSVD_TRY(gemm_f64(c, 'N', 'N', n, k, l, 1.0, Bt64.p, n, R.p, l, 0.0, V64.p, n));
    RSVD_TRY(f64_to_f32(c, n, k, V64.p, n, V, ldv));
    RSVD_HIP(hipMemcpyAsync(S, sig.p, sizeof(double) * k, hipMemcpyDeviceToDevice, c->stream));
  } else {
    // BB^T = B^T' B^T in FP64; eig; sigma = sqrt(lambda); U = Q Uhat;
    // V = B^T Uhat Sigma^-1; ascending, last k (rra.c:175-225)
    RSVD_TRY(gemm_f64(c, 'T', 'N', l, l, n, 1.0, Bt64.p, n, Bt64.p, n, 0.0, G.p, l));
    RSVD_TRY(symmetrize_upper_f64(c, l, G.p));
    RSVD_TRY(symeig_jacobi_f64(c, l, G.p, l, Uh.p, l, lam.p, 60, &sw));
    RSVD_TRY(sort_pairs_f64(c, l, lam.p, 0, Uh.p, l, l, nullptr, 0, 0));
    RSVD_TRY(sqrt_f64(c, l, lam.p, sig.p));
    const int64_t off = l - k;
    RSVD_TRY(f64_to_f32(c, l, k, Uh.p + off * l, l, X32.p, l));
    RSVD_TRY(gemm_bf16(c, "uv_bf16x3", 'N', 'N', m, k, l, 1.0f, Y.p, m, X32.p, l, 0.0f, U, ldu, 3));
    RSVD_TRY(scale_cols_f64(c, l, l, Uh.p, l, sig.p, 1));
    RSVD_TRY(f64_to_f32(c, l, k, Uh.p + off * l, l, X32.p, l));
    RSVD_TRY(gemm_bf16(c, "uv_bf16x3", 'N', 'N', n, k, l, 1.0f, Bt.p, n, X32.p, l, 0.0f, V, ldv, 3));
    RSVD_HIP(hipMemcpyAsync(S, sig.p + off, sizeof(double) * k, hipMemcpyDeviceToDevice,
                            c->stream));
  }
  if (info) {
    info->qr_fallbacks = fb;
    info->jacobi_sweeps = sw;
  }
  return 0;
}

}  // namespace rsvd

namespace rsvd {

int lowrank_residual(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                     const void* U, int64_t ldu, const double* S, const void* V, int64_t ldv,
                     int uv_f32, double* out) {
  DBuf<double> Us, V64, P, part;
  RSVD_TRY(Us.alloc((size_t)(m * k)));
  RSVD_TRY(V64.alloc((size_t)(n * k)));
  if (uv_f32) {
    RSVD_TRY(f32_to_f64(c, m, k, (const float*)U, ldu, Us.p, m));
    RSVD_TRY(f32_to_f64(c, n, k, (const float*)V, ldv, V64.p, n));
  } else {
    RSVD_TRY(copy2d_f64(c, m, k, (const double*)U, ldu, Us.p, m));
    RSVD_TRY(copy2d_f64(c, n, k, (const double*)V, ldv, V64.p, n));
  }
  RSVD_TRY(scale_cols_f64(c, m, k, Us.p, m, S, 0));
  const int64_t chunk = 2048;
  RSVD_TRY(P.alloc((size_t)(m * (chunk < n ? chunk : n))));
  RSVD_TRY(part.alloc(1));
  double total = 0.0;
  for (int64_t j0 = 0; j0 < n; j0 += chunk) {
    const int64_t w = (n - j0 < chunk) ? n - j0 : chunk;
    RSVD_TRY(gemm_f64(c, 'N', 'T', m, w, k, 1.0, Us.p, m, V64.p + j0, n, 0.0, P.p, m));
    RSVD_TRY(sumsq_diff_f64(c, m, w, A + j0 * lda, lda, P.p, m, part.p));
    double h = 0;
    RSVD_HIP(hipMemcpyAsync(&h, part.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
    RSVD_HIP(hipStreamSynchronize(c->stream));
    total += h;
  }
  *out = sqrt(total);
  return 0;
}

}  // namespace rsvd
