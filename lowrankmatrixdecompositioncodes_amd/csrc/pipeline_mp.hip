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

// *out = max_ij |G_ij - delta_ij| of an n x n FP64 matrix (one workgroup).
// max |G - I| of an n x n FP64 matrix (leading dimension n): one workgroup
// per column writes that column's maximum, then one workgroup takes the
// maximum of those (fmax is exact, so the split does not change the value).
__device__ __forceinline__ double block_max_256(double mx, double* red) {
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  return red[0];
}

__global__ void col_dev_identity_kernel(int64_t n, const double* __restrict__ G,
                                        double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t j = blockIdx.x;
  double mx = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256)
    mx = fmax(mx, fabs(G[i + j * n] - (i == j ? 1.0 : 0.0)));
  mx = block_max_256(mx, red);
  if (threadIdx.x == 0) part[j] = mx;
}

__global__ void max_of_kernel(int64_t n, const double* __restrict__ part,
                              double* __restrict__ out) {
  __shared__ double red[256];
  double mx = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) mx = fmax(mx, part[i]);
  mx = block_max_256(mx, red);
  if (threadIdx.x == 0) *out = mx;
}

// Y (m x n, fp32) <- orthonormal basis of its range.  passes = 1 (CholeskyQR)
// or 2 (CholeskyQR2).  precise = 1 forms the last pass's Gram on the fp32
// matrix cores and its Y R^-1 with fp32-exact products (three bf16 planes per
// operand); every other product is split-bf16 (the last pass corrects the
// earlier ones).  The last pass's Gram Q1^T Q1 is checked: if it is more than
// 0.25 from the identity (the split-bf16 first pass met a basis too
// ill-conditioned for its 16-bit products), one more exact pass follows, at
// most two.  Falls back to FP64 Householder on a Cholesky breakdown.
static int orth_f32(Context* c, int64_t m, int64_t n, float* Y, int64_t ldy, int passes,
                    int precise, int* fallback) {
  DBuf<float> G32, W, Rinv32;
  DBuf<double> G64, R64, Rinv64, diag0, dev;
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
  int total = passes;
  for (int pass = 0; pass < total; ++pass) {
    // the last pass fixes what the ones before it leave: only its products
    // need to be exact
    const bool exact = precise && pass + 1 >= passes;
    if (exact)
      RSVD_TRY(gemm_f32(c, 'T', 'N', n, n, m, 1.0f, src, lds, src, lds, 0.0f, G32.p, n));
    else
      RSVD_TRY(gemm_bf16(c, "gram_bf16x3", 'T', 'N', n, n, m, 1.0f, src, lds, src, lds, 0.0f,
                         G32.p, n, 3));
    RSVD_TRY(f32_to_f64(c, n, n, G32.p, n, G64.p, n));
    if (exact && pass > 0 && pass + 1 == total && total < passes + 2) {
      if (!dev.p) RSVD_TRY(dev.alloc((size_t)n + 1));  // result, then n column maxima
      hipLaunchKernelGGL(col_dev_identity_kernel, dim3((unsigned)n), dim3(256), 0, c->stream, n,
                         (const double*)G64.p, dev.p + 1);
      RSVD_LAUNCH_CHECK();
      hipLaunchKernelGGL(max_of_kernel, dim3(1), dim3(256), 0, c->stream, n,
                         (const double*)(dev.p + 1), dev.p);
      RSVD_LAUNCH_CHECK();
      double h = 0.0;
      RSVD_TRY(read_small(c, &h, dev.p, sizeof h));
      if (!(h <= 0.25)) ++total;  // (NaN included)
    }
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
    if (exact)
      RSVD_TRY(gemm_bf16_general(c, "orth_bf16x6", 'N', 'N', m, n, n, 1.0f, src, lds, Rinv32.p, n,
                                 0.0f, dst, ldd, 6));
    else
      RSVD_TRY(gemm_bf16(c, "orth_bf16x3", 'N', 'N', m, n, n, 1.0f, src, lds, Rinv32.p, n, 0.0f,
                         dst, ldd, 3));
    src = dst;
    lds = ldd;
  }
  if (src != Y) RSVD_TRY(copy2d_f32(c, m, n, src, lds, Y, ldy));
  return 0;
}

// U (m x k), S (k), V (n x k) from an orthonormal Q (m x l, fp32) of the fp32
// A: B^T = A^T Q in split-bf16, then the vnum branch of rra.c:133-225 with
// the small problem in FP64 and U, V products in split-bf16.
int svd_from_Q_mp(Context* c, const float* A, int64_t m, int64_t n, int64_t lda, const float* Q,
                  int64_t ldq, int64_t l, int64_t k, int vnum, const float* Bt_in, int64_t ldbt,
                  int exact, float* U, int64_t ldu, double* S, float* V, int64_t ldv, int* sw,
                  int* fb, const float* fp16_ascale) {
  if (k <= 0 || k > l) return set_error("svd_from_Q_mp: k=%lld outside [1, l=%lld]",
                                        (long long)k, (long long)l);
  if ((vnum == 1 || vnum > 2) && n < l)
    return set_error("vnum=1 needs n >= l (n=%lld, l=%lld)", (long long)n, (long long)l);
  // B^T = A^T Q (n x l), split-bf16, unless the caller has it
  const int usplit = exact ? 6 : 3;
  const char* utag = exact ? "uv_bf16x6" : "uv_bf16x3";
  DBuf<float> Btb;
  const float* Btp = Bt_in;
  int64_t ldb = ldbt;
  if (!Btp) {
    RSVD_TRY(Btb.alloc((size_t)(n * l)));
    if (fp16_ascale)
      RSVD_TRY(apass_fp16(c, "proj_fp16", true, n, l, m, 1.0f, A, lda, fp16_ascale, Q, ldq, 0.0f,
                          Btb.p, n));
    else
      RSVD_TRY(gemm_bf16(c, "proj_bf16x3", 'T', 'N', n, l, m, 1.0f, A, lda, Q, ldq, 0.0f, Btb.p,
                         n, 3));
    Btp = Btb.p;
    ldb = n;
  }
  DBuf<double> Bt64, G, Uh, lam, sig;
  RSVD_TRY(Bt64.alloc((size_t)(n * l)));
  RSVD_TRY(G.alloc((size_t)(l * l)));
  RSVD_TRY(Uh.alloc((size_t)(l * l)));
  RSVD_TRY(lam.alloc((size_t)l));
  RSVD_TRY(sig.alloc((size_t)l));
  RSVD_TRY(f32_to_f64(c, n, l, Btp, ldb, Bt64.p, n));
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
    *fb += f1;
    RSVD_TRY(jacobi_f64(c, l, l, R.p, l, Vh.p, l, 60, sw));
    RSVD_TRY(colnorms_f64(c, l, l, R.p, l, sig.p, 1));
    RSVD_TRY(sort_pairs_f64(c, l, sig.p, 1, R.p, l, l, Vh.p, l, l));
    RSVD_TRY(f64_to_f32(c, l, k, Vh.p, l, X32.p, l));
    RSVD_TRY(gemm_bf16_general(c, utag, 'N', 'N', m, k, l, 1.0f, Q, ldq, X32.p, l, 0.0f, U, ldu,
                               usplit));
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
    // U and V leave as fp32: the looser sweep test is enough
    RSVD_TRY(symeig_jacobi_f64(c, l, G.p, l, Uh.p, l, lam.p, 60, sw, 1));
    RSVD_TRY(sort_pairs_f64(c, l, lam.p, 0, Uh.p, l, l, nullptr, 0, 0));
    RSVD_TRY(sqrt_f64(c, l, lam.p, sig.p));
    const int64_t off = l - k;
    RSVD_TRY(f64_to_f32(c, l, k, Uh.p + off * l, l, X32.p, l));
    RSVD_TRY(gemm_bf16_general(c, utag, 'N', 'N', m, k, l, 1.0f, Q, ldq, X32.p, l, 0.0f, U, ldu,
                               usplit));
    RSVD_TRY(scale_cols_f64(c, l, l, Uh.p, l, sig.p, 1));
    RSVD_TRY(f64_to_f32(c, l, k, Uh.p + off * l, l, X32.p, l));
    RSVD_TRY(gemm_bf16_general(c, utag, 'N', 'N', n, k, l, 1.0f, Btp, ldb, X32.p, l, 0.0f, V, ldv,
                               usplit));
    RSVD_HIP(hipMemcpyAsync(S, sig.p + off, sizeof(double) * k, hipMemcpyDeviceToDevice,
                            c->stream));
  }
  return 0;
}

// One pass C = op(A) S (N columns) at the sketch precision; ascale: the fp8
// or fp16 chunk scales of A (prec = MP_FP8 / MP_FP16 only).
static int mp_pass(Context* c, int prec, bool ta, int64_t M, int64_t N, int64_t K, const float* A,
                   int64_t lda, const float* ascale, const float* S, int64_t lds, float* C,
                   int64_t ldc) {
  if (prec == MP_FP8)
    return apass_fp8(c, "apass_fp8", ta, M, N, K, 1.0f, A, lda, ascale, S, lds, 0.0f, C, ldc);
  if (prec == MP_FP16)
    return apass_fp16(c, "apass_fp16", ta, M, N, K, 1.0f, A, lda, ascale, S, lds, 0.0f, C, ldc);
  const int split = (prec == MP_BF16X3) ? 3 : 1;
  return gemm_bf16(c, split == 3 ? "apass_bf16x3" : "apass_bf16", ta ? 'T' : 'N', 'N', M, N, K,
                   1.0f, A, lda, S, lds, 0.0f, C, ldc, split);
}

int fixed_rank_mp(Context* c, const float* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                  int vnum, int q, int s, uint64_t seed, const float* dOmega, int prec, float* U,
                  int64_t ldu, double* S, float* V, int64_t ldv, MpInfo* info) {
  const int64_t l = (int64_t)k + p;
  if (k <= 0 || p < 0) return set_error("fixed_rank_mp: need k > 0 and p >= 0");
  if (l > (m < n ? m : n)) return set_error("fixed_rank_mp: k+p exceeds min(m,n)");
  if (s <= 0) return set_error("fixed_rank_mp: s must be positive");
  if (prec < MP_BF16 || prec > MP_FP16)
    return set_error("fixed_rank_mp: unknown precision %d", prec);
  // precision of pass i of npass (the sketch, then 2(q-1) power passes): the
  // two-level modes run every pass but the last in bf16 / fp8 and the last,
  // whose range becomes Q, in split-bf16
  const int npass = 1 + 2 * (q > 1 ? q - 1 : 0);
  auto prec_of = [&](int i) -> int {
    if (prec <= MP_FP8 || prec == MP_FP16) return prec;
    if (i + 1 == npass) return MP_BF16X3;
    return prec == MP_BF16_SPLIT ? MP_BF16 : MP_FP8;
  };
  // fp8 / fp16: the per-128-column scales of A, once per call
  DBuf<float> ascale;
  bool any_fp8 = false;
  for (int i = 0; i < npass; ++i) any_fp8 = any_fp8 || prec_of(i) == MP_FP8;
  if (any_fp8 || prec == MP_FP16) {
    RSVD_TRY(ascale.alloc((size_t)((n + 127) / 128)));
    if (prec == MP_FP16) RSVD_TRY(fp16_column_scales(c, m, n, A, lda, ascale.p));
    else RSVD_TRY(fp8_column_scales(c, m, n, A, lda, ascale.p));
  }
  // one sketch / power pass: C = op(A) S
  int ip = 0;
  auto pass = [&](bool ta, int64_t M, int64_t K, const float* S, int64_t lds_, float* C,
                  int64_t ldc) -> int {
    return mp_pass(c, prec_of(ip++), ta, M, l, K, A, lda, ascale.p, S, lds_, C, ldc);
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
  int sw = 0;
  RSVD_TRY(svd_from_Q_mp(c, A, m, n, lda, Y.p, m, l, k, vnum, nullptr, 0, 0, U, ldu, S, V, ldv,
                         &sw, &fb, prec == MP_FP16 ? ascale.p : nullptr));
  if (info) {
    info->qr_fallbacks = fb;
    info->jacobi_sweeps = sw;
  }
  return 0;
}

__global__ void add2d_f32_kernel(int64_t rows, int64_t cols, const float* __restrict__ src,
                                 int64_t lds, float* __restrict__ dst, int64_t ldd) {
  const int64_t total = rows * cols;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % rows, j = idx / rows;
    dst[i + j * ldd] += src[i + j * lds];
  }
}

// dst += src (rows x cols, fp32)
static int add2d_f32(Context* c, int64_t rows, int64_t cols, const float* src, int64_t lds,
                     float* dst, int64_t ldd) {
  if (rows * cols == 0) return 0;
  int64_t b = (rows * cols + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(add2d_f32_kernel, dim3((unsigned)b), dim3(256), 0, c->stream, rows, cols, src,
                     lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int qb_blocked_mp(Context* c, const float* M, int64_t m, int64_t n, int64_t ldm, int kstep,
                  int nstep, double TOL, int q, int s, int is_new, uint64_t seed,
                  const float* dRN, int prec, float* Q, int64_t ldq, float* Bt, int64_t ldbt,
                  int* frank, int* l_alloc, QbInfo* info) {
  int tolMode = 0;
  RSVD_TRY(qb_adjust(m, n, &kstep, &nstep, is_new, &tolMode));
  if (s <= 0) return set_error("randQB_mp: s must be positive (s=%d)", s);
  if (prec != MP_BF16 && prec != MP_BF16X3 && prec != MP_FP8)
    return set_error("randQB_mp: precision %d not available for the blocked QB", prec);
  const int64_t l = (int64_t)kstep * nstep;
  if (l_alloc) *l_alloc = (int)l;
  if (kstep > m) return set_error("randQB_mp: kstep=%d exceeds m=%lld", kstep, (long long)m);
  DBuf<float> R, Yp, AtQp, Bp, RN, QtQ, ascale;
  DBuf<double> Y64, sums;
  RSVD_TRY(R.alloc((size_t)(m * n)));
  RSVD_TRY(Yp.alloc((size_t)(m * kstep)));
  RSVD_TRY(AtQp.alloc((size_t)(n * kstep)));
  RSVD_TRY(Y64.alloc((size_t)(m * kstep)));
  RSVD_TRY(sums.alloc(1));
  RSVD_TRY(Bp.alloc((size_t)(n * 2 * kstep)));
  if (!dRN) RSVD_TRY(RN.alloc((size_t)(n * kstep)));
  if (prec == MP_FP8) RSVD_TRY(ascale.alloc((size_t)((n + 127) / 128)));
  RSVD_TRY(copy2d_f32(c, m, n, M, ldm, R.p, m));                      // rra.c:1630
  int fr = 0, fb = 0, f1 = 0, steps = 0;
  double normval = 0;
  for (int step = 0; step < nstep; ++step) {
    const int64_t c0 = (int64_t)step * kstep;
    const float* RNp = dRN ? dRN + c0 * n : RN.p;                      // rra.c:1639-1642
    if (!dRN) RSVD_TRY(gaussian_f32(c, seed, n, kstep, c0, n, RN.p, n));
    if (prec == MP_FP8) RSVD_TRY(fp8_column_scales(c, m, n, R.p, m, ascale.p));
    RSVD_TRY(mp_pass(c, prec, false, m, kstep, n, R.p, m, ascale.p, RNp, n, Yp.p, m));
    for (int j = 1; j <= q; ++j) {                                      // rra.c:1652-1686
      if ((2 * j - 2) % s == 0) RSVD_TRY(orth_f32(c, m, kstep, Yp.p, m, 1, 0, &fb));
      RSVD_TRY(mp_pass(c, prec, true, n, kstep, m, R.p, m, ascale.p, Yp.p, m, AtQp.p, n));
      if ((2 * j - 1) % s == 0) RSVD_TRY(orth_f32(c, n, kstep, AtQp.p, n, 1, 0, &fb));
      RSVD_TRY(mp_pass(c, prec, false, m, kstep, n, R.p, m, ascale.p, AtQp.p, n, Yp.p, m));
    }
    // Qp in FP64 (CholeskyQR2, Householder fallback), rra.c:1690; its fp32
    // rounding is orthonormal to ~1e-9, so R - Qp (Qp^T R) removes the range
    // of Qp from R to the rounding of R itself
    RSVD_TRY(f32_to_f64(c, m, kstep, Yp.p, m, Y64.p, m));
    RSVD_TRY(orth_f64(c, m, kstep, Y64.p, m, nullptr, &f1));
    fb += f1;
    if (step > 0) {
      // against every earlier block (rra.c:1703-1722 does so every second
      // step): R is held in fp32, so its rounding (~1e-7 ||A||) is a growing
      // share of R as R shrinks towards TOL and leaks into the new block
      RSVD_TRY(f64_to_f32(c, m, kstep, Y64.p, m, Yp.p, m));
      RSVD_TRY(QtQ.alloc((size_t)(c0 * kstep)));
      RSVD_TRY(gemm_bf16_general(c, "reorth_bf16x6", 'T', 'N', c0, kstep, m, 1.0f, Q, ldq, Yp.p,
                                 m, 0.0f, QtQ.p, c0, 6));
      RSVD_TRY(gemm_bf16_general(c, "reorth_bf16x6", 'N', 'N', m, kstep, c0, -1.0f, Q, ldq, QtQ.p,
                                 c0, 1.0f, Yp.p, m, 6));
      RSVD_TRY(f32_to_f64(c, m, kstep, Yp.p, m, Y64.p, m));
      RSVD_TRY(orth_f64(c, m, kstep, Y64.p, m, nullptr, &f1));
      fb += f1;
    }
    float* Qp = Q + c0 * ldq;
    RSVD_TRY(f64_to_f32(c, m, kstep, Y64.p, m, Qp, ldq));               // rra.c:1760
    // Bp^T = R^T Qp (split-bf16), rra.c:1741, and R -= Qp Bp with fp32-exact
    // products and the FP64 sum of squares of the new R, rra.c:1750 + the TOL
    // norm of rra.c:1771-1778.  Bp carries the split-bf16 error of a long sum
    // (~1e-6 relative); that error lies in range(Qp), where later blocks,
    // orthogonal to Qp, never reach it.  So from the second block on the same
    // pass also projects R on the previous block (R^T [Q_{p-1} Qp]) and the
    // deflation removes both: the previous block's residual share is cleared
    // one step late, to the accuracy of its (small) correction.
    const int64_t nb = step > 0 ? 2 * kstep : kstep;   // blocks in this deflation
    const float* Qd = Q + (c0 + kstep - nb) * ldq;
    RSVD_TRY(gemm_bf16(c, "proj_bf16x3", 'T', 'N', n, nb, m, 1.0f, R.p, m, Qd, ldq, 0.0f, Bp.p, n,
                       3));
    // the previous block's correction is ~1e-6 of Bp: its bf16 product is
    // exact enough, the new block's needs fp32-exact products
    const int64_t klo = (nb > kstep && kstep % 32 == 0) ? kstep : 0;
    RSVD_TRY(gemm_bf16_deflate(c, m, n, nb, Qd, ldq, Bp.p, n, R.p, m, sums.p, klo));
    if (Bt) {  // B rows of this block, and the previous block's correction
      RSVD_TRY(copy2d_f32(c, n, kstep, Bp.p + (nb - kstep) * n, n, Bt + c0 * ldbt, ldbt));
      if (step > 0) RSVD_TRY(add2d_f32(c, n, kstep, Bp.p, n, Bt + (c0 - kstep) * ldbt, ldbt));
    }
    fr = (step + 1) * kstep;                                            // rra.c:1770
    steps = step + 1;
    if (tolMode) {
      double ss = 0;
      RSVD_TRY(read_small(c, &ss, sums.p, sizeof ss));
      normval = sqrt(ss);
      if (normval < TOL) break;
    }
  }
  if (!tolMode && steps > 0) {
    double ss = 0;
    RSVD_TRY(read_small(c, &ss, sums.p, sizeof ss));
    normval = sqrt(ss);
  }
  if (frank) *frank = fr;
  if (info) {
    info->steps = steps;
    info->qr_fallbacks = fb;
    info->last_norm = normval;
    info->kstep = kstep;
    info->nstep = nstep;
    info->tol_mode = tolMode;
  }
  return 0;
}

int blockrand_mp(Context* c, const float* M, int64_t m, int64_t n, int64_t ldm, int k, int p,
                 double TOL, int vnum, int kstep, int q, int s, uint64_t seed, const float* dRN,
                 int prec, BlockrandMpOut* out) {
  int p_eff = 0, nstep = 0, rankMode = 0;
  RSVD_TRY(blockrand_plan(m, n, k, p, kstep, &p_eff, &nstep, &rankMode));
  p = p_eff;
  int ks = kstep, ns = nstep, tm = 0;
  RSVD_TRY(qb_adjust(m, n, &ks, &ns, 1, &tm));
  const int64_t lmax = (int64_t)ks * ns;
  DBuf<float> Q, Bt;
  RSVD_TRY(Q.alloc((size_t)(m * lmax)));
  RSVD_TRY(Bt.alloc((size_t)(n * lmax)));
  int frank = 0, lalloc = 0;
  QbInfo qi{};
  RSVD_TRY(qb_blocked_mp(c, M, m, n, ldm, kstep, nstep, TOL, q, s, 1, seed, dRN, prec, Q.p, m,
                         Bt.p, n, &frank, &lalloc, &qi));
  const int64_t l = qi.tol_mode ? frank : lalloc;                       // B->nrows
  if (rankMode == 1) frank = k;                                         // rra.c:273-279
  else frank = (int)llround((double)(frank / (frank + p + 1e-6)) * frank);
  out->frank = frank;
  out->l = (int)l;
  out->qb = qi;
  if (frank <= 0 || frank > l)
    return set_error("blockrand_mp: rank %d outside [1, l=%lld] (k=%d p=%d kstep=%d)", frank,
                     (long long)l, k, p, kstep);
  RSVD_TRY(out->U.alloc((size_t)(m * frank)));
  RSVD_TRY(out->S.alloc((size_t)frank));
  RSVD_TRY(out->V.alloc((size_t)(n * frank)));
  int sw = 0, fb = 0;
  // rra.c:288-380 forms B = Q^T M afresh; here B comes from the QB, where
  // every block's share was corrected once (a fresh split-bf16 pass over A
  // would carry a ~1e-6 relative error into a result held to TOL)
  RSVD_TRY(svd_from_Q_mp(c, M, m, n, ldm, Q.p, m, l, frank, vnum, Bt.p, n, 1, out->U.p, m,
                         out->S.p, out->V.p, n, &sw, &fb));
  out->svd.l = l;
  out->svd.jacobi_sweeps = sw;
  out->svd.qr_fallbacks = fb;
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
    RSVD_TRY(read_small(c, &h, part.p, sizeof h));
    total += h;
  }
  *out = sqrt(total);
  return 0;
}

}  // namespace rsvd
