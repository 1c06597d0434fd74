// qr.hip -- tall-skinny orthonormalisation for the rSVD / QB pipeline,
// replacing QR_factorization_getQ and compact_QR_factorization
// (MKL/mvf.c:1214-1263, dgeqrf + dorgqr).
//
// Main path: CholeskyQR2.  G = Y^T Y on the matrix cores (split-K GEMM), FP64
// blocked Cholesky G = R^T R fused with the triangular inverse (one launch per
// 32-row panel), Q1 = Y R^-1 (GEMM), and the same once more on Q1; R = R2 R1.  Breakdown (a pivot below
// tol * original diagonal, or non-finite) is reported back to the host and the
// column-by-column Householder QR (geqr2 + org2r, LAPACK sign convention)
// runs instead.
#include <cmath>

#include "rsvd_internal.h"

namespace rsvd {

constexpr double kCholRelTol = 1e-13;   // pivot / original diagonal

// ---------------------------------------------------------------- Householder
// Reflector for column j (rows j..m-1) of Y, LAPACK dlarfg convention:
// beta = -sign(alpha)*||x||, tau = (beta-alpha)/beta, v = x/(alpha-beta), v0=1.
__global__ void __launch_bounds__(1024)
hh_reflector_kernel(int64_t m, double* __restrict__ Y, int64_t ldy, int64_t j,
                    double* __restrict__ tau) {
  __shared__ double red[16];
  __shared__ double sh_scale;
  double* x = Y + j + j * ldy;
  const int64_t len = m - j;
  double s = 0;
  for (int64_t i = 1 + threadIdx.x; i < len; i += 1024) s += x[i] * x[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sig = 0;
    for (int w = 0; w < 16; ++w) sig += red[w];
    double alpha = x[0];
    if (sig == 0.0) {
      tau[j] = 0.0;
      sh_scale = 0.0;
    } else {
      double nrm = sqrt(alpha * alpha + sig);
      double beta = (alpha >= 0) ? -nrm : nrm;
      tau[j] = (beta - alpha) / beta;
      sh_scale = 1.0 / (alpha - beta);
      x[0] = beta;
    }
  }
  __syncthreads();
  const double sc = sh_scale;
  if (sc != 0.0)
    for (int64_t i = 1 + threadIdx.x; i < len; i += 1024) x[i] *= sc;
}

// Apply H_j = I - tau v v^T (v stored below the diagonal of column j, v0 = 1)
// to columns c0.. of Y restricted to rows j..m-1.  One workgroup per column.
__global__ void __launch_bounds__(256)
hh_apply_kernel(int64_t m, double* __restrict__ Y, int64_t ldy, int64_t j, int64_t c0,
                const double* __restrict__ tau, const double* __restrict__ V, int64_t ldv) {
  __shared__ double red[4];
  const int64_t col = c0 + blockIdx.x;
  const double t = tau[j];
  if (t == 0.0) return;
  const double* v = V + j + j * ldv;
  double* y = Y + j + col * ldy;
  const int64_t len = m - j;
  double s = 0;
  for (int64_t i = threadIdx.x; i < len; i += 256) s += (i == 0 ? 1.0 : v[i]) * y[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  const double w = t * (red[0] + red[1] + red[2] + red[3]);
  for (int64_t i = threadIdx.x; i < len; i += 256) y[i] -= (i == 0 ? 1.0 : v[i]) * w;
}

// org2r: column j of Q = H_j e_j after the later reflectors were applied.
__global__ void hh_form_col_kernel(int64_t m, double* __restrict__ Y, int64_t ldy, int64_t j,
                                   const double* __restrict__ tau) {
  const double t = tau[j];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    double* q = Y + j * ldy;
    if (i < j) q[i] = 0.0;
    else if (i == j) q[i] = 1.0 - t;
    else q[i] = -t * q[i];
  }
}

__global__ void copy_upper_kernel(int64_t n, const double* __restrict__ Y, int64_t ldy,
                                  double* __restrict__ R) {
  int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % n, j = idx / n;
    R[idx] = (i <= j) ? Y[i + j * ldy] : 0.0;
  }
}

int householder_qr_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* R) {
  if (n == 0 || m == 0) return 0;
  if (m < n) return set_error("householder_qr: m=%lld < n=%lld", (long long)m, (long long)n);
  DBuf<double> tau;
  RSVD_TRY(tau.alloc((size_t)n));
  for (int64_t j = 0; j < n; ++j) {
    hipLaunchKernelGGL(hh_reflector_kernel, dim3(1), dim3(1024), 0, c->stream, m, Y, ldy, j,
                       tau.p);
    RSVD_LAUNCH_CHECK();
    if (j + 1 < n) {
      hipLaunchKernelGGL(hh_apply_kernel, dim3((unsigned)(n - j - 1)), dim3(256), 0, c->stream, m,
                         Y, ldy, j, j + 1, (const double*)tau.p, (const double*)Y, ldy);
      RSVD_LAUNCH_CHECK();
    }
  }
  if (R) {
    int blocks = (int)((n * n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(copy_upper_kernel, dim3(blocks), dim3(256), 0, c->stream, n,
                       (const double*)Y, ldy, R);
    RSVD_LAUNCH_CHECK();
  }
  return org2r_f64(c, m, n, Y, ldy, tau.p);
}

int org2r_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, const double* tau) {
  if (n <= 0 || m <= 0) return 0;
  for (int64_t j = n - 1; j >= 0; --j) {
    if (j + 1 < n) {
      hipLaunchKernelGGL(hh_apply_kernel, dim3((unsigned)(n - j - 1)), dim3(256), 0, c->stream, m,
                         Y, ldy, j, j + 1, tau, (const double*)Y, ldy);
      RSVD_LAUNCH_CHECK();
    }
    int blocks = (int)((m + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(hh_form_col_kernel, dim3(blocks), dim3(256), 0, c->stream, m, Y, ldy, j,
                       tau);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

// ---------------------------------------------------------------- Cholesky + inverse
// Right-looking blocked Cholesky G = R^T R fused with X = R^-1, one launch
// per 32-row panel k (p = 32 k).  The serial chain is the panel count; every
// workgroup does O(1) tile work, so no launch has a loop over earlier panels.
// Each workgroup first factors the panel's diagonal block D = G_kk (already
// updated by panels < k) in LDS -- redundantly, which is cheaper than another
// launch -- then takes one role:
//   trailing (ti <= tj > k): solve R_kk^T R_k,t = G_k,t for t = ti, tj and
//       update G_ti,tj -= R_k,ti^T R_k,tj in place; the ti == tj tile writes
//       R_k,ti (and, in panel 0, the original diagonal for the pivot test);
//   accumulate (i <= k, tj > k): X_ik = -Acc_ik X_kk (X_kk when i == k) and
//       Acc_i,tj (+)= X_ik R_k,tj, so that Acc_ij = sum_{k=i}^{j-1} X_ik R_kj
//       when panel j starts;
//   finalize (i <= k): X_ik = -Acc_ik X_kk, or for i == k: X_kk, R_kk and the
//       zeros below both, and the pivot flag.
// X_ij = -(sum_{k=i}^{j-1} X_ik R_kj) X_jj is the column form of the
// triangular inverse.  Inside a launch no tile is both read and written.
namespace {

constexpr int CB = 32, CHP = CB + 1;

__device__ __forceinline__ void ci_load(const double* __restrict__ M, int64_t ld, int64_t r0,
                                        int64_t c0, int rows, int cols, double (*T)[CHP]) {
  for (int e = threadIdx.x; e < CB * CB; e += 256) {
    const int i = e & (CB - 1), j = e >> 5;
    T[i][j] = (i < rows && j < cols) ? M[(r0 + i) + (c0 + j) * ld] : 0.0;
  }
}

// 1 / d to about one ulp: the hardware reciprocal estimate and two Newton
// steps (shorter than a correctly rounded division on the serial path).
__device__ __forceinline__ double ci_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

// Unnormalised right-looking Cholesky of Dm (upper triangle valid, kb x kb):
// step j subtracts U_j,a U_j,c / U_j,j from the rows below, one barrier per
// step.  On exit Dm holds R_kk (zero below the diagonal and past kb) and
// sqd / invd its diagonal and reciprocals.  A pivot that is not above
// kCholRelTol * ref (or is not finite) sets *bad and is replaced (by 1 when
// it is not positive) so the factorisation runs to the end.
__device__ void ci_factor(double (*Dm)[CHP], int kb, const double* ref, double* piv,
                          double* sqd, double* invd, int* bad) {
  const int tid = threadIdx.x, a = tid & (CB - 1), c0 = (tid >> 5) * 4;
  double own[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) own[q] = Dm[a][c0 + q];
  for (int j = 0; j < kb; ++j) {
    double d = Dm[j][j];
    const bool ok = (d > kCholRelTol * ref[j]) && isfinite(d);
    if (!ok) d = (d > 0.0 && isfinite(d)) ? d : 1.0;
    if (tid == 0) {
      if (!ok) *bad = 1;
      piv[j] = d;
    }
    if (a > j && a < kb) {
      const double u = Dm[j][a] * ci_rcp(d);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (c0 + q >= a && c0 + q < kb) own[q] -= u * Dm[j][c0 + q];
      if (a == j + 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (c0 + q >= a && c0 + q < kb) Dm[a][c0 + q] = own[q];
      }
    }
    __syncthreads();
  }
  if (tid < kb) {
    const double sj = sqrt(piv[tid]);
    sqd[tid] = sj;
    invd[tid] = 1.0 / sj;
  }
  __syncthreads();
  // normalise: R[j][c] = U[j][c] / sqrt(U[j][j])
  for (int e = tid; e < CB * CB; e += 256) {
    const int i = e & (CB - 1), c = e >> 5;
    double v = 0.0;
    if (i < kb && c < kb) v = (c > i) ? Dm[i][c] * invd[i] : (c == i ? sqd[i] : 0.0);
    Dm[i][c] = v;
  }
  __syncthreads();
}

// Column c of R_kk^-1 into Xk by back substitution, one thread per column,
// the column in registers: x_r = -(sum_{u > r} R_ru x_u) / R_rr with x_c =
// 1 / R_cc and x_r = 0 for r > c (R_kk is zero past kb, so the padding
// columns come out zero).  Fully unrolled, so the R_kk reads (the same
// address in every lane) are independent of the recurrence and the serial
// path is one multiply-add per row.
__device__ void ci_inv_col(const double (*Rk)[CHP], const double* invd, int kb, int c,
                           double (*Xk)[CHP]) {
  double x[CB];
#pragma unroll
  for (int r = CB - 1; r >= 0; --r) {
    double s = 0.0;
#pragma unroll
    for (int u = CB - 1; u > r; --u) s = fma(Rk[r][u], x[u], s);
    const double ir = (r < kb) ? invd[r] : 0.0;
    x[r] = (r > c) ? 0.0 : ((r == c) ? ir : -s * ir);
  }
#pragma unroll
  for (int r = 0; r < CB; ++r) Xk[r][c] = (c < kb) ? x[r] : 0.0;
}

// Column c of P <- R_kk^-T P by forward substitution, one thread per column,
// fully unrolled (the column in registers; the last term of each row's sum is
// the one just computed).
__device__ void ci_fwd_col(const double (*Rk)[CHP], const double* invd, int kb, int c,
                           double (*P)[CHP]) {
  double y[CB];
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    double s = P[j][c];
#pragma unroll
    for (int i = 0; i < j; ++i) s = fma(-Rk[i][j], y[i], s);
    y[j] = (j < kb) ? s * invd[j] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < CB; ++j) P[j][c] = y[j];
}

// acc[q] += sum_j op(A)[a][j] B[j][c0 + q] for thread (a, c0) over the padded
// 32 x 32 tiles (op(A) = A^T when TA).
template <bool TA>
__device__ __forceinline__ void ci_mma(const double (*A)[CHP], const double (*B)[CHP],
                                       double (&acc)[4]) {
  const int a = threadIdx.x & (CB - 1), c0 = (threadIdx.x >> 5) * 4;
#pragma unroll 8
  for (int j = 0; j < CB; ++j) {
    const double av = TA ? A[j][a] : A[a][j];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += av * B[j][c0 + q];
  }
}

__global__ void __launch_bounds__(256)
cholinv_panel_kernel(int64_t n, double* __restrict__ G, double* __restrict__ R,
                     double* __restrict__ X, double* __restrict__ Acc,
                     double* __restrict__ diag0, int k, int* __restrict__ fail) {
  __shared__ double Dm[CB][CHP], Xk[CB][CHP], T1[CB][CHP], T2[CB][CHP], T3[CB][CHP],
      T4[CB][CHP];
  __shared__ double refd[CB], piv[CB], sqd[CB], invd[CB];
  __shared__ int bad;
  const int tid = threadIdx.x, a = tid & (CB - 1), c0 = (tid >> 5) * 4;
  const int64_t p = (int64_t)k * CB;
  const int kb = (int)((n - p) < CB ? (n - p) : CB);
  const int nb = (int)((n + CB - 1) / CB);
  const int T = nb - k - 1;
  const int nt = T * (T + 1) / 2, na = (k + 1) * T;
  auto bw = [&](int b) { return (int)((n - (int64_t)b * CB) < CB ? (n - (int64_t)b * CB) : CB); };
  int role, bid = blockIdx.x, ti = 0, tj = 0, i = 0;
  if (bid < nt) {
    role = 0;
    int u = 0;
    while (bid >= T - u) {
      bid -= T - u;
      ++u;
    }
    ti = k + 1 + u;
    tj = ti + bid;
  } else if (bid < nt + na) {
    role = 1;
    bid -= nt;
    i = bid / T;
    tj = k + 1 + bid % T;
  } else {
    role = 2;
    i = bid - nt - na;
  }
  const int64_t pi = (int64_t)i * CB, pti = (int64_t)ti * CB, ptj = (int64_t)tj * CB;

  // ---- loads (all issued before the serial factorisation)
  ci_load(G, n, p, p, kb, kb, Dm);
  if (role == 0) {
    ci_load(G, n, p, pti, kb, bw(ti), T1);
    if (tj != ti) ci_load(G, n, p, ptj, kb, bw(tj), T2);
    ci_load(G, n, pti, ptj, bw(ti), bw(tj), T3);
  } else if (role == 1) {
    ci_load(G, n, p, ptj, kb, bw(tj), T1);
    if (i < k) {
      ci_load(Acc, n, pi, p, CB, kb, T2);
      ci_load(Acc, n, pi, ptj, CB, bw(tj), T3);
    }
  } else if (i < k) {
    ci_load(Acc, n, pi, p, CB, kb, T2);
  }
  if (tid < kb) refd[tid] = (k == 0) ? G[(p + tid) * (n + 1)] : diag0[p + tid];
  if (tid == 0) bad = 0;
  __syncthreads();
  if (role == 0 && k == 0 && ti == tj && tid < bw(ti)) diag0[pti + tid] = T3[tid][tid];
  ci_factor(Dm, kb, refd, piv, sqd, invd, &bad);

  if (role == 0) {
    // R_k,ti and R_k,tj, then the trailing tile
    if (tid < CB) ci_fwd_col(Dm, invd, kb, tid, T1);
    else if (tj != ti && tid >= 64 && tid < 64 + CB) ci_fwd_col(Dm, invd, kb, tid - 64, T2);
    __syncthreads();
    const double(*Pj)[CHP] = (tj != ti) ? T2 : T1;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    ci_mma<true>(T1, Pj, acc);
    const int wi = bw(ti), wj = bw(tj);
    if (a < wi) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (c0 + q < wj) G[(pti + a) + (ptj + c0 + q) * n] = T3[a][c0 + q] - acc[q];
    }
    if (ti == tj) {
      for (int e = tid; e < CB * CB; e += 256) {
        const int r = e & (CB - 1), cc = e >> 5;
        if (r < kb && cc < wi) R[(p + r) + (pti + cc) * n] = T1[r][cc];
      }
    }
    return;
  }

  // roles 1 and 2 need X_kk; role 1 also R_k,tj
  if (tid < CB) ci_inv_col(Dm, invd, kb, tid, Xk);
  else if (role == 1 && tid >= 64 && tid < 64 + CB) ci_fwd_col(Dm, invd, kb, tid - 64, T1);
  __syncthreads();
  // X_ik = -Acc_ik X_kk (rows of block i), or X_kk
  const double(*Xi)[CHP] = Xk;
  if (i < k) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    ci_mma<false>(T2, Xk, acc);
#pragma unroll
    for (int q = 0; q < 4; ++q) T4[a][c0 + q] = -acc[q];
    __syncthreads();
    Xi = T4;
  }
  if (role == 1) {
    double acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = (i < k) ? T3[a][c0 + q] : 0.0;
    ci_mma<false>(Xi, T1, acc);
    const int rows = (i == k) ? kb : CB, wj = bw(tj);
    if (a < rows) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (c0 + q < wj) Acc[(pi + a) + (ptj + c0 + q) * n] = acc[q];
    }
    return;
  }
  // finalize
  const int rows = (i == k) ? kb : CB;
  for (int e = tid; e < CB * CB; e += 256) {
    const int r = e & (CB - 1), cc = e >> 5;
    if (r < rows && cc < kb) X[(pi + r) + (p + cc) * n] = Xi[r][cc];
  }
  if (i == k) {
    for (int e = tid; e < CB * CB; e += 256) {
      const int r = e & (CB - 1), cc = e >> 5;
      if (r < kb && cc < kb) R[(p + r) + (p + cc) * n] = Dm[r][cc];
    }
    const int64_t below = n - p - kb;
    for (int64_t e = tid; e < below * kb; e += 256) {
      const int64_t r = p + kb + e % below, cc = p + e / below;
      R[r + cc * n] = 0.0;
      X[r + cc * n] = 0.0;
    }
    if (tid == 0 && bad) atomicOr(fail, 1);
  }
}

}  // namespace

// R = chol(G) (upper) and Rinv = R^-1, both n x n with ld n (zero below the
// diagonal).  G's trailing part is overwritten; acc (n x n) and diag0 (n) are
// workspace.  Sets *dfail on a pivot that is not safely positive.
static int chol_inv_launch(Context* c, int64_t n, double* G, double* R, double* Rinv,
                           int* dfail, double* diag0, double* acc) {
  const int nb = (int)((n + CB - 1) / CB);
  KTimer kt(c, "chol_inv");
  for (int k = 0; k < nb; ++k) {
    const int T = nb - k - 1;
    const int grid = T * (T + 1) / 2 + (k + 1) * T + (k + 1);
    hipLaunchKernelGGL(cholinv_panel_kernel, dim3(grid), dim3(256), 0, c->stream, n, G, R, Rinv,
                       acc, diag0, k, dfail);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

int chol_inv_upper_f64(Context* c, int64_t n, double* G, double* R, double* Rinv, int* ok) {
  DBuf<int> fail;
  DBuf<double> diag0, acc;
  RSVD_TRY(fail.alloc(1));
  RSVD_TRY(diag0.alloc((size_t)n));
  RSVD_TRY(acc.alloc((size_t)(n * n)));
  RSVD_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), c->stream));
  RSVD_TRY(chol_inv_launch(c, n, G, R, Rinv, fail.p, diag0.p, acc.p));
  int h = 0;
  RSVD_TRY(read_small(c, &h, fail.p, sizeof(int)));
  *ok = h ? 0 : 1;
  return 0;
}

// ---------------------------------------------------------------- CholeskyQR2
// One CholeskyQR pass: G = Y^T Y, R = chol(G), Qout = Y R^-1.
static int cholqr_pass(Context* c, int64_t m, int64_t n, const double* Y, int64_t ldy,
                       double* Qout, int64_t ldq, double* G, double* R, double* Rinv,
                       double* diag0, double* acc, int* dfail) {
  RSVD_TRY(gemm_f64(c, 'T', 'N', n, n, m, 1.0, Y, ldy, Y, ldy, 0.0, G, n));
  RSVD_TRY(chol_inv_launch(c, n, G, R, Rinv, dfail, diag0, acc));
  RSVD_TRY(gemm_f64(c, 'N', 'N', m, n, n, 1.0, Y, ldy, Rinv, n, 0.0, Qout, ldq));
  return 0;
}

static int read_flag(Context* c, int* dflag, int* out) {
  RSVD_HIP(hipMemcpyAsync(out, dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  RSVD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int orth_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* R,
             int* used_fallback) {
  if (used_fallback) *used_fallback = 0;
  if (n == 0 || m == 0) return 0;
  if (m < n) return set_error("orth: m=%lld < n=%lld (thin QR needs m >= n)", (long long)m,
                              (long long)n);
  DBuf<double> W, G, G1, G2, Rinv, diag0, acc;
  DBuf<int> fail;
  RSVD_TRY(acc.alloc((size_t)(n * n)));
  RSVD_TRY(W.alloc((size_t)(m * n)));
  RSVD_TRY(G.alloc((size_t)(n * n)));
  RSVD_TRY(G1.alloc((size_t)(n * n)));
  RSVD_TRY(G2.alloc((size_t)(n * n)));
  RSVD_TRY(Rinv.alloc((size_t)(n * n)));
  RSVD_TRY(diag0.alloc((size_t)n));
  RSVD_TRY(fail.alloc(1));
  RSVD_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), c->stream));
  // pass 1: Q1 = Y R1^-1 into W (Y kept for the fallback)
  RSVD_TRY(cholqr_pass(c, m, n, Y, ldy, W.p, m, G.p, G1.p, Rinv.p, diag0.p, acc.p, fail.p));
  int h = 0;
  RSVD_TRY(read_flag(c, fail.p, &h));
  if (h) {
    if (used_fallback) *used_fallback = 1;
    return householder_qr_f64(c, m, n, Y, ldy, R);
  }
  // pass 2: Q = Q1 R2^-1 into Y
  RSVD_TRY(cholqr_pass(c, m, n, W.p, m, Y, ldy, G.p, G2.p, Rinv.p, diag0.p, acc.p, fail.p));
  RSVD_TRY(read_flag(c, fail.p, &h));
  if (h) {
    // range(Q1) = range(Y): Householder on Q1, R = R_hh * R1
    if (used_fallback) *used_fallback = 1;
    RSVD_TRY(copy2d_f64(c, m, n, W.p, m, Y, ldy));
    RSVD_TRY(householder_qr_f64(c, m, n, Y, ldy, R ? G2.p : nullptr));
    if (R) RSVD_TRY(gemm_f64(c, 'N', 'N', n, n, n, 1.0, G2.p, n, G1.p, n, 0.0, R, n));
    return 0;
  }
  if (R) RSVD_TRY(gemm_f64(c, 'N', 'N', n, n, n, 1.0, G2.p, n, G1.p, n, 0.0, R, n));
  return 0;
}

}  // namespace rsvd
