// qr.hip -- tall-skinny orthonormalisation for the rSVD / QB pipeline,
// replacing QR_factorization_getQ and compact_QR_factorization
// (MKL/mvf.c:1214-1263, dgeqrf + dorgqr).
//
// Main path: CholeskyQR2.  G = Y^T Y on the matrix cores (split-K GEMM), FP64
// blocked Cholesky G = R^T R, blocked triangular inverse, Q1 = Y R^-1 (GEMM),
// and the same once more on Q1; R = R2 R1.  Breakdown (a pivot below
// tol * original diagonal, or non-finite) is reported back to the host and the
// column-by-column Householder QR (geqr2 + org2r, LAPACK sign convention)
// runs instead.
#include <cmath>

#include "rsvd_internal.h"

namespace rsvd {

constexpr int NB = 32;                  // panel width of Cholesky / trtri
constexpr double kCholRelTol = 1e-13;   // pivot / original diagonal

// ---------------------------------------------------------------- Cholesky
// Factor the kb x kb diagonal block at (p,p) of G in LDS (upper, G = R^T R),
// then (every workgroup) solve R_pp^T X = G[p:p+kb, cols] for its 256-column
// chunk of the row panel.  The chunk is staged through LDS with coalesced
// loads and stores (each column of the panel is kb contiguous doubles), and
// the solves multiply by the reciprocal pivots.  Out of place: G is only read
// here, R only written, so the workgroups never race on the diagonal block.
// Workgroup 0 stores it.
__global__ void __launch_bounds__(256)
chol_panel_kernel(int64_t n, const double* __restrict__ G, double* __restrict__ R, int64_t p,
                  int kb, const double* __restrict__ diag0, int* __restrict__ fail) {
  __shared__ double D[NB][NB + 1];
  __shared__ double inv[NB];
  __shared__ double P[256][NB + 1];
  __shared__ int bad;
  const int tid = threadIdx.x;
  for (int e = tid; e < kb * kb; e += 256) {
    int i = e % kb, j = e / kb;
    D[i][j] = (i <= j) ? G[(p + i) + (p + j) * n] : 0.0;
  }
  // stage this workgroup's chunk of the row panel: column cc of the chunk is
  // G[p : p + kb, c0 + cc]
  const int64_t c0 = p + kb + (int64_t)blockIdx.x * 256;
  const int64_t ncols = (n - c0) < 256 ? (n - c0) : 256;
  for (int e = tid; e < NB * 256; e += 256) {
    const int j = e % NB, cc = e / NB;
    P[cc][j] = (j < kb && cc < ncols) ? G[(p + j) + (c0 + cc) * n] : 0.0;
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    if (tid == 0) {
      double d = D[j][j];
      double ref = diag0[p + j];
      if (!(d > kCholRelTol * ref) || !isfinite(d)) {
        bad = 1;
        d = (d > 0 && isfinite(d)) ? d : 1.0;
      }
      const double r = sqrt(d);
      D[j][j] = r;
      inv[j] = 1.0 / r;
    }
    __syncthreads();
    const double ri = inv[j];
    for (int c = j + 1 + tid; c < kb; c += 256) D[j][c] *= ri;
    __syncthreads();
    // trailing update of the upper triangle of D
    for (int e = tid; e < kb * kb; e += 256) {
      int i = e % kb, c = e / kb;
      if (i > j && c >= i) D[i][c] -= D[j][i] * D[j][c];
    }
    __syncthreads();
  }
  if (blockIdx.x == 0) {
    for (int e = tid; e < kb * kb; e += 256) {
      int i = e % kb, j = e / kb;
      R[(p + i) + (p + j) * n] = (i <= j) ? D[i][j] : 0.0;
    }
    if (tid == 0 && bad) atomicOr(fail, 1);
  }
  // forward substitution R_pp^T x = g for column tid of the chunk, in LDS
  if (tid < ncols) {
    double x[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < kb) {
        double s = P[tid][j];
        for (int i = 0; i < j; ++i) s -= D[i][j] * x[i];
        x[j] = s * inv[j];
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < kb) P[tid][j] = x[j];
  }
  __syncthreads();
  for (int e = tid; e < NB * 256; e += 256) {
    const int j = e % NB, cc = e / NB;
    if (j < kb && cc < ncols) R[(p + j) + (c0 + cc) * n] = P[cc][j];
  }
}

__global__ void save_diag_kernel(int64_t n, const double* __restrict__ G, double* __restrict__ d) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) d[i] = G[i + i * n];
}

// R = chol(G) (upper), G's trailing part is destroyed.
static int chol_launch(Context* c, int64_t n, double* G, double* R, int* dfail, double* diag0) {
  RSVD_TRY(fill2d_f64(c, n, n, 0.0, R, n));
  hipLaunchKernelGGL(save_diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     n, (const double*)G, diag0);
  RSVD_LAUNCH_CHECK();
  for (int64_t p = 0; p < n; p += NB) {
    const int kb = (int)((n - p < NB) ? n - p : NB);
    const int64_t rest = n - p - kb;
    const int nblk = (int)((rest + 255) / 256);
    hipLaunchKernelGGL(chol_panel_kernel, dim3(nblk > 0 ? nblk : 1), dim3(256), 0, c->stream, n,
                       (const double*)G, R, p, kb, (const double*)diag0, dfail);
    RSVD_LAUNCH_CHECK();
    if (rest > 0) {
      const double* X = R + p + (p + kb) * n;  // kb x rest, ld n
      double* T = G + (p + kb) + (p + kb) * n; // rest x rest, ld n
      RSVD_TRY(gemm_f64(c, 'T', 'N', rest, rest, kb, -1.0, X, n, X, n, 1.0, T, n));
    }
  }
  return 0;
}

int cholesky_upper_f64(Context* c, int64_t n, double* G, double* R, int* ok) {
  DBuf<int> fail;
  DBuf<double> diag0;
  RSVD_TRY(fail.alloc(1));
  RSVD_TRY(diag0.alloc((size_t)n));
  RSVD_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), c->stream));
  RSVD_TRY(chol_launch(c, n, G, R, fail.p, diag0.p));
  int h = 0;
  RSVD_HIP(hipMemcpyAsync(&h, fail.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  RSVD_HIP(hipStreamSynchronize(c->stream));
  *ok = h ? 0 : 1;
  return 0;
}

// ---------------------------------------------------------------- trtri
// Inverse of each NB x NB diagonal block (one workgroup per block), and zero
// the strictly lower part of Rinv.
__global__ void __launch_bounds__(256)
trtri_diag_kernel(int64_t n, const double* __restrict__ R, double* __restrict__ X) {
  __shared__ double Rs[NB][NB + 1];
  __shared__ double Xs[NB][NB + 1];
  const int64_t p = (int64_t)blockIdx.x * NB;
  const int kb = (int)((n - p < NB) ? n - p : NB);
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += 256) {
    int i = e % NB, j = e / NB;
    Rs[i][j] = (i < kb && j < kb && i <= j) ? R[(p + i) + (p + j) * n] : 0.0;
    Xs[i][j] = 0.0;
  }
  __syncthreads();
  if (tid < kb) {
    const int col = tid;  // back substitution for column `col` of X
    Xs[col][col] = 1.0 / Rs[col][col];
    for (int i = col - 1; i >= 0; --i) {
      double s = 0;
      for (int k = i + 1; k <= col; ++k) s += Rs[i][k] * Xs[k][col];
      Xs[i][col] = -s / Rs[i][i];
    }
  }
  __syncthreads();
  // write the diagonal block and zero everything below the diagonal blocks in
  // this block-column
  for (int e = tid; e < NB * NB; e += 256) {
    int i = e % NB, j = e / NB;
    if (i < kb && j < kb) X[(p + i) + (p + j) * n] = Xs[i][j];
  }
  for (int64_t e = tid; e < (n - p - kb) * kb; e += 256) {
    int64_t i = p + kb + e % (n - p - kb), j = p + e / (n - p - kb);
    X[i + j * n] = 0.0;
  }
}

// Block diagonal d: X_{i,i+d} = -X_ii * sum_{k=i+1}^{i+d} R_ik X_{k,i+d}.
__global__ void __launch_bounds__(256)
trtri_offdiag_kernel(int64_t n, const double* __restrict__ R, double* __restrict__ X, int d) {
  __shared__ double Ta[NB][NB + 1];
  __shared__ double Tb[NB][NB + 1];
  __shared__ double S[NB][NB + 1];
  const int bi = blockIdx.x, bj = bi + d;
  const int64_t pi = (int64_t)bi * NB, pj = (int64_t)bj * NB;
  const int ki = (int)((n - pi < NB) ? n - pi : NB);
  const int kj = (int)((n - pj < NB) ? n - pj : NB);
  const int tid = threadIdx.x;
  const int ti = tid % NB, tj0 = tid / NB;  // each thread: row ti, cols tj0 + 8*r
  double acc[NB / 8];
#pragma unroll
  for (int r = 0; r < NB / 8; ++r) acc[r] = 0.0;
  for (int bk = bi + 1; bk <= bj; ++bk) {
    const int64_t pk = (int64_t)bk * NB;
    const int kk = (int)((n - pk < NB) ? n - pk : NB);
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) {
      int i = e % NB, j = e / NB;
      Ta[i][j] = (i < ki && j < kk) ? R[(pi + i) + (pk + j) * n] : 0.0;
      Tb[i][j] = (i < kk && j < kj) ? X[(pk + i) + (pj + j) * n] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NB / 8; ++r) {
      int tj = tj0 + 8 * r;
      double s = 0;
      for (int k = 0; k < NB; ++k) s += Ta[ti][k] * Tb[k][tj];
      acc[r] += s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NB / 8; ++r) S[ti][tj0 + 8 * r] = acc[r];
  for (int e = tid; e < NB * NB; e += 256) {
    int i = e % NB, j = e / NB;
    Ta[i][j] = (i < ki && j < ki) ? X[(pi + i) + (pi + j) * n] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NB / 8; ++r) {
    int tj = tj0 + 8 * r;
    double s = 0;
    for (int k = 0; k < NB; ++k) s += Ta[ti][k] * S[k][tj];
    if (ti < ki && tj < kj) X[(pi + ti) + (pj + tj) * n] = -s;
  }
}

int trtri_upper_f64(Context* c, int64_t n, const double* R, double* Rinv) {
  const int nblk = (int)((n + NB - 1) / NB);
  hipLaunchKernelGGL(trtri_diag_kernel, dim3(nblk), dim3(256), 0, c->stream, n, R, Rinv);
  RSVD_LAUNCH_CHECK();
  for (int d = 1; d < nblk; ++d) {
    hipLaunchKernelGGL(trtri_offdiag_kernel, dim3(nblk - d), dim3(256), 0, c->stream, n, R, Rinv,
                       d);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

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
  for (int64_t j = n - 1; j >= 0; --j) {
    if (j + 1 < n) {
      hipLaunchKernelGGL(hh_apply_kernel, dim3((unsigned)(n - j - 1)), dim3(256), 0, c->stream, m,
                         Y, ldy, j, j + 1, (const double*)tau.p, (const double*)Y, ldy);
      RSVD_LAUNCH_CHECK();
    }
    int blocks = (int)((m + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(hh_form_col_kernel, dim3(blocks), dim3(256), 0, c->stream, m, Y, ldy, j,
                       (const double*)tau.p);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

// ---------------------------------------------------------------- CholeskyQR2
// One CholeskyQR pass: G = Y^T Y, R = chol(G), Qout = Y R^-1.
static int cholqr_pass(Context* c, int64_t m, int64_t n, const double* Y, int64_t ldy,
                       double* Qout, int64_t ldq, double* G, double* R, double* Rinv,
                       double* diag0, int* dfail) {
  RSVD_TRY(gemm_f64(c, 'T', 'N', n, n, m, 1.0, Y, ldy, Y, ldy, 0.0, G, n));
  RSVD_TRY(chol_launch(c, n, G, R, dfail, diag0));
  RSVD_TRY(trtri_upper_f64(c, n, R, Rinv));
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
  DBuf<double> W, G, G1, G2, Rinv, diag0;
  DBuf<int> fail;
  RSVD_TRY(W.alloc((size_t)(m * n)));
  RSVD_TRY(G.alloc((size_t)(n * n)));
  RSVD_TRY(G1.alloc((size_t)(n * n)));
  RSVD_TRY(G2.alloc((size_t)(n * n)));
  RSVD_TRY(Rinv.alloc((size_t)(n * n)));
  RSVD_TRY(diag0.alloc((size_t)n));
  RSVD_TRY(fail.alloc(1));
  RSVD_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), c->stream));
  // pass 1: Q1 = Y R1^-1 into W (Y kept for the fallback)
  RSVD_TRY(cholqr_pass(c, m, n, Y, ldy, W.p, m, G.p, G1.p, Rinv.p, diag0.p, fail.p));
  int h = 0;
  RSVD_TRY(read_flag(c, fail.p, &h));
  if (h) {
    if (used_fallback) *used_fallback = 1;
    return householder_qr_f64(c, m, n, Y, ldy, R);
  }
  // pass 2: Q = Q1 R2^-1 into Y
  RSVD_TRY(cholqr_pass(c, m, n, W.p, m, Y, ldy, G.p, G2.p, Rinv.p, diag0.p, fail.p));
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
