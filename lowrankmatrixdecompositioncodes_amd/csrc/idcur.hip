// idcur.hip -- interpolative (ID) and CUR decompositions on top of the blocked
// QB, the column-pivoted QR they need and the triangular solve for their
// coefficients.
//
//   pivoted_qr_f64  : pivotedQR_mkl (MKL/rra.c:924-975), LAPACK dgeqp3 semantics
//                     (Businger-Golub column pivoting, partial column norms
//                     downdated and recomputed on cancellation as in dlaqp2)
//   trsm_upper_f64  : upper_triangular_system_solve type 1 (MKL/mvf.c:1477-1494)
//   id_blockrand_f64, id_two_sided_blockrand_f64, cur_blockrand_f64 :
//                     rra.c:1969-2027, 2086-2112, 2262-2332
//
// Every arithmetic step runs on the GPU; the host keeps the control flow and
// the integer index vectors the reference returns.
#include <cmath>
#include <vector>

#include "idcur.h"

namespace rsvd {
namespace {

// sqrt(dlamch('Epsilon')) of dlaqp2: the downdated norm is recomputed when
// the update cancelled to below this fraction of its last exact value
constexpr double kTol3z = 1.0536712127723509e-08;

// vn1 = vn2 = ||A(:, j)||, perm = identity; one wave per column.
__global__ void __launch_bounds__(256)
pqr_init_kernel(int64_t m, int64_t n, const double* __restrict__ A, int64_t lda,
                double* __restrict__ vn1, double* __restrict__ vn2, int* __restrict__ perm) {
  const int lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const double* a = A + j * lda;
  double s = 0.0;
  for (int64_t r = lane; r < m; r += 64) s += a[r] * a[r];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    vn1[j] = sqrt(s);
    vn2[j] = vn1[j];
    perm[j] = (int)j;
  }
}

// Step i, part 1 (one workgroup): pivot = first index of max vn1[i..n)
// (idamax), swap it into column i (vn of the pivot's slot inherits column i's,
// as dlaqp2), then the reflector of A(i:m, i) in dlarfg's convention: beta =
// -sign(alpha) hypot(alpha, ||x||), tau = (beta - alpha) / beta, v = x /
// (alpha - beta), R_ii = beta; tau = 0 for a single row or a zero x.
__global__ void __launch_bounds__(1024)
pqr_pivot_kernel(int64_t m, int64_t n, double* __restrict__ A, int64_t lda, int64_t i,
                 double* __restrict__ vn1, double* __restrict__ vn2, int* __restrict__ perm,
                 double* __restrict__ tau) {
  __shared__ double sv[16];
  __shared__ int64_t si[16];
  __shared__ int64_t s_pvt;
  __shared__ double s_scale;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double best = -1.0;
  int64_t bidx = n;
  for (int64_t j = i + tid; j < n; j += 1024) {
    const double v = vn1[j];
    if (v > best) { best = v; bidx = j; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(best, off, 64);
    const int64_t oi = __shfl_xor(bidx, off, 64);
    if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
  }
  if (lane == 0) { sv[w] = best; si[w] = bidx; }
  __syncthreads();
  if (tid == 0) {
    double b = sv[0];
    int64_t bi = si[0];
    for (int k = 1; k < 16; ++k)
      if (sv[k] > b || (sv[k] == b && si[k] < bi)) { b = sv[k]; bi = si[k]; }
    s_pvt = (bi < n) ? bi : i;
  }
  __syncthreads();
  const int64_t pvt = s_pvt;
  if (pvt != i) {
    double* ci = A + i * lda;
    double* cp = A + pvt * lda;
    for (int64_t r = tid; r < m; r += 1024) {
      const double t = ci[r];
      ci[r] = cp[r];
      cp[r] = t;
    }
    if (tid == 0) {
      const int t = perm[pvt];
      perm[pvt] = perm[i];
      perm[i] = t;
      vn1[pvt] = vn1[i];
      vn2[pvt] = vn2[i];
    }
  }
  __syncthreads();
  double* x = A + i + i * lda;
  const int64_t len = m - i;
  double s = 0.0;
  for (int64_t r = 1 + tid; r < len; r += 1024) s += x[r] * x[r];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) sv[w] = s;
  __syncthreads();
  if (tid == 0) {
    double ss = 0.0;
    for (int k = 0; k < 16; ++k) ss += sv[k];
    const double alpha = x[0];
    if (len <= 1 || ss == 0.0) {
      tau[i] = 0.0;
      s_scale = 0.0;
    } else {
      const double xn = sqrt(ss);
      const double a = fabs(alpha), big = fmax(a, xn), sml = fmin(a, xn);
      const double h = (sml == 0.0) ? big : big * sqrt(1.0 + (sml / big) * (sml / big));
      const double beta = (alpha >= 0.0) ? -h : h;
      tau[i] = (beta - alpha) / beta;
      s_scale = 1.0 / (alpha - beta);
      x[0] = beta;
    }
  }
  __syncthreads();
  const double sc = s_scale;
  if (sc != 0.0)
    for (int64_t r = 1 + tid; r < len; r += 1024) x[r] *= sc;
}

// Step i, part 2: columns c > i get H_i applied (rows i..m-1) and their
// partial norms downdated (dlaqp2's rule, recomputed from the column when the
// downdate cancelled); one wave per column.
__global__ void __launch_bounds__(256)
pqr_update_kernel(int64_t m, int64_t n, double* __restrict__ A, int64_t lda, int64_t i,
                  double* __restrict__ vn1, double* __restrict__ vn2,
                  const double* __restrict__ tau) {
  const int lane = threadIdx.x & 63;
  const int64_t c = i + 1 + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= n) return;
  const double t = tau[i];
  const double* v = A + i * lda;  // v_i = 1, v_r = A(r, i) below
  double* y = A + c * lda;
  double yi = y[i];
  if (t != 0.0) {
    double s = 0.0;
    for (int64_t r = i + lane; r < m; r += 64) s += (r == i ? 1.0 : v[r]) * y[r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    const double wv = t * s;
    for (int64_t r = i + lane; r < m; r += 64) y[r] -= (r == i ? 1.0 : v[r]) * wv;
    yi -= wv;
  }
  const double n1 = vn1[c];
  if (n1 != 0.0) {
    double temp = fabs(yi) / n1;
    temp = fmax(1.0 - temp * temp, 0.0);
    const double r2 = n1 / vn2[c];
    if (temp * r2 * r2 <= kTol3z) {
      double s = 0.0;
      for (int64_t r = i + 1 + lane; r < m; r += 64) s += y[r] * y[r];
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) {
        const double nv = (i + 1 < m) ? sqrt(s) : 0.0;
        vn1[c] = nv;
        vn2[c] = nv;
      }
    } else if (lane == 0) {
      vn1[c] = n1 * sqrt(temp);
    }
  }
}

// Inverse of each 64 x 64 diagonal block of an upper triangular R (k x k):
// thread j solves R_bb x = e_j by back substitution.
__global__ void __launch_bounds__(64)
trsm_diag_inv_kernel(int64_t k, const double* __restrict__ R, int64_t ldr,
                     double* __restrict__ Dinv) {
  __shared__ double Rb[64][65];
  __shared__ double X[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int bs = (int)((k - r0 < 64) ? k - r0 : 64);
  const int j = threadIdx.x;
  for (int col = 0; col < 64; ++col)
    Rb[j][col] = (j < bs && col < bs) ? R[(r0 + j) + (r0 + col) * ldr] : 0.0;
  __syncthreads();
  for (int ii = 63; ii >= 0; --ii) {
    double x = 0.0;
    if (j < bs && ii < bs && ii <= j) {
      double s = (ii == j) ? 1.0 : 0.0;
      for (int kk = ii + 1; kk <= j; ++kk) s -= Rb[ii][kk] * X[kk][j];
      x = s / Rb[ii][ii];
    }
    X[ii][j] = x;
  }
  __syncthreads();
  double* D = Dinv + (size_t)blockIdx.x * 64 * 64;
  for (int ii = 0; ii < 64; ++ii) D[ii + j * 64] = X[ii][j];
}

// out(j, i) = A(i, idx[j]) for j < cnt (the transpose of the selected columns).
__global__ void gather_cols_t_kernel(int64_t m, int64_t cnt, const double* __restrict__ A,
                                     int64_t lda, const int* __restrict__ idx,
                                     double* __restrict__ out, int64_t ldo) {
  const int64_t total = m * cnt;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;  // consecutive threads read a column
    out[j + i * ldo] = A[i + (int64_t)idx[j] * lda];
  }
}

// out(:, j) = A(:, idx[j]) for j < cnt.
__global__ void gather_cols_kernel(int64_t m, int64_t cnt, const double* __restrict__ A,
                                   int64_t lda, const int* __restrict__ idx,
                                   double* __restrict__ out, int64_t ldo) {
  const int64_t total = m * cnt;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;
    out[i + j * ldo] = A[i + (int64_t)idx[j] * lda];
  }
}

// out(r, :) = A(idx[r], :) for r < cnt (cnt x n, ld ldo).
__global__ void gather_rows_kernel(int64_t cnt, int64_t n, const double* __restrict__ A,
                                   int64_t lda, const int* __restrict__ idx,
                                   double* __restrict__ out, int64_t ldo) {
  const int64_t total = cnt * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / cnt, r = e - j * cnt;
    out[r + j * ldo] = A[(int64_t)idx[r] + j * lda];
  }
}

// V = [I_k; T^T](inv, :) (n x k): row i is e_{inv[i]} when inv[i] < k, else
// row inv[i] - k of T^T, i.e. column inv[i] - k of T (k x (n - k), ld ldt).
__global__ void id_interp_matrix_kernel(int64_t n, int64_t k, const int* __restrict__ inv,
                                        const double* __restrict__ T, int64_t ldt,
                                        double* __restrict__ V, int64_t ldv) {
  const int64_t total = n * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / n, i = e - j * n;
    const int64_t r = inv[i];
    V[i + j * ldv] = (r < k) ? (r == j ? 1.0 : 0.0) : T[j + (r - k) * ldt];
  }
}

// A(i, j) = 0 for i > j (rows x cols, ld lda).
__global__ void zero_lower_kernel(int64_t rows, int64_t cols, double* __restrict__ A,
                                  int64_t lda) {
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / rows, i = e - j * rows;
    if (i > j) A[i + j * lda] = 0.0;
  }
}

int grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b < 1) b = 1;
  return (int)(b > 8192 ? 8192 : b);
}

}  // namespace

int pivoted_qr_f64(Context* c, int64_t m, int64_t n, double* A, int64_t lda, int* perm,
                   double* tau) {
  if (m <= 0 || n <= 0) return 0;
  if (lda < m) return set_error("pivoted_qr: lda %lld < m %lld", (long long)lda, (long long)m);
  DBuf<double> vn1, vn2;
  RSVD_TRY(vn1.alloc((size_t)n));
  RSVD_TRY(vn2.alloc((size_t)n));
  KTimer kt(c, "pivoted_qr");
  hipLaunchKernelGGL(pqr_init_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c->stream, m,
                     n, (const double*)A, lda, vn1.p, vn2.p, perm);
  RSVD_LAUNCH_CHECK();
  const int64_t kmax = m < n ? m : n;
  for (int64_t i = 0; i < kmax; ++i) {
    hipLaunchKernelGGL(pqr_pivot_kernel, dim3(1), dim3(1024), 0, c->stream, m, n, A, lda, i,
                       vn1.p, vn2.p, perm, tau);
    RSVD_LAUNCH_CHECK();
    const int64_t rest = n - i - 1;
    if (rest > 0) {
      hipLaunchKernelGGL(pqr_update_kernel, dim3((unsigned)((rest + 3) / 4)), dim3(256), 0,
                         c->stream, m, n, A, lda, i, vn1.p, vn2.p, (const double*)tau);
      RSVD_LAUNCH_CHECK();
    }
  }
  return 0;
}

int trsm_upper_f64(Context* c, int64_t k, int64_t nrhs, const double* R, int64_t ldr, double* B,
                   int64_t ldb) {
  if (k <= 0 || nrhs <= 0) return 0;
  constexpr int64_t NB = 64;
  const int64_t nb = (k + NB - 1) / NB;
  DBuf<double> Dinv, tmp;
  RSVD_TRY(Dinv.alloc((size_t)(nb * NB * NB)));
  RSVD_TRY(tmp.alloc((size_t)(NB * nrhs)));
  KTimer kt(c, "trsm");
  hipLaunchKernelGGL(trsm_diag_inv_kernel, dim3((unsigned)nb), dim3(64), 0, c->stream, k, R, ldr,
                     Dinv.p);
  RSVD_LAUNCH_CHECK();
  for (int64_t b = nb - 1; b >= 0; --b) {
    const int64_t r0 = b * NB, bs = (k - r0 < NB) ? k - r0 : NB, r1 = r0 + bs;
    if (r1 < k)  // B_b -= R(b, later) X(later)
      RSVD_TRY(gemm_f64(c, 'N', 'N', bs, nrhs, k - r1, -1.0, R + r0 + r1 * ldr, ldr, B + r1, ldb,
                        1.0, B + r0, ldb));
    RSVD_TRY(gemm_f64(c, 'N', 'N', bs, nrhs, bs, 1.0, Dinv.p + b * NB * NB, NB, B + r0, ldb, 0.0,
                      tmp.p, bs));
    RSVD_TRY(copy2d_f64(c, bs, nrhs, tmp.p, bs, B + r0, ldb));
  }
  return 0;
}

// The coefficients of an ID from R (the upper part of a pivoted QR, ld ldr,
// n columns): T = R(0:f, 0:f)^-1 R(0:f, f:n), f x (n - f).
static int id_coefficients(Context* c, const double* R, int64_t ldr, int64_t n, int64_t f,
                           DBuf<double>& T) {
  RSVD_TRY(T.alloc((size_t)(f * (n - f > 0 ? n - f : 1))));
  if (n - f > 0) {
    RSVD_TRY(copy2d_f64(c, f, n - f, R + f * ldr, ldr, T.p, f));
    RSVD_TRY(trsm_upper_f64(c, f, n - f, R, ldr, T.p, f));
  }
  return 0;
}

int id_blockrand_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                     double TOL, int kstep, int q, int s, const OmegaSrc& om, IdOut* out) {
  if (kstep <= 0) return set_error("id_blockrand: kstep must be positive (kstep=%d)", kstep);
  int nstep = (k + p) / kstep;                                          // rra.c:1974
  const int rankMode = k > 0;
  if (!rankMode) nstep = 0;                                             // rra.c:1980-1983
  int ks = kstep, ns = nstep, tm = 0;
  RSVD_TRY(qb_adjust(m, n, &ks, &ns, 1, &tm));
  const int64_t lmax = (int64_t)ks * ns;
  DBuf<double> Q, B;
  RSVD_TRY(Q.alloc((size_t)(m * lmax)));
  RSVD_TRY(B.alloc((size_t)(lmax * n)));
  int frank = 0, lalloc = 0;
  QbInfo qi{};
  RSVD_TRY(qb_blocked_f64(c, A, m, n, lda, kstep, nstep, TOL, q, s, 1, om, Q.p, m, B.p, lmax,
                          &frank, &lalloc, &qi));                      // rra.c:1988
  Q.reset();
  const int64_t l = qi.tol_mode ? frank : lalloc;                       // rows of B
  DBuf<int> perm;
  DBuf<double> tau;
  RSVD_TRY(perm.alloc((size_t)n));
  RSVD_TRY(tau.alloc((size_t)(l < n ? l : n)));
  RSVD_TRY(pivoted_qr_f64(c, l, n, B.p, lmax, perm.p, tau.p));          // rra.c:1997
  if (rankMode) frank = k;                                              // rra.c:1998-2003
  else frank = (int)llround((double)(frank / (frank + p + 1e-6)) * frank);
  const int64_t rows = l < n ? l : n;
  if (frank <= 0 || frank > rows)
    return set_error("id_blockrand: rank %d outside [1, %lld] (k=%d p=%d kstep=%d)", frank,
                     (long long)rows, k, p, kstep);
  out->frank = frank;
  out->n = n;
  RSVD_TRY(id_coefficients(c, B.p, lmax, n, frank, out->T));           // rra.c:2005-2020
  out->I.resize((size_t)n);
  RSVD_HIP(hipMemcpyAsync(out->I.data(), perm.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost,
                          c->stream));
  RSVD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int id_two_sided_blockrand_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda,
                               int k, int p, double TOL, int kstep, int q, int s,
                               const OmegaSrc& om, IdOut* col, IdOut* row) {
  RSVD_TRY(id_blockrand_f64(c, A, m, n, lda, k, p, TOL, kstep, q, s, om, col));  // rra.c:2095
  const int64_t f = col->frank;
  // MI^T = A(:, Icol(1:f))^T (f x m), then its full ID (rra.c:2100-2108)
  DBuf<int> icol;
  RSVD_TRY(icol.alloc((size_t)f));
  RSVD_HIP(hipMemcpyAsync(icol.p, col->I.data(), sizeof(int) * (size_t)f, hipMemcpyHostToDevice,
                          c->stream));
  DBuf<double> MIt;
  RSVD_TRY(MIt.alloc((size_t)(f * m)));
  hipLaunchKernelGGL(gather_cols_t_kernel, dim3(grid_for(f * m)), dim3(256), 0, c->stream, m, f,
                     A, lda, (const int*)icol.p, MIt.p, f);
  RSVD_LAUNCH_CHECK();
  RSVD_HIP(hipStreamSynchronize(c->stream));  // icol's host source is not pinned
  DBuf<int> perm;
  DBuf<double> tau;
  RSVD_TRY(perm.alloc((size_t)m));
  RSVD_TRY(tau.alloc((size_t)(f < m ? f : m)));
  RSVD_TRY(pivoted_qr_f64(c, f, m, MIt.p, f, perm.p, tau.p));
  if (f > (f < m ? f : m)) return set_error("id_two_sided: rank %lld above m", (long long)f);
  row->frank = (int)f;
  row->n = m;
  RSVD_TRY(id_coefficients(c, MIt.p, f, m, f, row->T));
  row->I.resize((size_t)m);
  RSVD_HIP(hipMemcpyAsync(row->I.data(), perm.p, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost,
                          c->stream));
  RSVD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int cur_blockrand_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                      double TOL, int kstep, int q, int s, const OmegaSrc& om, CurOut* out) {
  IdOut col, row;
  RSVD_TRY(id_two_sided_blockrand_f64(c, A, m, n, lda, k, p, TOL, kstep, q, s, om, &col, &row));
  const int64_t f = col.frank;
  out->frank = (int)f;
  std::vector<int> inv((size_t)n);
  for (int64_t i = 0; i < n; ++i) inv[(size_t)col.I[(size_t)i]] = (int)i;  // mvf.c:1195-1201
  DBuf<int> dinv, icol, irow;
  RSVD_TRY(dinv.alloc((size_t)n));
  RSVD_TRY(icol.alloc((size_t)f));
  RSVD_TRY(irow.alloc((size_t)f));
  RSVD_HIP(hipMemcpyAsync(dinv.p, inv.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice,
                          c->stream));
  RSVD_HIP(hipMemcpyAsync(icol.p, col.I.data(), sizeof(int) * (size_t)f, hipMemcpyHostToDevice,
                          c->stream));
  RSVD_HIP(hipMemcpyAsync(irow.p, row.I.data(), sizeof(int) * (size_t)f, hipMemcpyHostToDevice,
                          c->stream));
  // V = [I_k; T^T](Icol^-1, :), n x f (rra.c:2277-2292)
  DBuf<double> V;
  RSVD_TRY(V.alloc((size_t)(n * f)));
  hipLaunchKernelGGL(id_interp_matrix_kernel, dim3(grid_for(n * f)), dim3(256), 0, c->stream, n,
                     f, (const int*)dinv.p, (const double*)col.T.p, f, V.p, n);
  RSVD_LAUNCH_CHECK();
  // R = A(Irow(1:f), :), C = A(:, Icol(1:f)) (rra.c:2304-2311)
  RSVD_TRY(out->R.alloc((size_t)(f * n)));
  RSVD_TRY(out->C.alloc((size_t)(m * f)));
  hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for(f * n)), dim3(256), 0, c->stream, f, n, A,
                     lda, (const int*)irow.p, out->R.p, f);
  RSVD_LAUNCH_CHECK();
  hipLaunchKernelGGL(gather_cols_kernel, dim3(grid_for(m * f)), dim3(256), 0, c->stream, m, f, A,
                     lda, (const int*)icol.p, out->C.p, m);
  RSVD_LAUNCH_CHECK();
  // (R R^T) U^T = R V (rra.c:2316-2323), R R^T symmetric positive definite:
  // Cholesky R R^T = L^T L with its inverse, U^T = L^-1 L^-T (R V)
  DBuf<double> G, L, Linv, RV, W;
  RSVD_TRY(G.alloc((size_t)(f * f)));
  RSVD_TRY(L.alloc((size_t)(f * f)));
  RSVD_TRY(Linv.alloc((size_t)(f * f)));
  RSVD_TRY(RV.alloc((size_t)(f * f)));
  RSVD_TRY(W.alloc((size_t)(f * f)));
  RSVD_TRY(gemm_f64(c, 'N', 'T', f, f, n, 1.0, out->R.p, f, out->R.p, f, 0.0, G.p, f));
  RSVD_TRY(gemm_f64(c, 'N', 'N', f, f, n, 1.0, out->R.p, f, V.p, n, 0.0, RV.p, f));
  int ok = 0;
  RSVD_TRY(chol_inv_upper_f64(c, f, G.p, L.p, Linv.p, &ok));
  if (!ok) return set_error("cur: R R^T is numerically singular (rank %lld)", (long long)f);
  RSVD_TRY(gemm_f64(c, 'T', 'N', f, f, f, 1.0, Linv.p, f, RV.p, f, 0.0, W.p, f));
  RSVD_TRY(gemm_f64(c, 'N', 'N', f, f, f, 1.0, Linv.p, f, W.p, f, 0.0, RV.p, f));  // U^T
  RSVD_TRY(out->U.alloc((size_t)(f * f)));
  RSVD_TRY(transpose_f64(c, f, f, RV.p, f, out->U.p, f));
  RSVD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int zero_lower_f64(Context* c, int64_t rows, int64_t cols, double* A, int64_t lda) {
  if (rows <= 0 || cols <= 0) return 0;
  hipLaunchKernelGGL(zero_lower_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, c->stream,
                     rows, cols, A, lda);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int gather_cols_f64(Context* c, int64_t m, int64_t cnt, const double* A, int64_t lda,
                    const int* idx, double* out, int64_t ldo) {
  if (m <= 0 || cnt <= 0) return 0;
  hipLaunchKernelGGL(gather_cols_kernel, dim3(grid_for(m * cnt)), dim3(256), 0, c->stream, m, cnt,
                     A, lda, idx, out, ldo);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int gather_rows_f64(Context* c, int64_t cnt, int64_t n, const double* A, int64_t lda,
                    const int* idx, double* out, int64_t ldo) {
  if (cnt <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for(cnt * n)), dim3(256), 0, c->stream, cnt, n,
                     A, lda, idx, out, ldo);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int id_interp_matrix_f64(Context* c, int64_t n, int64_t k, const int* inv, const double* T,
                         int64_t ldt, double* V, int64_t ldv) {
  if (n <= 0 || k <= 0) return 0;
  hipLaunchKernelGGL(id_interp_matrix_kernel, dim3(grid_for(n * k)), dim3(256), 0, c->stream, n,
                     k, inv, T, ldt, V, ldv);
  RSVD_LAUNCH_CHECK();
  return 0;
}

}  // namespace rsvd
