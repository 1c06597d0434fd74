// qr.hip -- tall-skinny orthonormalisation for the rSVD / QB pipeline,
// replacing QR_factorization_getQ and compact_QR_factorization
// (MKL/mvf.c:1214-1263, dgeqrf + dorgqr).
//
// Main path: CholeskyQR2.  G = Y^T Y on the matrix cores (split-K GEMM), FP64
// blocked Cholesky G = R^T R fused with the triangular inverse (one launch per
// 32-row panel), Q1 = Y R^-1 (GEMM), and the same once more on Q1; R = R2 R1.
// Breakdown (a pivot below tol * original diagonal, or non-finite) is
// reported back to the host and the blocked Householder QR (geqrf + orgqr,
// LAPACK sign convention, hhqr.hip) runs instead.
#include <cmath>

#include "rsvd_internal.h"

namespace rsvd {

constexpr double kCholRelTol = 1e-13;   // pivot / original diagonal

// ---------------------------------------------------------------- Householder
// (blocked geqrf / orgqr: hhqr.hip)
__global__ void copy_upper_kernel(int64_t n, const double* __restrict__ Y, int64_t ldy,
                                  double* __restrict__ R) {
  int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % n, j = idx / n;
    R[idx] = (i <= j) ? Y[i + j * ldy] : 0.0;
  }
}

int householder_qr_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* R,
                       double* vwork) {
  if (n == 0 || m == 0) return 0;
  if (m < n) return set_error("householder_qr: m=%lld < n=%lld", (long long)m, (long long)n);
  DBuf<double> tau, Tall;
  RSVD_TRY(tau.alloc((size_t)n));
  RSVD_TRY(Tall.alloc((size_t)((n + 31) / 32) * 32 * 32));
  RSVD_TRY(geqrf_f64(c, m, n, Y, ldy, tau.p, vwork, Tall.p));
  if (R) {
    int blocks = (int)((n * n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(copy_upper_kernel, dim3(blocks), dim3(256), 0, c->stream, n,
                       (const double*)Y, ldy, R);
    RSVD_LAUNCH_CHECK();
  }
  return orgqr_f64(c, m, n, Y, ldy, tau.p, vwork, Tall.p);
}

int org2r_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, const double* tau) {
  return orgqr_f64(c, m, n, Y, ldy, tau, nullptr);
}

// ---------------------------------------------------------------- Cholesky + inverse
// Right-looking blocked Cholesky G = R^T R fused with X = R^-1, one launch
// per 32-row panel k (p = 32 k).  The serial chain is the panel count; every
// workgroup does O(1) tile work, so no launch has a loop over earlier panels.
// Each workgroup first factors the panel's diagonal block D = G_kk (already
// updated by panels < k) in LDS and inverts it, X_kk = R_kk^-1 --
// redundantly, which is cheaper than another launch -- then takes one role:
//   trailing (ti <= tj > k): R_k,t = X_kk^T G_k,t for t = ti, tj and
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

// A 32 x 32 tile of M at (r0, c0) (rows x cols valid, zero beyond) in two
// halves: ci_fetch issues the thread's four loads (entries tid + 256 q; the
// padding reads an in-bounds element and drops it), ci_put stores them to
// LDS, so that every tile of a launch is fetched before the first wait.
__device__ __forceinline__ void ci_fetch(const double* __restrict__ M, int64_t ld, int64_t r0,
                                         int64_t c0, int rows, int cols, double (&v)[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + 256 * q, i = e & (CB - 1), j = e >> 5;
    const bool in = i < rows && j < cols;
    const double x = M[(r0 + (in ? i : 0)) + (c0 + (in ? j : 0)) * ld];
    v[q] = in ? x : 0.0;
  }
}
__device__ __forceinline__ void ci_put(const double (&v)[4], double (*T)[CHP]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + 256 * q;
    T[e & (CB - 1)][e >> 5] = v[q];
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

// Unnormalised right-looking Cholesky of Dm (upper triangle valid, kb x kb),
// four pivots per barrier.  A step takes the pivot rows P = j .. j + 3: every
// thread eliminates the 4 x 4 pivot block itself (the same instructions on
// the same LDS entries in every thread, so all copies agree) and the pivot
// rows at the columns it needs (its row a and its four columns), then
// subtracts sum_{w in P} U_wa U_wc / U_ww from its entries of the rows below
// P.  A pivot that is not above kCholRelTol * ref (or is not finite) sets
// *bad and is replaced (by 1 when it is not positive) so the factorisation
// runs to the end.  Rows stay in registers; the rows of the next step are
// stored before the barrier that ends a step, all rows at the end.  On exit Dm
// holds R_kk (zero below the diagonal and past kb) and sqd / invd its diagonal
// and reciprocals.
__device__ void ci_factor(double (*Dm)[CHP], int kb, const double* ref, double* piv,
                          double* sqd, double* invd, int* bad) {
  constexpr int PB = 4;
  const int tid = threadIdx.x, a = tid & (CB - 1), c0 = (tid >> 5) * 4;
  double own[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) own[q] = Dm[a][c0 + q];
  // All eight steps run on full 4 x 4 blocks: the rows and columns past kb
  // are zero, their pivots are taken as 1 and change nothing, so the step
  // body has no bounds on the block.
  for (int j = 0; j < CB; j += PB) {
    // S of the pivot block (upper part) and of the pivot rows at column a
    // (x = 0) and at columns c0 .. c0 + 3 (x = 1 .. 4)
    double U[PB][PB], X[PB][5], rinv[PB];
#pragma unroll
    for (int u = 0; u < PB; ++u) {
#pragma unroll
      for (int v = u; v < PB; ++v) U[u][v] = Dm[j + u][j + v];
      X[u][0] = Dm[j + u][a];
#pragma unroll
      for (int q = 0; q < 4; ++q) X[u][1 + q] = Dm[j + u][c0 + q];
    }
    // eliminate inside the pivot block, pivot by pivot
#pragma unroll
    for (int u = 0; u < PB; ++u) {
#pragma unroll
      for (int w = 0; w < u; ++w) {
        const double l = U[w][u] * rinv[w];
#pragma unroll
        for (int v = u; v < PB; ++v) U[u][v] -= l * U[w][v];
#pragma unroll
        for (int x = 0; x < 5; ++x) X[u][x] -= l * X[w][x];
      }
      double d = U[u][u];
      const bool pad = j + u >= kb;
      const bool ok = pad | ((d > kCholRelTol * ref[pad ? 0 : j + u]) & isfinite(d));
      d = pad ? 1.0 : (ok ? d : ((d > 0.0 && isfinite(d)) ? d : 1.0));
      rinv[u] = ci_rcp(d);
      if (tid == 0) {
        if (!ok) *bad = 1;
        piv[j + u] = d;
      }
    }
    if (a >= j + PB && a < kb) {
      // rows below P: the Schur complement update
#pragma unroll
      for (int u = 0; u < PB; ++u) {
        const double l = X[u][0] * rinv[u];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (c0 + q >= a && c0 + q < kb) own[q] -= l * X[u][1 + q];
      }
      if (a < j + 2 * PB) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (c0 + q >= a && c0 + q < kb) Dm[a][c0 + q] = own[q];
      }
    } else if (a >= j && a < j + PB) {
      // a pivot row: its eliminated values are final
#pragma unroll
      for (int u = 0; u < PB; ++u)
        if (a == j + u) {
#pragma unroll
          for (int q = 0; q < 4; ++q) own[q] = X[u][1 + q];
        }
    }
    __syncthreads();
  }
  if (a < kb) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c0 + q >= a && c0 + q < kb) Dm[a][c0 + q] = own[q];
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

// Xk = R_kk^-1 (upper, zero past kb) by blocked inversion on all threads:
// the four 8 x 8 diagonal blocks by back substitution (a thread per column),
// then X_12 = -X_11 R_12 X_22 for the two 16 x 16 diagonal blocks and for the
// whole tile, each as two small products with a barrier between.  Tw is
// scratch.  Ends with a barrier.
__device__ void ci_tri_inv(const double (*Rk)[CHP], const double* invd, int kb,
                           double (*Xk)[CHP], double (*Tw)[CHP]) {
  const int tid = threadIdx.x;
  for (int e = tid; e < CB * CB; e += 256) Xk[e & (CB - 1)][e >> 5] = 0.0;
  __syncthreads();
  if (tid < CB) {
    const int c = tid, b0 = c & ~7;
    double x[8];
#pragma unroll
    for (int r = 7; r >= 0; --r) {
      double s = 0.0;
#pragma unroll
      for (int u = 7; u > r; --u) s = fma(Rk[b0 + r][b0 + u], x[u], s);
      const double ir = (b0 + r < kb) ? invd[b0 + r] : 0.0;
      x[r] = (b0 + r > c) ? 0.0 : ((b0 + r == c) ? ir : -s * ir);
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) Xk[b0 + r][c] = (c < kb) ? x[r] : 0.0;
  }
  __syncthreads();
  // 8 -> 16: blocks at b0 = 0 and 16, sub-blocks of 8
  if (tid < 128) {
    const int b0 = (tid >> 6) * 16, r = (tid >> 3) & 7, c = tid & 7;
    double s = 0.0;  // T = R_12 X_22
#pragma unroll
    for (int u = 0; u < 8; ++u) s = fma(Rk[b0 + r][b0 + 8 + u], Xk[b0 + 8 + u][b0 + 8 + c], s);
    Tw[b0 + r][c] = s;
  }
  __syncthreads();
  if (tid < 128) {
    const int b0 = (tid >> 6) * 16, r = (tid >> 3) & 7, c = tid & 7;
    double s = 0.0;  // X_12 = -X_11 T
#pragma unroll
    for (int u = 0; u < 8; ++u) s = fma(Xk[b0 + r][b0 + u], Tw[b0 + u][c], s);
    Xk[b0 + r][b0 + 8 + c] = -s;
  }
  __syncthreads();
  // 16 -> 32
  {
    const int r = tid >> 4, c = tid & 15;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) s = fma(Rk[r][16 + u], Xk[16 + u][16 + c], s);
    Tw[r][c] = s;
  }
  __syncthreads();
  {
    const int r = tid >> 4, c = tid & 15;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) s = fma(Xk[r][u], Tw[u][c], s);
    Xk[r][16 + c] = -s;
  }
  __syncthreads();
}

typedef double ci_d4 __attribute__((ext_vector_type(4)));

// acc += the wave's 16 x 16 quadrant (wave / 2, wave % 2) of op(A) B over the
// padded 32 x 32 tiles (op(A) = A^T when TA), on the FP64 matrix cores: eight
// v_mfma_f64_16x16x4f64 steps (srcA lane l: op(A)[l % 16][k = l / 16], srcB
// lane l: B[k = l / 16][l % 16]).  Register r of lane l is the entry at
// ci_row(r), ci_col() of the tile.
template <bool TA>
__device__ __forceinline__ ci_d4 ci_mma(const double (*A)[CHP], const double (*B)[CHP],
                                        ci_d4 acc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4, ti = wave >> 1, tj = wave & 1;
#pragma unroll
  for (int s = 0; s < CB / 4; ++s) {
    const int k = 4 * s + lk;
    const double av = TA ? A[k][16 * ti + li] : A[16 * ti + li][k];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, B[k][16 * tj + li], acc, 0, 0, 0);
  }
  return acc;
}
__device__ __forceinline__ int ci_row(int r) {
  return 16 * (threadIdx.x >> 7) + 4 * r + ((threadIdx.x & 63) >> 4);
}
__device__ __forceinline__ int ci_col() {
  return 16 * ((threadIdx.x >> 6) & 1) + (threadIdx.x & 15);
}

__global__ void __launch_bounds__(256)
cholinv_panel_kernel(int64_t n, double* __restrict__ G, double* __restrict__ R,
                     double* __restrict__ X, double* __restrict__ Acc,
                     double* __restrict__ diag0, int k, int* __restrict__ fail) {
  __shared__ double Dm[CB][CHP], Xk[CB][CHP], T1[CB][CHP], T2[CB][CHP], T3[CB][CHP],
      T4[CB][CHP];
  __shared__ double refd[CB], piv[CB], sqd[CB], invd[CB];
  __shared__ int bad;
  const int tid = threadIdx.x;
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

  // ---- loads: every tile of the workgroup fetched before the first wait
  // (one memory latency ahead of the serial factorisation)
  {
    double vd[4], v1[4], v2[4], v3[4];
    const double rv = (tid < kb) ? ((k == 0) ? G[(p + tid) * (n + 1)] : diag0[p + tid]) : 0.0;
    ci_fetch(G, n, p, p, kb, kb, vd);
    if (role == 0) {
      ci_fetch(G, n, p, pti, kb, bw(ti), v1);
      if (tj != ti) ci_fetch(G, n, p, ptj, kb, bw(tj), v2);
      ci_fetch(G, n, pti, ptj, bw(ti), bw(tj), v3);
      ci_put(v1, T1);
      if (tj != ti) ci_put(v2, T2);
      ci_put(v3, T3);
    } else if (role == 1) {
      ci_fetch(G, n, p, ptj, kb, bw(tj), v1);
      if (i < k) {
        ci_fetch(Acc, n, pi, p, CB, kb, v2);
        ci_fetch(Acc, n, pi, ptj, CB, bw(tj), v3);
      }
      ci_put(v1, T1);
      if (i < k) {
        ci_put(v2, T2);
        ci_put(v3, T3);
      }
    } else if (i < k) {
      ci_fetch(Acc, n, pi, p, CB, kb, v2);
      ci_put(v2, T2);
    }
    ci_put(vd, Dm);
    if (tid < kb) refd[tid] = rv;
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  if (role == 0 && k == 0 && ti == tj && tid < bw(ti)) diag0[pti + tid] = T3[tid][tid];
  ci_factor(Dm, kb, refd, piv, sqd, invd, &bad);

  ci_tri_inv(Dm, invd, kb, Xk, T4);

  const ci_d4 z4 = {0.0, 0.0, 0.0, 0.0};
  const int cl = ci_col();
  if (role == 0) {
    // R_k,ti = X_kk^T G_k,ti into T4 and R_k,tj into Dm (R_kk is not needed
    // here any more), then the trailing tile
    const ci_d4 pa = ci_mma<true>(Xk, T1, z4);
    const ci_d4 pb = (tj != ti) ? ci_mma<true>(Xk, T2, z4) : z4;
    // Dm's last reads (inside ci_tri_inv) were before its closing barrier
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T4[ci_row(r)][cl] = pa[r];
      if (tj != ti) Dm[ci_row(r)][cl] = pb[r];
    }
    __syncthreads();
    const double(*Pj)[CHP] = (tj != ti) ? Dm : T4;
    const ci_d4 acc = ci_mma<true>(T4, Pj, z4);
    const int wi = bw(ti), wj = bw(tj);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rw = ci_row(r);
      if (rw < wi && cl < wj) G[(pti + rw) + (ptj + cl) * n] = T3[rw][cl] - acc[r];
    }
    if (ti == tj) {
      for (int e = tid; e < CB * CB; e += 256) {
        const int r = e & (CB - 1), cc = e >> 5;
        if (r < kb && cc < wi) R[(p + r) + (pti + cc) * n] = T4[r][cc];
      }
    }
    return;
  }

  // X_ik = -Acc_ik X_kk (rows of block i) into T4, or X_kk itself; role 1
  // also R_k,tj = X_kk^T G_k,tj into Dm
  const double(*Xi)[CHP] = Xk;
  {
    const ci_d4 pb = (role == 1) ? ci_mma<true>(Xk, T1, z4) : z4;
    const ci_d4 xa = (i < k) ? ci_mma<false>(T2, Xk, z4) : z4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (role == 1) Dm[ci_row(r)][cl] = pb[r];
      if (i < k) T4[ci_row(r)][cl] = -xa[r];
    }
    __syncthreads();
    if (i < k) Xi = T4;
  }
  if (role == 1) {
    ci_d4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = (i < k) ? T3[ci_row(r)][cl] : 0.0;
    acc = ci_mma<false>(Xi, Dm, acc);
    const int rows = (i == k) ? kb : CB, wj = bw(tj);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rw = ci_row(r);
      if (rw < rows && cl < wj) Acc[(pi + rw) + (ptj + cl) * n] = acc[r];
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
    return householder_qr_f64(c, m, n, Y, ldy, R, W.p);
  }
  // pass 2: Q = Q1 R2^-1 into Y
  RSVD_TRY(cholqr_pass(c, m, n, W.p, m, Y, ldy, G.p, G2.p, Rinv.p, diag0.p, acc.p, fail.p));
  RSVD_TRY(read_flag(c, fail.p, &h));
  if (h) {
    // range(Q1) = range(Y): Householder on Q1, R = R_hh * R1
    if (used_fallback) *used_fallback = 1;
    RSVD_TRY(copy2d_f64(c, m, n, W.p, m, Y, ldy));
    RSVD_TRY(householder_qr_f64(c, m, n, Y, ldy, R ? G2.p : nullptr, W.p));
    if (R) RSVD_TRY(gemm_f64(c, 'N', 'N', n, n, n, 1.0, G2.p, n, G1.p, n, 0.0, R, n));
    return 0;
  }
  if (R) RSVD_TRY(gemm_f64(c, 'N', 'N', n, n, n, 1.0, G2.p, n, G1.p, n, 0.0, R, n));
  return 0;
}

}  // namespace rsvd
