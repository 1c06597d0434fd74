// hhqr.hip -- blocked Householder QR (dgeqrf + dorgqr semantics) on the GPU:
// the robust orthonormalisation behind CholeskyQR2's breakdown fallback and
// the Q of the column-pivoted QR.  Replaces the column-by-column geqr2 / org2r
// (MKL/mvf.c:1214-1263 calls LAPACK dgeqrf + dorgqr).
//
// Layout: Y (m x n, column-major, m >= n) in place; reflector j is
// v_j = [0 (rows < j); 1; Y(j+1:m, j)] with tau_j, H_j = I - tau_j v_j v_j^T,
// R in the upper triangle, LAPACK's dlarfg sign convention
// (beta = -sign(alpha) ||x||).
//
// Panels of kNb = 32 columns.  Panel factorisation: the rows of the panel are
// split over G workgroups (a contiguous chunk each, one to a few rows per
// thread); column j takes one launch: every workgroup sums the G partial
// norms of column j and partial dot products of column j with the panel's
// later columns (same order in every workgroup, so all of them agree bit for
// bit), forms beta, tau and v, updates its rows of the later columns and,
// from the rows it just updated, writes its partials for column j + 1.  The
// dot products are taken against the unscaled column x (v^T y = y_j +
// x^T y / (alpha - beta)), which is what lets one pass do both.  No atomics:
// the result does not depend on the order workgroups run in.  The trailing matrix is updated per
// panel with the compact WY form H_0 ... H_{nb-1} = I - V T V^T (T upper
// triangular, LAPACK dlarft forward / columnwise) in three FP64 GEMMs on the
// matrix cores.  Q = H_0 ... H_{n-1} [I; 0] is formed panel by panel from the
// last one (dorgqr order): the later columns get I - V T V^T by GEMMs, the
// panel's own columns are [0; E - V (T V1^T)] (E = [I; 0], V1 the top
// nb x nb of V).
#include "rsvd_internal.h"

namespace rsvd {

namespace {

constexpr int kNb = 32;
constexpr int kThr = 256;
constexpr int kMaxGroups = 512;

// The panel's rows [j0, m) over G workgroups: `chunk` rows each, a multiple
// of kThr.
struct PanelGrid {
  int groups;
  int64_t chunk;
};

PanelGrid panel_grid(int64_t rows) {
  const int64_t per = (rows + kMaxGroups - 1) / kMaxGroups;
  const int64_t chunk = ((per + kThr - 1) / kThr) * kThr;
  return {(int)((rows + chunk - 1) / chunk), chunk};
}

// Fixed-order sum of x over the workgroup (every workgroup running this on
// the same data gets the same bits).
__device__ double block_sum(double x, double* red) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// Sum over the wave's 64 lanes of each of a[0..32): reduce-scatter by halves
// (lanes that differ in bit b exchange the half the other keeps), then the
// last pair.  Lane l returns the total for index l >> 1.
template <int H>
__device__ __forceinline__ void halve(double (&a)[kNb], int lane) {
  const bool up = (lane & (2 * H)) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double send = up ? a[i] : a[i + H];
    const double keep = up ? a[i + H] : a[i];
    a[i] = keep + __shfl_xor(send, 2 * H, 64);
  }
}

__device__ double wave_sum32(double (&a)[kNb]) {
  const int lane = threadIdx.x & 63;
  halve<16>(a, lane);
  halve<8>(a, lane);
  halve<4>(a, lane);
  halve<2>(a, lane);
  halve<1>(a, lane);
  return a[0] + __shfl_xor(a[0], 1, 64);
}

// Per-column state passed from one launch to the next (two copies, by the
// parity of the column: a launch reads one and writes the other):
//   pn[g]        partial sum over workgroup g's rows i > j of x_i^2 (x =
//                column j as the previous launch left it);
//   pd[g][k]     partial sum over the same rows of x_i Y(i, j0 + k), k > c;
//   row[0]       x_j (alpha), row[k] = Y(j, j0 + k) for k > c: row j as the
//                previous launch left it (the owner of row j rewrites it in
//                this launch while the others still read it).
struct ColState {
  double* pn;
  double* pd;
  double* row;
};

// Partials and row j0 of the panel's first column (j = j0, c = 0) into st.
__global__ void __launch_bounds__(kThr)
hh_start_kernel(int64_t m, const double* __restrict__ Y, int64_t ldy, int64_t j0, int nb,
                int64_t chunk, ColState st) {
  __shared__ double red[4];
  __shared__ double wred[4][kNb];
  const int64_t base = j0 + blockIdx.x * chunk;
  const int64_t end = base + chunk < m ? base + chunk : m;
  double s = 0.0, acc[kNb];
#pragma unroll
  for (int k = 0; k < kNb; ++k) acc[k] = 0.0;
  for (int64_t i = base + threadIdx.x; i < end; i += kThr) {
    const double x = Y[i + j0 * ldy];
    if (i == j0) {
      st.row[0] = x;
      for (int k = 1; k < nb; ++k) st.row[k] = Y[i + (j0 + k) * ldy];
      continue;
    }
    s += x * x;
#pragma unroll
    for (int k = 1; k < kNb; ++k)
      if (k < nb) acc[k] += x * Y[i + (j0 + k) * ldy];
  }
  s = block_sum(s, red);
  const double tot = wave_sum32(acc);
  const int lane = threadIdx.x & 63;
  if ((lane & 1) == 0) wred[threadIdx.x >> 6][lane >> 1] = tot;
  __syncthreads();
  if (threadIdx.x < kNb) {
    const int k = threadIdx.x;
    st.pd[(int64_t)blockIdx.x * kNb + k] = (wred[0][k] + wred[1][k]) + (wred[2][k] + wred[3][k]);
  }
  if (threadIdx.x == 0) st.pn[blockIdx.x] = s;
}

// Column c of the panel at j0 (j = j0 + c), one launch: every workgroup sums
// the partials of `in` (fixed order: all of them get the same bits), forms
// LAPACK's reflector (beta = -sign(alpha) sqrt(alpha^2 + sig), tau =
// (beta - alpha) / beta, v = x / (alpha - beta)) and
// w_k = tau v^T Y(:, j0 + k) = tau (Y(j, j0 + k) + sum_{i > j} x_i Y(i, j0 + k)
// / (alpha - beta)), writes v and beta over column j and applies
// Y(:, j0 + k) -= v w_k to its rows; then the partials and row j + 1 of
// column c + 1 into `out`, from the updated rows it still holds.
__global__ void __launch_bounds__(kThr)
hh_column_kernel(int64_t m, double* __restrict__ Y, int64_t ldy, int64_t j0, int c, int nb,
                 int64_t chunk, int groups, ColState in, ColState out,
                 double* __restrict__ tau) {
  __shared__ double red[4];
  __shared__ double psum[kThr / kNb][kNb];
  __shared__ double wred[4][kNb];
  __shared__ double w[kNb];
  const int64_t j = j0 + c;
  double s = 0.0;
  for (int g = threadIdx.x; g < groups; g += kThr) s += in.pn[g];
  const double sig = block_sum(s, red);
  {
    const int k = threadIdx.x & (kNb - 1), sub = threadIdx.x / kNb;
    double d = 0.0;
    // loads issued ahead of the (in-order) adds
#pragma unroll 8
    for (int g = sub; g < groups; g += kThr / kNb) d += in.pd[(int64_t)g * kNb + k];
    psum[sub][k] = d;
  }
  const double alpha = in.row[0];
  double t = 0.0, beta = alpha, scale = 1.0;
  if (sig != 0.0) {
    const double nrm = sqrt(alpha * alpha + sig);
    beta = alpha >= 0.0 ? -nrm : nrm;
    t = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  __syncthreads();
  if (threadIdx.x < kNb) {
    const int k = threadIdx.x;
    double d = 0.0;
#pragma unroll
    for (int u = 0; u < kThr / kNb; ++u) d += psum[u][k];
    w[k] = (k > c && k < nb) ? t * (in.row[k] + scale * d) : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) tau[j] = t;
  __syncthreads();
  const int64_t base = j0 + blockIdx.x * chunk;
  const int64_t end = base + chunk < m ? base + chunk : m;
  const int cn = c + 1;
  const bool next = cn < nb;
  double sn = 0.0, acc[kNb];
#pragma unroll
  for (int k = 0; k < kNb; ++k) acc[k] = 0.0;
  for (int64_t i = base + threadIdx.x; i < end; i += kThr) {
    if (i < j) continue;
    double v = 1.0;
    if (i == j) {
      Y[i + j * ldy] = beta;
    } else {
      v = Y[i + j * ldy] * scale;
      Y[i + j * ldy] = v;
    }
    double y[kNb];
#pragma unroll
    for (int k = 0; k < kNb; ++k) {
      y[k] = 0.0;
      if (k > c && k < nb) {
        y[k] = Y[i + (j0 + k) * ldy] - v * w[k];
        Y[i + (j0 + k) * ldy] = y[k];
      }
    }
    if (!next || i <= j) continue;
    // column c + 1 as updated: x' = y[cn]
    double xn = 0.0;
#pragma unroll
    for (int k = 0; k < kNb; ++k)
      if (k == cn) xn = y[k];
    if (i == j + 1) {
      out.row[0] = xn;
#pragma unroll
      for (int k = 0; k < kNb; ++k)
        if (k > cn && k < nb) out.row[k] = y[k];
      continue;
    }
    sn += xn * xn;
#pragma unroll
    for (int k = 0; k < kNb; ++k)
      if (k > cn && k < nb) acc[k] += xn * y[k];
  }
  sn = block_sum(sn, red);
  const double tot = wave_sum32(acc);
  const int lane = threadIdx.x & 63;
  if ((lane & 1) == 0) wred[threadIdx.x >> 6][lane >> 1] = tot;
  __syncthreads();
  if (next && threadIdx.x < kNb) {
    const int k = threadIdx.x;
    out.pd[(int64_t)blockIdx.x * kNb + k] = (wred[0][k] + wred[1][k]) + (wred[2][k] + wred[3][k]);
  }
  if (next && threadIdx.x == 0) out.pn[blockIdx.x] = sn;
}

// V (rows x nb, ld ldv) = the panel's reflectors as a dense matrix: unit
// diagonal, zeros above, Y's entries below.
__global__ void __launch_bounds__(kThr)
hh_extract_v_kernel(int64_t rows, int nb, const double* __restrict__ Y, int64_t ldy,
                    double* __restrict__ V, int64_t ldv) {
  const int64_t total = rows * nb;
  for (int64_t e = blockIdx.x * (int64_t)kThr + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kThr) {
    const int64_t i = e % rows, k = e / rows;
    V[i + k * ldv] = i < k ? 0.0 : (i == k ? 1.0 : Y[i + k * ldy]);
  }
}

// T (nb x nb, ld kNb, upper) of H_0 ... H_{nb-1} = I - V T V^T from
// S = V^T V and tau (dlarft, forward, columnwise):
// T(i, i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) S(0:i, i).
__global__ void __launch_bounds__(kThr)
hh_larft_kernel(int nb, const double* __restrict__ S, int64_t lds, const double* __restrict__ tau,
                double* __restrict__ T) {
  __shared__ double Ts[kNb][kNb + 1];
  __shared__ double Ss[kNb][kNb + 1];
  for (int e = threadIdx.x; e < kNb * kNb; e += kThr) {
    const int q = e % kNb, i = e / kNb;
    Ts[q][i] = 0.0;
    Ss[q][i] = (q < nb && i < nb) ? S[q + i * lds] : 0.0;
  }
  __syncthreads();
  // step i: thread (r, u) = (tid / 8, tid % 8) sums the terms q = u, u + 8,
  // ... of sum_{q = r}^{i - 1} T(r, q) S(q, i); the 8 lanes of r combine by
  // a fixed butterfly
  const int r = threadIdx.x >> 3, u = threadIdx.x & 7;
  for (int i = 0; i < nb; ++i) {
    double x = 0.0;
    if (r < i)
      for (int q = u; q < i; q += 8) x += Ts[r][q] * Ss[q][i];
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    if (u == 0) {
      if (r < i) Ts[r][i] = -tau[i] * x;
      if (r == i) Ts[i][i] = tau[i];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < kNb * kNb; e += kThr)
    T[(e % kNb) + (e / kNb) * kNb] = Ts[e % kNb][e / kNb];
}

// The panel's own columns of Q: rows < j0 zero, rows j0 + i (i < rows) =
// E(i, :) - V(i, :) M with M = T V1^T formed in LDS by every workgroup.
__global__ void __launch_bounds__(kThr)
hh_form_panel_kernel(int64_t rows, int64_t j0, int nb, const double* __restrict__ V,
                     int64_t ldv, const double* __restrict__ T, double* __restrict__ Q,
                     int64_t ldq) {
  __shared__ double Ms[kNb][kNb + 1];
  for (int e = threadIdx.x; e < kNb * kNb; e += kThr) {
    const int k = e % kNb, cc = e / kNb;  // M(k, cc) = sum_q T(k, q) V1(cc, q)
    double x = 0.0;
    if (k < nb && cc < nb)
      for (int q = k; q <= cc && q < nb; ++q) x += T[k + q * kNb] * V[cc + q * ldv];
    Ms[k][cc] = x;
  }
  __syncthreads();
  const int64_t total = j0 + rows;
  for (int64_t r = blockIdx.x * (int64_t)kThr + threadIdx.x; r < total;
       r += (int64_t)gridDim.x * kThr) {
    if (r < j0) {
      for (int cc = 0; cc < nb; ++cc) Q[r + (j0 + cc) * ldq] = 0.0;
      continue;
    }
    const int64_t i = r - j0;
    double vr[kNb];
#pragma unroll
    for (int k = 0; k < kNb; ++k) vr[k] = k < nb ? V[i + k * ldv] : 0.0;
    for (int cc = 0; cc < nb; ++cc) {
      double x = (i == cc) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < kNb; ++k) x -= vr[k] * Ms[k][cc];
      Q[r + (j0 + cc) * ldq] = x;
    }
  }
}

int grid_1d(int64_t total) {
  int64_t b = (total + kThr - 1) / kThr;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

// Buffers of one factorisation: V (rows x kNb), S, T, the W products and the
// partial sums.
struct HhWork {
  DBuf<double> own_v, S, T, W, W2, col;
  double* V = nullptr;
  int alloc(int64_t m, int64_t n, double* vwork) {
    if (vwork) V = vwork;
    else {
      RSVD_TRY(own_v.alloc((size_t)(m * kNb)));
      V = own_v.p;
    }
    RSVD_TRY(S.alloc((size_t)(kNb * kNb)));
    RSVD_TRY(T.alloc((size_t)(kNb * kNb)));
    RSVD_TRY(W.alloc((size_t)((n > 0 ? n : 1) * kNb)));
    RSVD_TRY(W2.alloc((size_t)((n > 0 ? n : 1) * kNb)));
    RSVD_TRY(col.alloc(2 * ((size_t)kMaxGroups * (kNb + 1) + kNb)));
    return 0;
  }
  // the two copies of the per-column state
  ColState state(int par) const {
    double* b = col.p + par * ((size_t)kMaxGroups * (kNb + 1) + kNb);
    return {b, b + kMaxGroups, b + (size_t)kMaxGroups * (kNb + 1)};
  }
};

// V of the panel at j0 (rows j0..m-1, nb columns) into wk.V (ld rows) and,
// unless T is null, its T into T (kNb x kNb).
int panel_vt(Context* c, int64_t m, int64_t j0, int nb, const double* Y, int64_t ldy,
             const double* tau, HhWork& wk, double* T) {
  const int64_t rows = m - j0;
  hipLaunchKernelGGL(hh_extract_v_kernel, dim3(grid_1d(rows * nb)), dim3(kThr), 0, c->stream,
                     rows, nb, Y + j0 + j0 * ldy, ldy, wk.V, rows);
  RSVD_LAUNCH_CHECK();
  if (!T) return 0;
  RSVD_TRY(gemm_f64(c, 'T', 'N', nb, nb, rows, 1.0, wk.V, rows, wk.V, rows, 0.0, wk.S.p, nb));
  hipLaunchKernelGGL(hh_larft_kernel, dim3(1), dim3(kThr), 0, c->stream, nb,
                     (const double*)wk.S.p, (int64_t)nb, tau + j0, T);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// C (rows x nc, ld ldc) <- (I - V op(T) V^T) C: op(T) = T^T for the QR's
// trailing update (H^T C), T for the formation of Q (H C).
int apply_block_reflector(Context* c, int64_t rows, int64_t nc, int nb, HhWork& wk,
                          const double* T, bool trans, double* C, int64_t ldc) {
  if (nc <= 0) return 0;
  // W = C^T V (nc x nb); W2 = W op(T)^T; C -= V W2^T
  RSVD_TRY(gemm_f64(c, 'T', 'N', nc, nb, rows, 1.0, C, ldc, wk.V, rows, 0.0, wk.W.p, nc));
  RSVD_TRY(gemm_f64(c, 'N', trans ? 'N' : 'T', nc, nb, nb, 1.0, wk.W.p, nc, T, kNb, 0.0,
                    wk.W2.p, nc));
  RSVD_TRY(gemm_f64(c, 'N', 'T', rows, nc, nb, -1.0, wk.V, rows, wk.W2.p, nc, 1.0, C, ldc));
  return 0;
}

}  // namespace

int geqrf_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* tau,
              double* vwork, double* Tall) {
  if (n <= 0 || m <= 0) return 0;
  if (m < n) return set_error("geqrf: m=%lld < n=%lld", (long long)m, (long long)n);
  HhWork wk;
  RSVD_TRY(wk.alloc(m, n, vwork));
  for (int64_t j0 = 0; j0 < n; j0 += kNb) {
    const int nb = (int)(n - j0 < kNb ? n - j0 : kNb);
    const PanelGrid pg = panel_grid(m - j0);
    hipLaunchKernelGGL(hh_start_kernel, dim3(pg.groups), dim3(kThr), 0, c->stream, m,
                       (const double*)Y, ldy, j0, nb, pg.chunk, wk.state(0));
    RSVD_LAUNCH_CHECK();
    for (int cc = 0; cc < nb; ++cc) {
      hipLaunchKernelGGL(hh_column_kernel, dim3(pg.groups), dim3(kThr), 0, c->stream, m, Y, ldy,
                         j0, cc, nb, pg.chunk, pg.groups, wk.state(cc & 1),
                         wk.state((cc + 1) & 1), tau);
      RSVD_LAUNCH_CHECK();
    }
    const int64_t nc = n - j0 - nb;
    double* T = Tall ? Tall + (j0 / kNb) * kNb * kNb : wk.T.p;
    if (nc > 0 || Tall) RSVD_TRY(panel_vt(c, m, j0, nb, Y, ldy, tau, wk, T));
    if (nc > 0)
      RSVD_TRY(apply_block_reflector(c, m - j0, nc, nb, wk, T, true, Y + j0 + (j0 + nb) * ldy,
                                     ldy));
  }
  return 0;
}

int orgqr_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, const double* tau,
              double* vwork, const double* Tall) {
  if (n <= 0 || m <= 0) return 0;
  if (m < n) return set_error("orgqr: m=%lld < n=%lld", (long long)m, (long long)n);
  HhWork wk;
  RSVD_TRY(wk.alloc(m, n, vwork));
  for (int64_t j0 = ((n - 1) / kNb) * kNb; j0 >= 0; j0 -= kNb) {
    const int nb = (int)(n - j0 < kNb ? n - j0 : kNb);
    const double* T = Tall ? Tall + (j0 / kNb) * kNb * kNb : wk.T.p;
    RSVD_TRY(panel_vt(c, m, j0, nb, Y, ldy, tau, wk, Tall ? nullptr : wk.T.p));
    // the later columns (rows j0..j0+nb-1 of them are zero at this point)
    RSVD_TRY(apply_block_reflector(c, m - j0, n - j0 - nb, nb, wk, T, false,
                                   Y + j0 + (j0 + nb) * ldy, ldy));
    hipLaunchKernelGGL(hh_form_panel_kernel, dim3(grid_1d(m)), dim3(kThr), 0, c->stream, m - j0,
                       j0, nb, (const double*)wk.V, m - j0, T, Y, ldy);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace rsvd
