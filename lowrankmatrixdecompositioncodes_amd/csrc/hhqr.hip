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
// thread); column j takes two launches -- (1) every workgroup sums the G
// partial norms of column j (same order in every workgroup, so all of them
// agree bit for bit), forms beta, tau and v on its rows and writes its partial
// dot products of v with the panel's later columns; (2) every workgroup sums
// those partials (fixed order), updates its rows of the later columns and
// writes its partial norm of column j + 1.  No atomics: the result does not
// depend on the order workgroups run in.  The trailing matrix is updated per
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

// Partial sum of squares of column col over rows (r0, m) on this
// workgroup's chunk into part[g]; the owner of row r0 stores Y(r0, col) in
// *alpha (read by the next launches: the diagonal entry itself is rewritten
// with beta while other workgroups may still be running).
__global__ void __launch_bounds__(kThr)
hh_norm_kernel(int64_t m, double* __restrict__ Y, int64_t ldy, int64_t j0, int64_t chunk,
               int64_t r0, int64_t col, double* __restrict__ part, double* __restrict__ alpha) {
  __shared__ double red[4];
  const int64_t base = j0 + blockIdx.x * chunk;
  const int64_t end = base + chunk < m ? base + chunk : m;
  const double* y = Y + col * ldy;
  double s = 0.0;
  for (int64_t i = base + threadIdx.x; i < end; i += kThr) {
    if (i > r0) s += y[i] * y[i];
    if (i == r0) *alpha = y[i];
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Column c of the panel at j0 (global column j = j0 + c): reflector from the
// summed norm partials, v written over the column below the diagonal, beta on
// it (the owner of row j), tau (workgroup 0), and this workgroup's partial
// dot products v^T Y(:, j0 + k) for k in (c, nb) into dots[g * kNb + k].
__global__ void __launch_bounds__(kThr)
hh_reflect_kernel(int64_t m, double* __restrict__ Y, int64_t ldy, int64_t j0, int c, int nb,
                  int64_t chunk, int groups, const double* __restrict__ part,
                  const double* __restrict__ alpha_p, double* __restrict__ tau,
                  double* __restrict__ dots) {
  __shared__ double red[4];
  __shared__ double wred[4][kNb];
  const int64_t j = j0 + c;
  double s = 0.0;
  for (int g = threadIdx.x; g < groups; g += kThr) s += part[g];
  const double sig = block_sum(s, red);
  const double alpha = *alpha_p;
  double t = 0.0, beta = alpha, scale = 1.0;
  if (sig != 0.0) {
    const double nrm = sqrt(alpha * alpha + sig);
    beta = alpha >= 0.0 ? -nrm : nrm;
    t = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) tau[j] = t;
  const int64_t base = j0 + blockIdx.x * chunk;
  const int64_t end = base + chunk < m ? base + chunk : m;
  double* yj = Y + j * ldy;
  double acc[kNb];
#pragma unroll
  for (int k = 0; k < kNb; ++k) acc[k] = 0.0;
  for (int64_t i = base + threadIdx.x; i < end; i += kThr) {
    if (i < j) continue;
    double v = 1.0;
    if (i == j) {
      yj[i] = beta;
    } else {
      v = yj[i] * scale;
      yj[i] = v;
    }
#pragma unroll
    for (int k = 0; k < kNb; ++k)
      if (k > c && k < nb) acc[k] += v * Y[i + (j0 + k) * ldy];
  }
  const double tot = wave_sum32(acc);
  const int lane = threadIdx.x & 63;
  if ((lane & 1) == 0) wred[threadIdx.x >> 6][lane >> 1] = tot;
  __syncthreads();
  if (threadIdx.x < kNb) {
    const int k = threadIdx.x;
    dots[(int64_t)blockIdx.x * kNb + k] = (wred[0][k] + wred[1][k]) + (wred[2][k] + wred[3][k]);
  }
}

// Apply H_j to the panel's later columns on this workgroup's rows
// (w_k = tau_j * sum of the dot partials, fixed order), then the partial norm
// of column j + 1 below row j + 1 and its new diagonal entry, as
// hh_norm_kernel.
__global__ void __launch_bounds__(kThr)
hh_update_kernel(int64_t m, double* __restrict__ Y, int64_t ldy, int64_t j0, int c, int nb,
                 int64_t chunk, int groups, const double* __restrict__ tau,
                 const double* __restrict__ dots, double* __restrict__ part,
                 double* __restrict__ alpha) {
  __shared__ double red[4];
  __shared__ double psum[kThr / kNb][kNb];
  __shared__ double w[kNb];
  const int64_t j = j0 + c;
  {
    const int k = threadIdx.x & (kNb - 1), sub = threadIdx.x / kNb;
    double s = 0.0;
    for (int g = sub; g < groups; g += kThr / kNb) s += dots[(int64_t)g * kNb + k];
    psum[sub][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < kNb) {
    const int k = threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < kThr / kNb; ++u) s += psum[u][k];
    w[k] = tau[j] * s;
  }
  __syncthreads();
  const int64_t base = j0 + blockIdx.x * chunk;
  const int64_t end = base + chunk < m ? base + chunk : m;
  const double* yj = Y + j * ldy;
  const int64_t jn = j + 1;
  const bool next = c + 1 < nb;
  double s = 0.0;
  for (int64_t i = base + threadIdx.x; i < end; i += kThr) {
    if (i < j) continue;
    const double v = (i == j) ? 1.0 : yj[i];
#pragma unroll
    for (int k = 0; k < kNb; ++k)
      if (k > c && k < nb) Y[i + (j0 + k) * ldy] -= v * w[k];
    if (next) {
      const double x = Y[i + jn * ldy];
      if (i > jn) s += x * x;
      if (i == jn) *alpha = x;
    }
  }
  s = block_sum(s, red);
  if (next && threadIdx.x == 0) part[blockIdx.x] = s;
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
__global__ void __launch_bounds__(64)
hh_larft_kernel(int nb, const double* __restrict__ S, int64_t lds, const double* __restrict__ tau,
                double* __restrict__ T) {
  __shared__ double Ts[kNb][kNb + 1];
  const int r = threadIdx.x;
  if (r < kNb)
    for (int q = 0; q < kNb; ++q) Ts[r][q] = 0.0;
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    const double ti = tau[i];
    double x = 0.0;
    if (r < i)
      for (int q = r; q < i; ++q) x += Ts[r][q] * S[q + i * lds];
    __syncthreads();
    if (r < i) Ts[r][i] = -ti * x;
    if (r == i) Ts[i][i] = ti;
    __syncthreads();
  }
  if (r < kNb)
    for (int q = 0; q < kNb; ++q) T[r + q * kNb] = Ts[r][q];
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
  DBuf<double> own_v, S, T, W, W2, part, dots, alpha;
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
    RSVD_TRY(part.alloc(kMaxGroups));
    RSVD_TRY(dots.alloc((size_t)kMaxGroups * kNb));
    RSVD_TRY(alpha.alloc(1));
    return 0;
  }
};

// V of the panel at j0 (rows j0..m-1, nb columns) into wk.V (ld rows) and its
// T into wk.T.
int panel_vt(Context* c, int64_t m, int64_t j0, int nb, const double* Y, int64_t ldy,
             const double* tau, HhWork& wk) {
  const int64_t rows = m - j0;
  hipLaunchKernelGGL(hh_extract_v_kernel, dim3(grid_1d(rows * nb)), dim3(kThr), 0, c->stream,
                     rows, nb, Y + j0 + j0 * ldy, ldy, wk.V, rows);
  RSVD_LAUNCH_CHECK();
  RSVD_TRY(gemm_f64(c, 'T', 'N', nb, nb, rows, 1.0, wk.V, rows, wk.V, rows, 0.0, wk.S.p, nb));
  hipLaunchKernelGGL(hh_larft_kernel, dim3(1), dim3(64), 0, c->stream, nb, (const double*)wk.S.p,
                     (int64_t)nb, tau + j0, wk.T.p);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// C (rows x nc, ld ldc) <- (I - V op(T) V^T) C: op(T) = T^T for the QR's
// trailing update (H^T C), T for the formation of Q (H C).
int apply_block_reflector(Context* c, int64_t rows, int64_t nc, int nb, HhWork& wk, bool trans,
                          double* C, int64_t ldc) {
  if (nc <= 0) return 0;
  // W = C^T V (nc x nb); W2 = W op(T)^T; C -= V W2^T
  RSVD_TRY(gemm_f64(c, 'T', 'N', nc, nb, rows, 1.0, C, ldc, wk.V, rows, 0.0, wk.W.p, nc));
  RSVD_TRY(gemm_f64(c, 'N', trans ? 'N' : 'T', nc, nb, nb, 1.0, wk.W.p, nc, wk.T.p, kNb, 0.0,
                    wk.W2.p, nc));
  RSVD_TRY(gemm_f64(c, 'N', 'T', rows, nc, nb, -1.0, wk.V, rows, wk.W2.p, nc, 1.0, C, ldc));
  return 0;
}

}  // namespace

int geqrf_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* tau,
              double* vwork) {
  if (n <= 0 || m <= 0) return 0;
  if (m < n) return set_error("geqrf: m=%lld < n=%lld", (long long)m, (long long)n);
  HhWork wk;
  RSVD_TRY(wk.alloc(m, n, vwork));
  for (int64_t j0 = 0; j0 < n; j0 += kNb) {
    const int nb = (int)(n - j0 < kNb ? n - j0 : kNb);
    const PanelGrid pg = panel_grid(m - j0);
    hipLaunchKernelGGL(hh_norm_kernel, dim3(pg.groups), dim3(kThr), 0, c->stream, m, Y, ldy, j0,
                       pg.chunk, j0, j0, wk.part.p, wk.alpha.p);
    RSVD_LAUNCH_CHECK();
    for (int cc = 0; cc < nb; ++cc) {
      hipLaunchKernelGGL(hh_reflect_kernel, dim3(pg.groups), dim3(kThr), 0, c->stream, m, Y, ldy,
                         j0, cc, nb, pg.chunk, pg.groups, (const double*)wk.part.p,
                         (const double*)wk.alpha.p, tau, wk.dots.p);
      RSVD_LAUNCH_CHECK();
      if (cc + 1 < nb) {
        hipLaunchKernelGGL(hh_update_kernel, dim3(pg.groups), dim3(kThr), 0, c->stream, m, Y,
                           ldy, j0, cc, nb, pg.chunk, pg.groups, (const double*)tau,
                           (const double*)wk.dots.p, wk.part.p, wk.alpha.p);
        RSVD_LAUNCH_CHECK();
      }
    }
    const int64_t nc = n - j0 - nb;
    if (nc > 0) {
      RSVD_TRY(panel_vt(c, m, j0, nb, Y, ldy, tau, wk));
      RSVD_TRY(apply_block_reflector(c, m - j0, nc, nb, wk, true, Y + j0 + (j0 + nb) * ldy, ldy));
    }
  }
  return 0;
}

int orgqr_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, const double* tau,
              double* vwork) {
  if (n <= 0 || m <= 0) return 0;
  if (m < n) return set_error("orgqr: m=%lld < n=%lld", (long long)m, (long long)n);
  HhWork wk;
  RSVD_TRY(wk.alloc(m, n, vwork));
  for (int64_t j0 = ((n - 1) / kNb) * kNb; j0 >= 0; j0 -= kNb) {
    const int nb = (int)(n - j0 < kNb ? n - j0 : kNb);
    RSVD_TRY(panel_vt(c, m, j0, nb, Y, ldy, tau, wk));
    // the later columns (rows j0..j0+nb-1 of them are zero at this point)
    RSVD_TRY(apply_block_reflector(c, m - j0, n - j0 - nb, nb, wk, false,
                                   Y + j0 + (j0 + nb) * ldy, ldy));
    hipLaunchKernelGGL(hh_form_panel_kernel, dim3(grid_1d(m)), dim3(kThr), 0, c->stream, m - j0,
                       j0, nb, (const double*)wk.V, m - j0, (const double*)wk.T.p, Y, ldy);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace rsvd
