// jacobi.hip -- small dense eigen/SVD solver of the rSVD pipeline on the GPU,
// replacing LAPACKE_dsyev on BB^T (MKL/mvf.c:1206-1209, vnum = 2) and
// LAPACKE_dgesvd on R-hat (MKL/mvf.c:1270-1284, vnum = 1).
//
// One-sided (Hestenes) Jacobi with block tournament ordering.  The n columns
// of W are cut into blocks of b columns; a launch ("round") pairs the blocks
// with the circle method and gives each pair to one workgroup, which stages
// its 2b columns (and the matching columns of the accumulated rotation V) in
// LDS, runs one cyclic inner sweep over all 2b(2b-1)/2 column pairs (one wave
// per pair, rounds of the circle method again) and writes the block back.
// nblocks-1 rounds make one outer sweep; sweeps repeat until every pair met
// in a sweep was already orthogonal to |x.y| <= tol*|x||y|.
//
// For a symmetric positive semidefinite G (BB^T), W = G*V converges to
// Uhat*Lambda with V = Uhat, so eigenvalues are the column norms of W and the
// eigenvectors are the columns of V.  For R-hat, W = R*V converges to
// Uhat*Sigma and V to Vhat.
#include <cmath>
#include <cstring>

#include "rsvd_internal.h"

namespace rsvd {

// Player at position `pos` in round r of the circle method with N players.
__device__ __forceinline__ int circle_player(int pos, int r, int N) {
  return pos == 0 ? 0 : 1 + ((pos - 1 + r) % (N - 1));
}

template <bool USE_LDS>
__global__ void __launch_bounds__(1024)
jacobi_round_kernel(int64_t L, int64_t n, double* __restrict__ W, int64_t ldw,
                    double* __restrict__ V, int64_t ldv, int64_t Lv, int b, int nblocks_real,
                    int nplayers, int round, double tol,
                    unsigned long long* __restrict__ offmax) {
  extern __shared__ double sm[];
  const int pairs = nplayers / 2;
  const int w = blockIdx.x;
  if (w >= pairs) return;
  const int pa = circle_player(w, round, nplayers);
  const int pb = circle_player(nplayers - 1 - w, round, nplayers);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int S = 2 * b;
  const int64_t colLen = L + Lv;

  auto slot_col = [&](int s) -> int64_t {
    int blk = (s < b) ? pa : pb;
    int off = (s < b) ? s : s - b;
    if (blk >= nblocks_real) return -1;
    int64_t col = (int64_t)blk * b + off;
    return col < n ? col : -1;
  };

  if (USE_LDS) {
    for (int s = 0; s < S; ++s) {
      int64_t col = slot_col(s);
      double* dst = sm + (int64_t)s * colLen;
      if (col < 0) continue;
      for (int64_t r = threadIdx.x; r < L; r += blockDim.x) dst[r] = W[r + col * ldw];
      for (int64_t r = threadIdx.x; r < Lv; r += blockDim.x) dst[L + r] = V[r + col * ldv];
    }
    __syncthreads();
  }

  double mymax = 0.0;
  for (int t = 0; t < S - 1; ++t) {
    if (wave < b) {
      const int sx = circle_player(wave, t, S);
      const int sy = circle_player(S - 1 - wave, t, S);
      const int64_t cx = slot_col(sx), cy = slot_col(sy);
      if (cx >= 0 && cy >= 0) {
        double *wx, *wy, *vx, *vy;
        if (USE_LDS) {
          wx = sm + (int64_t)sx * colLen;
          wy = sm + (int64_t)sy * colLen;
          vx = wx + L;
          vy = wy + L;
        } else {
          wx = W + cx * ldw;
          wy = W + cy * ldw;
          vx = V ? V + cx * ldv : nullptr;
          vy = V ? V + cy * ldv : nullptr;
        }
        double a = 0, bb = 0, g = 0;
        for (int64_t r = lane; r < L; r += 64) {
          double x = wx[r], y = wy[r];
          a += x * x;
          bb += y * y;
          g += x * y;
        }
        for (int off = 32; off > 0; off >>= 1) {
          a += __shfl_xor(a, off, 64);
          bb += __shfl_xor(bb, off, 64);
          g += __shfl_xor(g, off, 64);
        }
        double rel = 0.0;
        if (a > 0.0 && bb > 0.0) rel = fabs(g) / (sqrt(a) * sqrt(bb));
        if (rel > mymax) mymax = rel;
        if (rel > tol) {
          double zeta = (bb - a) / (2.0 * g);
          double tt;
          if (fabs(zeta) > 1e150) tt = 0.5 / zeta;
          else tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          double cs = 1.0 / sqrt(1.0 + tt * tt);
          double sn = cs * tt;
          for (int64_t r = lane; r < L; r += 64) {
            double x = wx[r], y = wy[r];
            wx[r] = cs * x - sn * y;
            wy[r] = sn * x + cs * y;
          }
          if (Lv > 0) {
            for (int64_t r = lane; r < Lv; r += 64) {
              double x = vx[r], y = vy[r];
              vx[r] = cs * x - sn * y;
              vy[r] = sn * x + cs * y;
            }
          }
        }
      }
    }
    __syncthreads();
  }

  if (USE_LDS) {
    for (int s = 0; s < S; ++s) {
      int64_t col = slot_col(s);
      const double* src = sm + (int64_t)s * colLen;
      if (col < 0) continue;
      for (int64_t r = threadIdx.x; r < L; r += blockDim.x) W[r + col * ldw] = src[r];
      for (int64_t r = threadIdx.x; r < Lv; r += blockDim.x) V[r + col * ldv] = src[L + r];
    }
  }
  // workgroup max of the measured off-orthogonality
  __shared__ double red[16];
  for (int off = 32; off > 0; off >>= 1) mymax = fmax(mymax, __shfl_xor(mymax, off, 64));
  if (lane == 0) red[wave] = mymax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, red[i]);
    atomicMax(offmax, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void set_identity_kernel(int64_t n, double* __restrict__ V, int64_t ldv) {
  int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % n, j = idx / n;
    V[i + j * ldv] = (i == j) ? 1.0 : 0.0;
  }
}

// dynamic LDS budget for the staged columns (the CU has 160 KiB; a little
// is left for the kernel's static reduction buffer)
static constexpr size_t kMaxLds = 152 * 1024;

int jacobi_f64(Context* c, int64_t L, int64_t n, double* W, int64_t ldw, double* V, int64_t ldv,
               int max_sweeps, int* sweeps_out) {
  if (sweeps_out) *sweeps_out = 0;
  if (n <= 1 || L == 0) return 0;
  const int64_t Lv = V ? n : 0;
  if (V) {
    int blocks = (int)((n * n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(set_identity_kernel, dim3(blocks), dim3(256), 0, c->stream, n, V, ldv);
    RSVD_LAUNCH_CHECK();
  }
  // block size: largest b in {16, 8} whose 2b staged columns fit in LDS
  int b = 16;
  bool lds = false;
  for (int cand : {16, 8}) {
    if ((size_t)2 * cand * (size_t)(L + Lv) * sizeof(double) <= kMaxLds) {
      b = cand;
      lds = true;
      break;
    }
  }
  if ((int64_t)b > (n + 1) / 2) {
    b = (int)((n + 1) / 2);
  }
  const int nblocks = (int)((n + b - 1) / b);
  const int nplayers = nblocks + (nblocks & 1);
  const size_t shmem = lds ? (size_t)2 * b * (size_t)(L + Lv) * sizeof(double) : 0;
  static bool attr_set = false;
  if (!attr_set) {
    RSVD_HIP(hipFuncSetAttribute((const void*)jacobi_round_kernel<true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds));
    attr_set = true;
  }
  const double tol = 2.220446049250313e-16 * sqrt((double)L);
  DBuf<unsigned long long> off;
  RSVD_TRY(off.alloc(1));
  unsigned long long h = 0;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    RSVD_HIP(hipMemsetAsync(off.p, 0, sizeof(unsigned long long), c->stream));
    for (int r = 0; r < nplayers - 1; ++r) {
      KTimer kt(c, "jacobi_round");
      if (lds)
        hipLaunchKernelGGL(jacobi_round_kernel<true>, dim3(nplayers / 2), dim3(64 * b), shmem,
                           c->stream, L, n, W, ldw, V, ldv, Lv, b, nblocks, nplayers, r, tol,
                           off.p);
      else
        hipLaunchKernelGGL(jacobi_round_kernel<false>, dim3(nplayers / 2), dim3(64 * b), 0,
                           c->stream, L, n, W, ldw, V, ldv, Lv, b, nblocks, nplayers, r, tol,
                           off.p);
      RSVD_LAUNCH_CHECK();
    }
    RSVD_HIP(hipMemcpyAsync(&h, off.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    RSVD_HIP(hipStreamSynchronize(c->stream));
    double mx;
    memcpy(&mx, &h, sizeof mx);
    if (!(mx > tol)) {
      ++sweep;
      break;
    }
  }
  if (sweeps_out) *sweeps_out = sweep;
  return 0;
}

// One wave per column: nrm[j] = ||W(:,j)||, optionally W(:,j) /= nrm[j].
__global__ void colnorms_kernel(int64_t L, int64_t n, double* __restrict__ W, int64_t ldw,
                                double* __restrict__ nrm, int normalise) {
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;
  double* w = W + j * ldw;
  double s = 0;
  for (int64_t r = lane; r < L; r += 64) s += w[r] * w[r];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  const double nv = sqrt(s);
  if (lane == 0) nrm[j] = nv;
  if (normalise && nv > 0.0) {
    const double inv = 1.0 / nv;
    for (int64_t r = lane; r < L; r += 64) w[r] *= inv;
  }
}

int colnorms_f64(Context* c, int64_t L, int64_t n, double* W, int64_t ldw, double* nrm,
                 int normalise) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(colnorms_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c->stream, L, n,
                     W, ldw, nrm, normalise);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// rank[j] = position of vals[j] in the sorted order (stable).
__global__ void rank_kernel(int64_t n, const double* __restrict__ v, int descending,
                            int64_t* __restrict__ rank) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const double x = v[j];
    int64_t r = 0;
    for (int64_t i = 0; i < n; ++i) {
      const double y = v[i];
      bool before = descending ? (y > x) : (y < x);
      if (before || (y == x && i < j)) ++r;
    }
    rank[j] = r;
  }
}

__global__ void permute_cols_kernel(int64_t rows, int64_t n, const double* __restrict__ src,
                                    int64_t lds, const int64_t* __restrict__ rank,
                                    double* __restrict__ dst, int64_t ldd) {
  int64_t total = rows * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % rows, j = idx / rows;
    dst[i + rank[j] * ldd] = src[i + j * lds];
  }
}

__global__ void permute_vals_kernel(int64_t n, const double* __restrict__ src,
                                    const int64_t* __restrict__ rank, double* __restrict__ dst) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    dst[rank[j]] = src[j];
}

static int permute_matrix(Context* c, int64_t rows, int64_t n, double* X, int64_t ldx,
                          const int64_t* rank) {
  if (!X || rows == 0) return 0;
  DBuf<double> tmp;
  RSVD_TRY(tmp.alloc((size_t)(rows * n)));
  int blocks = (int)((rows * n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(permute_cols_kernel, dim3(blocks), dim3(256), 0, c->stream, rows, n,
                     (const double*)X, ldx, rank, tmp.p, rows);
  RSVD_LAUNCH_CHECK();
  RSVD_TRY(copy2d_f64(c, rows, n, tmp.p, rows, X, ldx));
  return 0;
}

int sort_pairs_f64(Context* c, int64_t n, double* vals, int descending, double* X, int64_t rowsX,
                   int64_t ldx, double* Y, int64_t rowsY, int64_t ldy) {
  if (n <= 1) return 0;
  DBuf<int64_t> rank;
  DBuf<double> tv;
  RSVD_TRY(rank.alloc((size_t)n));
  RSVD_TRY(tv.alloc((size_t)n));
  int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL(rank_kernel, dim3(blocks), dim3(256), 0, c->stream, n, (const double*)vals,
                     descending, rank.p);
  RSVD_LAUNCH_CHECK();
  hipLaunchKernelGGL(permute_vals_kernel, dim3(blocks), dim3(256), 0, c->stream, n,
                     (const double*)vals, (const int64_t*)rank.p, tv.p);
  RSVD_LAUNCH_CHECK();
  RSVD_HIP(hipMemcpyAsync(vals, tv.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  RSVD_TRY(permute_matrix(c, rowsX, n, X, ldx, rank.p));
  RSVD_TRY(permute_matrix(c, rowsY, n, Y, ldy, rank.p));
  return 0;
}

}  // namespace rsvd
