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

// Rotation (cs, sn) zeroing the inner product of two columns with squared
// norms a, b and inner product g (Rutishauser's formulas).
__device__ __forceinline__ void jacobi_angle(double a, double b, double g, double* cs,
                                             double* sn) {
  double zeta = (b - a) / (2.0 * g);
  double tt;
  if (fabs(zeta) > 1e150) tt = 0.5 / zeta;
  else tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  *cs = 1.0 / sqrt(1.0 + tt * tt);
  *sn = *cs * tt;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// LDS path: the 2b columns of W for the block pair are staged in LDS, the
// rotations of `inner` cyclic inner sweeps are applied to them and
// accumulated into J (2b x 2b, LDS); V's block columns are then updated once,
// row by row, as V(:, blk) <- V(:, blk) * J.
__global__ void __launch_bounds__(1024)
jacobi_round_lds_kernel(int64_t L, int64_t n, double* __restrict__ W, int64_t ldw,
                        double* __restrict__ V, int64_t ldv, int64_t Lv, int b, int nblocks_real,
                        int nplayers, int round, double tol, int inner,
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
  double* Ws = sm;                       // S columns of length L
  double* J = sm + (int64_t)S * L;       // S x S, column-major
  __shared__ int64_t cols[32];
  __shared__ int any_rot;
  if (threadIdx.x < S) {
    const int s = threadIdx.x;
    const int blk = (s < b) ? pa : pb;
    const int off = (s < b) ? s : s - b;
    int64_t col = (int64_t)blk * b + off;
    cols[s] = (blk < nblocks_real && col < n) ? col : -1;
  }
  for (int e = threadIdx.x; e < S * S; e += blockDim.x) J[e] = ((e % S) == (e / S)) ? 1.0 : 0.0;
  __syncthreads();
  // coalesced staging: consecutive threads read consecutive rows of a column
  const int64_t tot = (int64_t)S * L;
  for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
    const int s = (int)(e / L);
    const int64_t r = e - (int64_t)s * L;
    const int64_t col = cols[s];
    Ws[e] = (col >= 0) ? W[r + col * ldw] : 0.0;
  }
  __syncthreads();

  double first_max = 0.0;
  for (int sweep = 0; sweep < inner; ++sweep) {
    if (threadIdx.x == 0) any_rot = 0;
    __syncthreads();
    for (int t = 0; t < S - 1; ++t) {
      if (wave < b) {
        const int sx = circle_player(wave, t, S);
        const int sy = circle_player(S - 1 - wave, t, S);
        if (cols[sx] >= 0 && cols[sy] >= 0) {
          double* wx = Ws + (int64_t)sx * L;
          double* wy = Ws + (int64_t)sy * L;
          double a = 0, bb = 0, g = 0;
          for (int64_t r = lane; r < L; r += 64) {
            const double x = wx[r], y = wy[r];
            a += x * x;
            bb += y * y;
            g += x * y;
          }
          a = wave_sum(a);
          bb = wave_sum(bb);
          g = wave_sum(g);
          double rel = 0.0;
          if (a > 0.0 && bb > 0.0) rel = fabs(g) / (sqrt(a) * sqrt(bb));
          if (sweep == 0 && rel > first_max) first_max = rel;
          if (rel > tol) {
            double cs, sn;
            jacobi_angle(a, bb, g, &cs, &sn);
            for (int64_t r = lane; r < L; r += 64) {
              const double x = wx[r], y = wy[r];
              wx[r] = cs * x - sn * y;
              wy[r] = sn * x + cs * y;
            }
            double* jx = J + sx * S;
            double* jy = J + sy * S;
            for (int r = lane; r < S; r += 64) {
              const double x = jx[r], y = jy[r];
              jx[r] = cs * x - sn * y;
              jy[r] = sn * x + cs * y;
            }
            if (lane == 0) any_rot = 1;
          }
        }
      }
      __syncthreads();
    }
    if (!any_rot) break;
  }
  // write W back
  for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
    const int s = (int)(e / L);
    const int64_t r = e - (int64_t)s * L;
    const int64_t col = cols[s];
    if (col >= 0) W[r + col * ldw] = Ws[e];
  }
  // V(r, cols) <- V(r, cols) * J, one row per thread
  if (Lv > 0) {
    for (int64_t r = threadIdx.x; r < Lv; r += blockDim.x) {
      double row[32];
      for (int s = 0; s < S; ++s) row[s] = (cols[s] >= 0) ? V[r + cols[s] * ldv] : 0.0;
      for (int s = 0; s < S; ++s) {
        if (cols[s] < 0) continue;
        double acc = 0.0;
        for (int u = 0; u < S; ++u) acc += row[u] * J[u + s * S];
        V[r + cols[s] * ldv] = acc;
      }
    }
  }
  __shared__ double red[16];
  for (int off = 32; off > 0; off >>= 1) first_max = fmax(first_max, __shfl_xor(first_max, off, 64));
  if (lane == 0) red[wave] = first_max;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, red[i]);
    atomicMax(offmax, (unsigned long long)__double_as_longlong(m));
  }
}

// Global-memory path for tall W (L too large to stage): each wave rotates its
// column pair in place in W and V.
__global__ void __launch_bounds__(1024)
jacobi_round_global_kernel(int64_t L, int64_t n, double* __restrict__ W, int64_t ldw,
                           double* __restrict__ V, int64_t ldv, int64_t Lv, int b,
                           int nblocks_real, int nplayers, int round, double tol,
                           unsigned long long* __restrict__ offmax) {
  const int pairs = nplayers / 2;
  const int w = blockIdx.x;
  if (w >= pairs) return;
  const int pa = circle_player(w, round, nplayers);
  const int pb = circle_player(nplayers - 1 - w, round, nplayers);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int S = 2 * b;
  auto slot_col = [&](int s) -> int64_t {
    int blk = (s < b) ? pa : pb;
    int off = (s < b) ? s : s - b;
    if (blk >= nblocks_real) return -1;
    int64_t col = (int64_t)blk * b + off;
    return col < n ? col : -1;
  };
  double mymax = 0.0;
  for (int t = 0; t < S - 1; ++t) {
    if (wave < b) {
      const int sx = circle_player(wave, t, S);
      const int sy = circle_player(S - 1 - wave, t, S);
      const int64_t cx = slot_col(sx), cy = slot_col(sy);
      if (cx >= 0 && cy >= 0) {
        double* wx = W + cx * ldw;
        double* wy = W + cy * ldw;
        double a = 0, bb = 0, g = 0;
        for (int64_t r = lane; r < L; r += 64) {
          const double x = wx[r], y = wy[r];
          a += x * x;
          bb += y * y;
          g += x * y;
        }
        a = wave_sum(a);
        bb = wave_sum(bb);
        g = wave_sum(g);
        double rel = 0.0;
        if (a > 0.0 && bb > 0.0) rel = fabs(g) / (sqrt(a) * sqrt(bb));
        if (rel > mymax) mymax = rel;
        if (rel > tol) {
          double cs, sn;
          jacobi_angle(a, bb, g, &cs, &sn);
          for (int64_t r = lane; r < L; r += 64) {
            const double x = wx[r], y = wy[r];
            wx[r] = cs * x - sn * y;
            wy[r] = sn * x + cs * y;
          }
          if (Lv > 0) {
            double* vx = V + cx * ldv;
            double* vy = V + cy * ldv;
            for (int64_t r = lane; r < Lv; r += 64) {
              const double x = vx[r], y = vy[r];
              vx[r] = cs * x - sn * y;
              vy[r] = sn * x + cs * y;
            }
          }
        }
      }
    }
    __syncthreads();
  }
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

// Block-Gram round: for the 2*WB columns of a block pair, H = W^T W (FP64
// MFMA over 64-row chunks staged in LDS), diagonalise H by parallel two-sided
// Jacobi in LDS accumulating J (H <- J^T H J), then W <- W J and V <- V J.
// Equivalent to one-sided Jacobi on the pair's columns, with the O(L) column
// work done once per round instead of once per 2x2 rotation.
typedef double d4g_t __attribute__((ext_vector_type(4)));

template <int WB>
__global__ void __launch_bounds__(256)
jacobi_gram_round_kernel(int64_t L, int64_t n, double* __restrict__ W, int64_t ldw,
                         double* __restrict__ V, int64_t ldv, int64_t Lv, int nblocks_real,
                         int nplayers, int round, double tol, int inner,
                         unsigned long long* __restrict__ offmax) {
  constexpr int S = 2 * WB;            // columns handled by this workgroup
  constexpr int HP = S + 1;            // padded LDS row
  constexpr int CH = 64;               // rows per Gram chunk
  __shared__ double Hs[S * HP];
  __shared__ double Js[S * HP];
  __shared__ double Ws[CH * HP];
  __shared__ double rc[S / 2], rs[S / 2];
  __shared__ int64_t cols[S];
  __shared__ double red[4];
  __shared__ int any_rot;
  const int pairs = nplayers / 2;
  const int w = blockIdx.x;
  if (w >= pairs) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int pa = circle_player(w, round, nplayers);
  const int pb = circle_player(nplayers - 1 - w, round, nplayers);
  if (tid < S) {
    const int blk = (tid < WB) ? pa : pb;
    const int off = (tid < WB) ? tid : tid - WB;
    const int64_t col = (int64_t)blk * WB + off;
    cols[tid] = (blk < nblocks_real && col < n) ? col : -1;
  }
  for (int e = tid; e < S * HP; e += 256) Js[e] = 0.0;
  __syncthreads();
  if (tid < S) Js[tid * HP + tid] = 1.0;

  // ---- H = W^T W on the matrix cores: wave w owns output tile (w/2, w%2)
  constexpr int NT = S / 16;           // tiles per side (2 for S = 32)
  d4g_t acc = {0, 0, 0, 0};
  const int ti = (wave / NT) % NT, tj = wave % NT;
  const bool has_tile = wave < NT * NT;
  for (int64_t r0 = 0; r0 < L; r0 += CH) {
    __syncthreads();
    for (int e = tid; e < CH * S; e += 256) {
      const int c = e / CH, r = e % CH;
      const int64_t col = cols[c];
      Ws[r * HP + c] = (col >= 0 && r0 + r < L) ? W[(r0 + r) + col * ldw] : 0.0;
    }
    __syncthreads();
    if (has_tile) {
#pragma unroll 4
      for (int k = 0; k < CH; k += 4) {
        const int kr = k + (lane >> 4);
        const double a = Ws[kr * HP + 16 * ti + (lane & 15)];
        const double b = Ws[kr * HP + 16 * tj + (lane & 15)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
    }
  }
  if (has_tile) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * ti + 4 * r + (lane >> 4);  // f64 16x16 accumulator layout
      const int j = 16 * tj + (lane & 15);
      Hs[i * HP + j] = acc[r];
    }
  }
  __syncthreads();

  // ---- two-sided Jacobi on H (S x S), rotations accumulated in J.  Step t:
  // (1) every thread of a pair computes the pair's rotation from H (all reads
  // of H happen before the barrier), rotates its entries of rows p, q of H
  // into registers and columns p, q of J; (2) barrier, write rows back; (3)
  // barrier, rotate columns p, q of H with the published angles.
  double first_max = 0.0;
  constexpr int PER = ((S / 2) * S + 255) / 256;  // entries per thread per phase
  for (int sweep = 0; sweep < inner; ++sweep) {
    if (tid == 0) any_rot = 0;
    for (int t = 0; t < S - 1; ++t) {
      double nh[PER][2];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = tid + 256 * k;
        if (e >= (S / 2) * S) break;
        const int i = e / S, j = e % S;
        const int p = circle_player(i, t, S), q = circle_player(S - 1 - i, t, S);
        double cs = 1.0, sn = 0.0;
        if (cols[p] >= 0 && cols[q] >= 0) {
          const double a = Hs[p * HP + p], b = Hs[q * HP + q], g = Hs[p * HP + q];
          double rel = 0.0;
          if (a > 0.0 && b > 0.0) rel = fabs(g) / (sqrt(a) * sqrt(b));
          if (j == 0 && sweep == 0 && rel > first_max) first_max = rel;
          if (rel > tol) jacobi_angle(a, b, g, &cs, &sn);
        }
        if (j == 0) {
          rc[i] = cs;
          rs[i] = sn;
          if (sn != 0.0) any_rot = 1;
        }
        const double hp = Hs[p * HP + j], hq = Hs[q * HP + j];
        nh[k][0] = cs * hp - sn * hq;
        nh[k][1] = sn * hp + cs * hq;
        const double jp = Js[j * HP + p], jq = Js[j * HP + q];
        Js[j * HP + p] = cs * jp - sn * jq;
        Js[j * HP + q] = sn * jp + cs * jq;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = tid + 256 * k;
        if (e >= (S / 2) * S) break;
        const int i = e / S, j = e % S;
        const int p = circle_player(i, t, S), q = circle_player(S - 1 - i, t, S);
        Hs[p * HP + j] = nh[k][0];
        Hs[q * HP + j] = nh[k][1];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = tid + 256 * k;
        if (e >= (S / 2) * S) break;
        const int i = e / S, j = e % S;
        const int p = circle_player(i, t, S), q = circle_player(S - 1 - i, t, S);
        const double cs = rc[i], sn = rs[i];
        const double hp = Hs[j * HP + p], hq = Hs[j * HP + q];
        Hs[j * HP + p] = cs * hp - sn * hq;
        Hs[j * HP + q] = sn * hp + cs * hq;
      }
      __syncthreads();
    }
    if (!any_rot) break;
  }

  // W(:, cols) <- W(:, cols) J ; V(:, cols) <- V(:, cols) J, one row per thread
  const int64_t tot_rows = L + Lv;
  for (int64_t rr = tid; rr < tot_rows; rr += 256) {
    double* __restrict__ X = (rr < L) ? W : V;
    const int64_t ldx = (rr < L) ? ldw : ldv;
    const int64_t r = (rr < L) ? rr : rr - L;
    double row[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int64_t cs_ = cols[s];
      row[s] = (cs_ >= 0) ? X[r + cs_ * ldx] : 0.0;
    }
    for (int s0 = 0; s0 < S; s0 += 4) {
      double o0 = 0, o1 = 0, o2 = 0, o3 = 0;
#pragma unroll
      for (int u = 0; u < S; ++u) {
        const double x = row[u];
        const double* jr = &Js[u * HP + s0];
        o0 += x * jr[0];
        o1 += x * jr[1];
        o2 += x * jr[2];
        o3 += x * jr[3];
      }
      if (cols[s0] >= 0) X[r + cols[s0] * ldx] = o0;
      if (cols[s0 + 1] >= 0) X[r + cols[s0 + 1] * ldx] = o1;
      if (cols[s0 + 2] >= 0) X[r + cols[s0 + 2] * ldx] = o2;
      if (cols[s0 + 3] >= 0) X[r + cols[s0 + 3] * ldx] = o3;
    }
  }

  for (int off = 32; off > 0; off >>= 1) first_max = fmax(first_max, __shfl_xor(first_max, off, 64));
  if (lane == 0) red[wave] = first_max;
  __syncthreads();
  if (tid == 0) {
    double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
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
               int max_sweeps, int* sweeps_out, double rel_tol) {
  if (sweeps_out) *sweeps_out = 0;
  if (n <= 1 || L == 0) return 0;
  const int64_t Lv = V ? n : 0;
  if (V) {
    int blocks = (int)((n * n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(set_identity_kernel, dim3(blocks), dim3(256), 0, c->stream, n, V, ldv);
    RSVD_LAUNCH_CHECK();
  }
  constexpr int WB = 16;
  const int nblocks = (int)((n + WB - 1) / WB);
  const int nplayers = nblocks + (nblocks & 1);
  // Stop when every pair met in a sweep was orthogonal to rel_tol relative
  // (singular values then carry O(rel_tol^2) relative error, vectors
  // O(rel_tol / relative gap)); never tighter than the rounding floor.
  double tol = 2.220446049250313e-16 * sqrt((double)L);
  if (rel_tol > tol) tol = rel_tol;
  DBuf<unsigned long long> off;
  RSVD_TRY(off.alloc(1));
  unsigned long long h = 0;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    RSVD_HIP(hipMemsetAsync(off.p, 0, sizeof(unsigned long long), c->stream));
    for (int r = 0; r < nplayers - 1; ++r) {
      KTimer kt(c, "jacobi_round");
      hipLaunchKernelGGL(jacobi_gram_round_kernel<WB>, dim3(nplayers / 2), dim3(256), 0,
                         c->stream, L, n, W, ldw, V, ldv, Lv, nblocks, nplayers, r, tol, 1,
                         off.p);
      RSVD_LAUNCH_CHECK();
    }
    RSVD_TRY(read_small(c, &h, off.p, sizeof(h)));
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
    // NaN ranks as +Inf and ties go by index: a strict total order, so the
    // ranks are a permutation whatever the values (a NaN would otherwise
    // compare false both ways and two entries would share a rank)
    const double x = isnan(v[j]) ? INFINITY : v[j];
    int64_t r = 0;
    for (int64_t i = 0; i < n; ++i) {
      const double y = isnan(v[i]) ? INFINITY : v[i];
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
