// gemm.hip -- FP64 / FP32 GEMM on CDNA4 matrix cores for the A-streaming passes
// of the rSVD / blocked QB pipeline (Y = A*Omega, Z = A^T*Y, B = Q^T*A, the
// deflation A -= Qp*Bp, U = Q*Uhat, ...), replacing the cblas_dgemm calls of
// MKL/mvf.c:538-561.
//
// Tiling: a 256-thread workgroup (4 waves, 2x2) owns a BM x BN tile of C and
// walks K in BK slices staged through LDS.  Each wave owns a (BM/2) x (BN/2)
// sub-tile built from 16x16 MFMA blocks (v_mfma_f64_16x16x4f64 or
// v_mfma_f32_16x16x4f32).  The MFMA "A" operand is fed from op(B) and the "B"
// operand from op(A), so the accumulator lanes run along M: 16 consecutive
// lanes hold 16 consecutive rows of one column of C and the epilogue writes
// whole 128-byte (f64) segments of column-major C.
//
// Operand lane layout of the 16x16x4 MFMAs (per 64-lane wave):
//   srcA: lane l holds X[i = l%16][k = l/16]
//   srcB: lane l holds Y[k = l/16][j = l%16]
//   dst : lane l, register r holds D[i][j = l%16] with
//         i = 4*(l/16) + r   for v_mfma_f32_16x16x4f32
//         i = 4*r + l/16     for v_mfma_f64_16x16x4f64
// Long-K products with few output tiles are split along K into a workspace
// and summed in a fixed order by a second kernel (deterministic).
#include <climits>

#include <rocblas/rocblas.h>

#include "rsvd_internal.h"

namespace rsvd {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // accumulator row held by lane group g (= lane/16), register r
  static __device__ __forceinline__ int row(int g, int r) { return 4 * r + g; }
};
template <>
struct Mfma<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

constexpr int kThreads = 256;

// C[z] = alpha * op(A)[:, kslice z] * op(B)[kslice z, :] + beta * C
// When sumsq != nullptr, each workgroup writes the sum of squares of the values
// it stored into sumsq[tm + gridDim.x*tn] for its tile (tm, tn) (double).
template <typename T, bool TA, bool TB, int BM, int BN, int BK>
__global__ void __launch_bounds__(kThreads)
gemm_kernel(int64_t M, int64_t N, int64_t K, int64_t kchunk, T alpha, const T* __restrict__ A,
            int64_t lda, const T* __restrict__ B, int64_t ldb, T beta, T* __restrict__ C,
            int64_t ldc, int64_t zstride, double* __restrict__ sumsq, int sym) {
  constexpr int PAD = 4;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int EA = BM * BK / kThreads;
  constexpr int EB = BN * BK / kThreads;
  static_assert(EA * kThreads == BM * BK && EB * kThreads == BN * BK, "tile/threads");
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];

  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;
  // Tile order.  Workgroups are handed to the 8 XCDs round-robin by
  // dispatch order (b % 8), each XCD with its own L2: XCD x takes a
  // contiguous run of tiles, walked with the dimension that has fewer tiles
  // fastest, so the workgroups that run together on one XCD share the large
  // operand's panel (the n / l column tiles of one row block of A run side by
  // side and A leaves HBM once instead of once per column tile).
  int64_t tm, tn;
  {
    const int64_t gx = gridDim.x, gy = gridDim.y, nt = gx * gy;
    const int64_t b = blockIdx.x + gx * (int64_t)blockIdx.y;
    const int64_t xcd = b & 7, j = b >> 3, q = nt >> 3, r = nt & 7;
    const int64_t t = xcd * q + (xcd < r ? xcd : r) + j;
    if (gy <= gx) {
      tn = t % gy;
      tm = t / gy;
    } else {
      tm = t % gx;
      tn = t / gx;
    }
  }
  const int64_t m0 = tm * BM;
  const int64_t n0 = tn * BN;
  // a Gram (sym): tiles wholly below the diagonal are mirrored afterwards
  if (sym && m0 >= n0 + BN) return;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = (kbeg + kchunk < K) ? kbeg + kchunk : K;
  T* Cz = C + (int64_t)blockIdx.z * zstride;

  acc_t acc[TN][TM];
#pragma unroll
  for (int a = 0; a < TN; ++a)
#pragma unroll
    for (int b = 0; b < TM; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  T ra[EA], rb[EB];
  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      int idx = tid + kThreads * e;
      int mm, kk;
      if (!TA) { mm = idx % BM; kk = idx / BM; }
      else { kk = idx % BK; mm = idx / BK; }
      int64_t gm = m0 + mm, gk = k0 + kk;
      T v = T(0);
      if (gm < M && gk < kend) v = TA ? A[gk + gm * lda] : A[gm + gk * lda];
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      int idx = tid + kThreads * e;
      int nn, kk;
      if (!TB) { kk = idx % BK; nn = idx / BK; }
      else { nn = idx % BN; kk = idx / BN; }
      int64_t gn = n0 + nn, gk = k0 + kk;
      T v = T(0);
      if (gn < N && gk < kend) v = TB ? B[gn + gk * ldb] : B[gk + gn * ldb];
      rb[e] = v;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      int idx = tid + kThreads * e;
      int mm, kk;
      if (!TA) { mm = idx % BM; kk = idx / BM; }
      else { kk = idx % BK; mm = idx / BK; }
      As[kk][mm] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      int idx = tid + kThreads * e;
      int nn, kk;
      if (!TB) { kk = idx % BK; nn = idx / BK; }
      else { nn = idx % BN; kk = idx / BN; }
      Bs[kk][nn] = rb[e];
    }
  };

  const int li = lane & 15;
  const int lk = lane >> 4;
  if (kbeg < kend) load_tiles(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (k0 + BK < kend) load_tiles(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + lk;
      T av[TN], bv[TM];
#pragma unroll
      for (int a = 0; a < TN; ++a) av[a] = Bs[kr][wn0 + a * 16 + li];
#pragma unroll
      for (int b = 0; b < TM; ++b) bv[b] = As[kr][wm0 + b * 16 + li];
#pragma unroll
      for (int a = 0; a < TN; ++a)
#pragma unroll
        for (int b = 0; b < TM; ++b) acc[a][b] = Mfma<T>::mma(av[a], bv[b], acc[a][b]);
    }
  }

  double ss = 0.0;
#pragma unroll
  for (int a = 0; a < TN; ++a) {
#pragma unroll
    for (int b = 0; b < TM; ++b) {
      const int64_t gm = m0 + wm0 + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gn = n0 + wn0 + a * 16 + Mfma<T>::row(lk, r);
        if (gm < M && gn < N) {
          T* cp = Cz + gm + gn * ldc;
          T v = alpha * acc[a][b][r];
          if (beta != T(0)) v += beta * (*cp);
          *cp = v;
          ss += (double)v * (double)v;
        }
      }
    }
  }
  if (sumsq) {
    __shared__ double red[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
    if (lane == 0) red[wave] = ss;
    __syncthreads();
    if (tid == 0) {
      double t = 0;
      for (int w = 0; w < kThreads / 64; ++w) t += red[w];
      sumsq[tm + (int64_t)gridDim.x * tn] = t;
    }
  }
}

// C = sum_z W[z] (+ beta*C), W[z] is M x N with leading dimension M.
template <typename T>
__global__ void splitk_reduce_kernel(int64_t M, int64_t N, int splits, const T* __restrict__ W,
                                     T beta, T* __restrict__ C, int64_t ldc) {
  int64_t total = M * N;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % M, j = idx / M;
    T s = T(0);
    // loads issued ahead of the (in-order) adds
#pragma unroll 8
    for (int z = 0; z < splits; ++z) s += W[(int64_t)z * total + idx];
    T* cp = C + i + j * ldc;
    if (beta != T(0)) s += beta * (*cp);
    *cp = s;
  }
}

// Fixed-order sum of n doubles into *out (single workgroup).
__global__ void sum_partials_kernel(int64_t n, const double* __restrict__ part,
                                    double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// C(i, j) <- C(j, i) for i > j: the strictly lower triangle from the upper one
template <typename T>
__global__ void mirror_lower_kernel(int64_t n, T* __restrict__ C, int64_t ldc) {
  const int64_t total = n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx % n, j = idx / n;
    if (i > j) C[i + j * ldc] = C[j + i * ldc];
  }
}

template <typename T, int BM, int BN, int BK>
static int launch_gemm(Context* c, char ta, char tb, int64_t M, int64_t N, int64_t K, T alpha,
                       const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C,
                       int64_t ldc, double* sumsq_partials, const char* tag) {
  const int64_t gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
  if (gy > 65535) return set_error("gemm: N=%lld too large for the grid", (long long)N);
  // A^T A with beta = 0: only the tiles that reach the diagonal or above are
  // formed, the strictly lower triangle is copied from the upper one afterwards
  const bool tA0 = (ta == 'T' || ta == 't'), tB0 = (tb == 'T' || tb == 't');
  const bool sym = A == B && lda == ldb && tA0 && !tB0 && M == N && beta == T(0) &&
                   !sumsq_partials;
  int64_t tiles = gx * gy;
  if (sym) {
    tiles = 0;
    for (int64_t bx = 0; bx < gx; ++bx)
      for (int64_t by = 0; by < gy; ++by) tiles += (bx * BM < by * BN + BN) ? 1 : 0;
  }
  // split K when the output has too few tiles to fill the CUs
  int splits = 1;
  const int64_t target = 2LL * (c->num_cus > 0 ? c->num_cus : 256);
  if (!sumsq_partials && tiles < target && K >= 4 * 256) {
    int64_t maxs = K / 256;
    const int64_t cap = BN <= 32 ? 128 : 64;  // skinny tiles: more, shorter slices
    if (maxs > cap) maxs = cap;
    splits = pick_k_slices(tiles, target, maxs);
  }
  int64_t kchunk = K;
  if (splits > 1) {
    kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + BK - 1) / BK * BK;
    splits = (int)((K + kchunk - 1) / kchunk);
  }
  DBuf<T> work;
  T* Cout = C;
  int64_t ldo = ldc, zstride = 0;
  T a_eff = alpha, b_eff = beta;
  if (splits > 1) {
    RSVD_TRY(work.alloc((size_t)(M * N * splits)));
    Cout = work.p;
    ldo = M;
    zstride = M * N;
    b_eff = T(0);
  }
  dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)splits);
  {
    KTimer kt(c, tag);
    const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
#define RSVD_GEMM_LAUNCH(TA_, TB_)                                                          \
  hipLaunchKernelGGL((gemm_kernel<T, TA_, TB_, BM, BN, BK>), grid, dim3(kThreads), 0,         \
                     c->stream, M, N, K, kchunk, a_eff, A, lda, B, ldb, b_eff, Cout, ldo,    \
                     zstride, sumsq_partials, sym ? 1 : 0)
    if (!tA && !tB) RSVD_GEMM_LAUNCH(false, false);
    else if (tA && !tB) RSVD_GEMM_LAUNCH(true, false);
    else if (!tA && tB) RSVD_GEMM_LAUNCH(false, true);
    else RSVD_GEMM_LAUNCH(true, true);
#undef RSVD_GEMM_LAUNCH
    RSVD_LAUNCH_CHECK();
  }
  if (splits > 1) {
    int64_t total = M * N;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3(blocks), dim3(256), 0, c->stream, M, N,
                       splits, (const T*)work.p, beta, C, ldc);
    RSVD_LAUNCH_CHECK();
  }
  if (sym) {
    const int64_t total = M * M;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(mirror_lower_kernel<T>, dim3(blocks), dim3(256), 0, c->stream, M, C, ldc);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

static int check_dims(int64_t M, int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc,
                      char ta, char tb) {
  if (M < 0 || N < 0 || K < 0) return set_error("gemm: negative dimension");
  const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
  if (lda < (tA ? (K > 0 ? K : 1) : (M > 0 ? M : 1)))
    return set_error("gemm: lda=%lld too small", (long long)lda);
  if (ldb < (tB ? (N > 0 ? N : 1) : (K > 0 ? K : 1)))
    return set_error("gemm: ldb=%lld too small", (long long)ldb);
  if (ldc < (M > 0 ? M : 1)) return set_error("gemm: ldc=%lld too small", (long long)ldc);
  return 0;
}

template <typename T>
__global__ void scale_kernel(int64_t M, int64_t N, T beta, T* C, int64_t ldc) {
  int64_t total = M * N;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % M, j = idx / M;
    C[i + j * ldc] = beta == T(0) ? T(0) : beta * C[i + j * ldc];
  }
}

// With rsvd_set_fp64_vendor_gemm(1), the A-streaming FP64 products of the
// plain path (Y = A X, Z = A^T Y, B^T = A^T Q: K in the thousands, 10+ GFLOP)
// run on rocBLAS's dgemm instead of gemm_kernel: a vendor reference the
// plain method can be timed against (off by default).  Never inside an
// allocation-free call (rocBLAS takes its own workspace), nor for dimensions
// beyond rocBLAS's 32-bit integers.
static int vendor_f64(Context* c, int64_t M, int64_t N, int64_t K, int64_t lda, int64_t ldb,
                      int64_t ldc, bool* use) {
  *use = false;
  if (!c->fp64_vendor || c->scratch_mode) return 0;
  if (2.0 * (double)M * (double)N * (double)K < 1e10 || M * N < (int64_t(1) << 22) || K < 4096)
    return 0;
  for (int64_t v : {M, N, K, lda, ldb, ldc})
    if (v > INT_MAX) return 0;
  if (!c->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      return set_error("rocblas_create_handle failed");
    if (rocblas_set_stream(h, c->stream) != rocblas_status_success) {
      rocblas_destroy_handle(h);
      return set_error("rocblas_set_stream failed");
    }
    c->blas = h;
  }
  *use = true;
  return 0;
}

template <typename T>
static int gemm_any(Context* c, char ta, char tb, int64_t M, int64_t N, int64_t K, T alpha,
                    const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc,
                    const char* tag) {
  RSVD_TRY(check_dims(M, N, K, lda, ldb, ldc, ta, tb));
  if (M == 0 || N == 0) return 0;
  if constexpr (sizeof(T) == 8) {
    bool vendor = false;
    if (K > 0) RSVD_TRY(vendor_f64(c, M, N, K, lda, ldb, ldc, &vendor));
    if (vendor) {
      KTimer kt(c, tag);
      const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
      const double a = alpha, b = beta;
      const rocblas_status st = rocblas_dgemm(
          (rocblas_handle)c->blas, tA ? rocblas_operation_transpose : rocblas_operation_none,
          tB ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)M,
          (rocblas_int)N, (rocblas_int)K, &a, A, (rocblas_int)lda, B, (rocblas_int)ldb, &b, C,
          (rocblas_int)ldc);
      if (st != rocblas_status_success)
        return set_error("rocblas_dgemm failed: %s", rocblas_status_to_string(st));
      return 0;
    }
  }
  if (K == 0) {
    int blocks = (int)((M * N + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks), dim3(256), 0, c->stream, M, N, beta, C,
                       ldc);
    RSVD_LAUNCH_CHECK();
    return 0;
  }
  // N <= 32 (a product with a panel of 32 or fewer columns): a 128 x 32 tile,
  // no half-empty 128 x 64 tile
  if (N <= 32)
    return launch_gemm<T, 128, 32, 16>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                       nullptr, tag);
  return launch_gemm<T, 128, 64, 16>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                     nullptr, tag);
}

int gemm_f64(Context* c, char ta, char tb, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
             int64_t ldc) {
  return gemm_any<double>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                          "gemm_f64");
}

int gemm_f64_tag(Context* c, const char* tag, char ta, char tb, int64_t M, int64_t N, int64_t K,
                 double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                 double beta, double* C, int64_t ldc) {
  return gemm_any<double>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tag);
}

int gemm_f32(Context* c, char ta, char tb, int64_t M, int64_t N, int64_t K, float alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C,
             int64_t ldc) {
  return gemm_any<float>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, "gemm_f32");
}

int gemm_f64_deflate(Context* c, int64_t M, int64_t N, int64_t K, const double* Q, int64_t ldq,
                     const double* B, int64_t ldb, double* C, int64_t ldc, double* sumsq) {
  RSVD_TRY(check_dims(M, N, K, ldq, ldb, ldc, 'N', 'N'));
  if (M == 0 || N == 0) {
    RSVD_HIP(hipMemsetAsync(sumsq, 0, sizeof(double), c->stream));
    return 0;
  }
  const int64_t gx = (M + 127) / 128, gy = (N + 63) / 64;
  DBuf<double> part;
  RSVD_TRY(part.alloc((size_t)(gx * gy)));
  RSVD_TRY((launch_gemm<double, 128, 64, 16>(c, 'N', 'N', M, N, K > 0 ? K : 1, -1.0, Q, ldq, B,
                                             ldb, 1.0, C, ldc, part.p, "gemm_f64_deflate")));
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, c->stream, gx * gy,
                     (const double*)part.p, sumsq);
  RSVD_LAUNCH_CHECK();
  return 0;
}

}  // namespace rsvd
