// blas1.hip -- Gaussian test-matrix generation and the O(mn) streaming helpers
// of the rSVD / QB path (copy, fill, transpose, column scaling, precision
// conversion, FP64 sums of squares for the Frobenius norms).
//
// Gaussian stream (replaces MKL/mvf.c:458-486, which uses VSL MCG31 + ICDF in
// float seeded from time(NULL)): element (i, j) of a column-major matrix with R
// rows is element j*R + i of the counter-based stream of rng.h.
#include "rng.h"
#include "rsvd_internal.h"

namespace rsvd {

// A thread takes one Philox block (four consecutive elements of the stream)
// of a column's element range [e0, e0 + rows), e0 = (col0 + j) rows_total,
// and writes the entries it covers: one block per four elements instead of
// one per element.  Grid: x over the qn = rows / 4 + 2 blocks that can touch
// a column, y over columns.
template <typename T>
__global__ void gaussian_kernel(uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                                int64_t rows_total, T* __restrict__ out, int64_t ld) {
  const int64_t qn = rows / 4 + 2;
  for (int64_t j = blockIdx.y; j < cols; j += gridDim.y) {
    const int64_t e0 = (col0 + j) * rows_total;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < qn;
         q += (int64_t)gridDim.x * blockDim.x) {
      const int64_t b = (e0 >> 2) + q;
      if (4 * b >= e0 + rows) continue;
      double g[4];
      gaussian_block4(seed, (uint64_t)b, g);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int64_t i = 4 * b + w - e0;
        if (i >= 0 && i < rows) out[i + j * ld] = (T)g[w];
      }
    }
  }
}

// A(i,j) += s * element (j*rows + i) of the stream for `seed`.
__global__ void add_gaussian_kernel(uint64_t seed, int64_t rows, int64_t cols, double s,
                                    double* __restrict__ A, int64_t lda) {
  int64_t total = rows * cols;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % rows, j = idx / rows;
    A[i + j * lda] += s * gaussian_at(seed, (uint64_t)idx);
  }
}

static inline int grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

// Grid of the elementwise 2-D kernels: y over columns, x over rows (about
// four elements a thread per column), so no thread divides by the row count.
static inline dim3 grid2d(int64_t rows, int64_t cols) {
  int64_t gx = (rows + 1023) / 1024;
  if (gx > 64) gx = 64;
  if (gx < 1) gx = 1;
  int64_t gy = cols;
  if (gy > 4096) gy = 4096;
  if (gy < 1) gy = 1;
  return dim3((unsigned)gx, (unsigned)gy);
}

#define RSVD_FOR_2D(rows, cols)                                                        \
  for (int64_t j = blockIdx.y; j < (cols); j += gridDim.y)                             \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (rows);       \
         i += (int64_t)gridDim.x * blockDim.x)

int gaussian_f64(Context* c, uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                 int64_t rows_total, double* out, int64_t ld) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(gaussian_kernel<double>, grid2d(rows / 4 + 2, cols), dim3(256), 0,
                     c->stream, seed, rows, cols, col0, rows_total, out, ld);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int gaussian_f32(Context* c, uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                 int64_t rows_total, float* out, int64_t ld) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(gaussian_kernel<float>, grid2d(rows / 4 + 2, cols), dim3(256), 0,
                     c->stream, seed, rows, cols, col0, rows_total, out, ld);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int add_gaussian_f64(Context* c, uint64_t seed, int64_t rows, int64_t cols, double s, double* A,
                     int64_t lda) {
  if (rows * cols == 0) return 0;
  int64_t b = (rows * cols + 255) / 256;
  hipLaunchKernelGGL(add_gaussian_kernel, dim3((unsigned)(b > 16384 ? 16384 : b)), dim3(256), 0,
                     c->stream, seed, rows, cols, s, A, lda);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------- streaming copy
// dst[0:n) = src[0:n) with 16-byte loads and stores: n * 16 bytes read and
// written, nothing else -- the reference point for HBM byte counters.
__global__ void stream_copy_kernel(int64_t n, const float4* __restrict__ src,
                                   float4* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int stream_copy(Context* c, const void* src, void* dst, int64_t bytes) {
  if (bytes == 0) return 0;
  if (bytes % 16 || (reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16)
    return set_error("stream copy: size and pointers must be multiples of 16 bytes");
  const int64_t n = bytes / 16;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, n,
                     static_cast<const float4*>(src), static_cast<float4*>(dst));
  RSVD_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------- 2-D helpers
template <typename S, typename D>
__global__ void copy2d_kernel(int64_t rows, int64_t cols, const S* __restrict__ src, int64_t lds,
                              D* __restrict__ dst, int64_t ldd) {
  RSVD_FOR_2D(rows, cols) dst[i + j * ldd] = (D)src[i + j * lds];
}

int copy2d_f64(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds, double* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<double, double>), grid2d(rows, cols), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int copy2d_f32(Context* c, int64_t rows, int64_t cols, const float* src, int64_t lds, float* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<float, float>), grid2d(rows, cols), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int f32_to_f64(Context* c, int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<float, double>), grid2d(rows, cols), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int f64_to_f32(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds, float* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<double, float>), grid2d(rows, cols), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}

__global__ void fill2d_kernel(int64_t rows, int64_t cols, double v, double* __restrict__ dst,
                              int64_t ldd) {
  RSVD_FOR_2D(rows, cols) dst[i + j * ldd] = v;
}
int fill2d_f64(Context* c, int64_t rows, int64_t cols, double v, double* dst, int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(fill2d_kernel, grid2d(rows, cols), dim3(256), 0, c->stream, rows, cols, v,
                     dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// 32x32 tiles through LDS.
__global__ void transpose_kernel(int64_t rows, int64_t cols, const double* __restrict__ src,
                                 int64_t lds, double* __restrict__ dst, int64_t ldd) {
  __shared__ double t[32][33];
  int64_t i0 = (int64_t)blockIdx.x * 32, j0 = (int64_t)blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: ty 0..7
  for (int r = ty; r < 32; r += 8) {
    int64_t i = i0 + tx, j = j0 + r;
    t[r][tx] = (i < rows && j < cols) ? src[i + j * lds] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int64_t jj = j0 + tx, ii = i0 + r;  // dst(jj, ii) = src(ii, jj)
    if (jj < cols && ii < rows) dst[jj + ii * ldd] = t[tx][r];
  }
}
int transpose_f64(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds,
                  double* dst, int64_t ldd) {
  if (rows * cols == 0) return 0;
  dim3 g((unsigned)((rows + 31) / 32), (unsigned)((cols + 31) / 32));
  hipLaunchKernelGGL(transpose_kernel, g, dim3(256), 0, c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}

__global__ void scale_cols_kernel(int64_t rows, int64_t cols, double* __restrict__ a, int64_t lda,
                                  const double* __restrict__ s, int invert) {
  for (int64_t j = blockIdx.y; j < cols; j += gridDim.y) {
    const double f = invert ? 1.0 / s[j] : s[j];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * blockDim.x)
      a[i + j * lda] *= f;
  }
}
int scale_cols_f64(Context* c, int64_t rows, int64_t cols, double* a, int64_t lda,
                   const double* colscale, int invert) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(scale_cols_kernel, grid2d(rows, cols), dim3(256), 0, c->stream, rows, cols,
                     a, lda, colscale, invert);
  RSVD_LAUNCH_CHECK();
  return 0;
}

__global__ void fill_int_kernel(int* __restrict__ p, int v) {
  if (threadIdx.x == 0) *p = v;
}

// *flag |= 1 when any of x[0 .. n) is NaN or infinite.
__global__ void nonfinite_kernel(int64_t n, const double* __restrict__ x, int* __restrict__ flag) {
  bool bad = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    bad = bad || !isfinite(x[i]);
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

int check_finite_f64(Context* c, const double* x, int64_t n, const char* what) {
  if (n <= 0) return 0;
  DBuf<int> flag;
  RSVD_TRY(flag.alloc(1));
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, c->stream, flag.p, 0);
  RSVD_LAUNCH_CHECK();
  hipLaunchKernelGGL(nonfinite_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, n, x, flag.p);
  RSVD_LAUNCH_CHECK();
  int h = 0;
  RSVD_TRY(read_small(c, &h, flag.p, sizeof h));
  if (h)
    return set_error("%s: non-finite singular values (the input holds NaN or Inf); the outputs "
                     "are not usable", what);
  return 0;
}

// Sum of squares, deterministic: fixed grid of partial sums, then one block.
template <bool DIFF>
__global__ void sumsq_stage1(int64_t rows, int64_t cols, const double* __restrict__ a,
                             int64_t lda, const double* __restrict__ b, int64_t ldb,
                             double* __restrict__ part) {
  __shared__ double red[256];
  int64_t total = rows * cols;
  double s = 0;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % rows, j = idx / rows;
    double v = a[i + j * lda];
    if (DIFF) v -= b[i + j * ldb];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void sumsq_stage2(int n, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <bool DIFF>
static int sumsq_impl(Context* c, int64_t rows, int64_t cols, const double* a, int64_t lda,
                      const double* b, int64_t ldb, double* out) {
  const int nb = 1024;
  DBuf<double> part;
  RSVD_TRY(part.alloc(nb));
  hipLaunchKernelGGL(sumsq_stage1<DIFF>, dim3(nb), dim3(256), 0, c->stream, rows, cols, a, lda, b,
                     ldb, part.p);
  RSVD_LAUNCH_CHECK();
  hipLaunchKernelGGL(sumsq_stage2, dim3(1), dim3(256), 0, c->stream, nb, (const double*)part.p,
                     out);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int sumsq_f64(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds,
              double* out_scalar) {
  return sumsq_impl<false>(c, rows, cols, src, lds, nullptr, 0, out_scalar);
}
int sumsq_diff_f64(Context* c, int64_t rows, int64_t cols, const double* a, int64_t lda,
                   const double* b, int64_t ldb, double* out_scalar) {
  return sumsq_impl<true>(c, rows, cols, a, lda, b, ldb, out_scalar);
}

}  // namespace rsvd

namespace rsvd {

__global__ void add_diag_kernel(int64_t n, double v, double* __restrict__ a, int64_t lda) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i + i * lda] += v;
}
int add_diag_f64(Context* c, int64_t n, double v, double* a, int64_t lda) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     n, v, a, lda);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// out[j] = X(:,j) . Y(:,j), one wave per column.
__global__ void coldot_kernel(int64_t L, int64_t n, const double* __restrict__ X, int64_t ldx,
                              const double* __restrict__ Y, int64_t ldy, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;
  double s = 0;
  for (int64_t r = lane; r < L; r += 64) s += X[r + j * ldx] * Y[r + j * ldy];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[j] = s;
}
int coldot_f64(Context* c, int64_t L, int64_t n, const double* X, int64_t ldx, const double* Y,
               int64_t ldy, double* out) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(coldot_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c->stream, L, n,
                     X, ldx, Y, ldy, out);
  RSVD_LAUNCH_CHECK();
  return 0;
}

__global__ void add_scalar_kernel(int64_t n, double v, double* __restrict__ a) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] += v;
}
int add_scalar_f64(Context* c, int64_t n, double v, double* a) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(add_scalar_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     c->stream, n, v, a);
  RSVD_LAUNCH_CHECK();
  return 0;
}

}  // namespace rsvd
