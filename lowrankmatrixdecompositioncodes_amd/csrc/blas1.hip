// blas1.hip -- Gaussian test-matrix generation and the O(mn) streaming helpers
// of the rSVD / QB path (copy, fill, transpose, column scaling, precision
// conversion, FP64 sums of squares for the Frobenius norms).
//
// Gaussian stream (replaces MKL/mvf.c:458-486, which uses VSL MCG31 + ICDF in
// float seeded from time(NULL)): element e of a column-major matrix with R
// rows is lane e%4 of the Philox4x64-10 block with counter (e/4,0,0,0) and key
// (seed,0); the 64-bit word u maps to ((u>>11)+0.5)*2^-53, then through the
// AS241 inverse normal CDF, then is rounded to float and widened.  The
// arithmetic below is written contraction-free so the device produces the
// same bits as a plain IEEE double evaluation on the host.
#include "rsvd_internal.h"

namespace rsvd {

__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t lo0 = 0xD2E7470EE14C6C93ULL * c[0];
    uint64_t hi0 = __umul64hi(0xD2E7470EE14C6C93ULL, c[0]);
    uint64_t lo1 = 0xCA5A826395121157ULL * c[2];
    uint64_t hi1 = __umul64hi(0xCA5A826395121157ULL, c[2]);
    uint64_t n0 = hi1 ^ c[1] ^ k0;
    uint64_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B97F4A7C15ULL;
    k1 += 0xBB67AE8584CAA73BULL;
  }
}

// log(x), x > 0 finite: x = f*2^e, f in [sqrt(1/2), sqrt(2)), 2*atanh series.
__device__ double dev_log(double x) {
#pragma clang fp contract(off)
  int e;
  double f = frexp(x, &e);
  if (f < 0.70710678118654752440) {
    f = f * 2.0;
    e -= 1;
  }
  double t = (f - 1.0) / (f + 1.0);
  double t2 = t * t;
  double s = 1.0 / 27.0;
  for (int j = 25; j >= 1; j -= 2) s = s * t2 + 1.0 / (double)j;
  double lf = 2.0 * t * s;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  return ((double)e * ln2_lo + lf) + (double)e * ln2_hi;
}

// Wichura (1988) AS241 PPND16 inverse standard normal CDF.
__device__ double dev_ndtri(double p) {
#pragma clang fp contract(off)
  double q = p - 0.5, r, x;
  if (fabs(q) <= 0.425) {
    r = 0.180625 - q * q;
    x = q *
        (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
              6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r +
            1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r +
          1.3314166789178437745e+2) * r + 3.3871328727963666080e+0) /
        (((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r +
              3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r +
            5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r +
          4.2313330701600911252e+1) * r + 1.0);
    return x;
  }
  r = (q < 0.0) ? p : 1.0 - p;
  r = sqrt(-dev_log(r));
  if (r <= 5.0) {
    r = r - 1.6;
    x = (((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r +
              2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r +
            3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r +
          4.63033784615654529590e+0) * r + 1.42343711074968357734e+0) /
        (((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r +
              1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r +
            6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r +
          2.05319162663775882187e+0) * r + 1.0);
  } else {
    r = r - 5.0;
    x = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r +
              1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r +
            2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r +
          5.46378491116411436990e+0) * r + 6.65790464350110377720e+0) /
        (((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r +
              1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r +
            1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r +
          5.99832206555887937690e-1) * r + 1.0);
  }
  return (q < 0.0) ? -x : x;
}

__device__ __forceinline__ double gaussian_at(uint64_t seed, uint64_t e) {
  uint64_t c[4] = {e >> 2, 0, 0, 0};
  philox4x64(c, seed, 0);
  uint64_t w = c[e & 3];
  double u = ((double)(w >> 11) + 0.5) * 0x1.0p-53;
  return (double)(float)dev_ndtri(u);
}

template <typename T>
__global__ void gaussian_kernel(uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                                int64_t rows_total, T* __restrict__ out, int64_t ld) {
  int64_t total = rows * cols;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % rows, j = idx / rows;
    uint64_t e = (uint64_t)((col0 + j) * rows_total + i);
    out[i + j * ld] = (T)gaussian_at(seed, e);
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

int gaussian_f64(Context* c, uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                 int64_t rows_total, double* out, int64_t ld) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(gaussian_kernel<double>, dim3(grid_for(rows * cols)), dim3(256), 0,
                     c->stream, seed, rows, cols, col0, rows_total, out, ld);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int gaussian_f32(Context* c, uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                 int64_t rows_total, float* out, int64_t ld) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(gaussian_kernel<float>, dim3(grid_for(rows * cols)), dim3(256), 0,
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

// ---------------------------------------------------------------- 2-D helpers
template <typename S, typename D>
__global__ void copy2d_kernel(int64_t rows, int64_t cols, const S* __restrict__ src, int64_t lds,
                              D* __restrict__ dst, int64_t ldd) {
  int64_t total = rows * cols;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % rows, j = idx / rows;
    dst[i + j * ldd] = (D)src[i + j * lds];
  }
}

int copy2d_f64(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds, double* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<double, double>), dim3(grid_for(rows * cols)), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int copy2d_f32(Context* c, int64_t rows, int64_t cols, const float* src, int64_t lds, float* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<float, float>), dim3(grid_for(rows * cols)), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int f32_to_f64(Context* c, int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<float, double>), dim3(grid_for(rows * cols)), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}
int f64_to_f32(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds, float* dst,
               int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL((copy2d_kernel<double, float>), dim3(grid_for(rows * cols)), dim3(256), 0,
                     c->stream, rows, cols, src, lds, dst, ldd);
  RSVD_LAUNCH_CHECK();
  return 0;
}

__global__ void fill2d_kernel(int64_t rows, int64_t cols, double v, double* __restrict__ dst,
                              int64_t ldd) {
  int64_t total = rows * cols;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % rows, j = idx / rows;
    dst[i + j * ldd] = v;
  }
}
int fill2d_f64(Context* c, int64_t rows, int64_t cols, double v, double* dst, int64_t ldd) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(fill2d_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, c->stream, rows,
                     cols, v, dst, ldd);
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
  int64_t total = rows * cols;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % rows, j = idx / rows;
    double f = invert ? 1.0 / s[j] : s[j];
    a[i + j * lda] *= f;
  }
}
int scale_cols_f64(Context* c, int64_t rows, int64_t cols, double* a, int64_t lda,
                   const double* colscale, int invert) {
  if (rows * cols == 0) return 0;
  hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, c->stream,
                     rows, cols, a, lda, colscale, invert);
  RSVD_LAUNCH_CHECK();
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
