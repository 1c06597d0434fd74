// binio.hip -- the reference's binary matrix files, read in bulk.
//
// File layout (all variants): two integers rows, columns, then the entries row
// after row.
//   format 0: int32 header, double entries (MKL/mvf.c:77-99 loader, :110-133
//             writer; make_matrix_binary.m:32-41, read_matrix_binary.m:6-15);
//   format 1: int64 header, double entries
//             (multi_core_mkl_code_64bit/matrix_vector_functions_intel_mkl.c:78-101);
//   format 2: int64 header, float entries
//             (oneapi_code/matrix_vector_functions_one_api.c:71-95).
// The reference reads and writes one element per fread / fwrite.  Here the
// host `mat` loader reads blocks of rows and transposes them into column-major
// storage; the device loader streams blocks of rows through two pinned host
// buffers into HBM (the file read of one block overlaps the copy and the
// on-device transpose of the previous one), converting to double or float.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rsvd_internal.h"
#include "rsvdpack.h"
#include "rsvdpack_ext.h"

namespace rsvd {
namespace {

constexpr size_t kBlockBytes = size_t(32) << 20;  // bytes of file rows per block

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

size_t entry_bytes(int format) { return format == 2 ? 4 : 8; }

int read_header(FILE* fp, const char* fname, int format, int64_t* m, int64_t* n) {
  if (format < 0 || format > 2) return set_error("binary file: unknown format %d", format);
  if (format == 0) {
    int32_t h[2];
    if (fread(h, sizeof(int32_t), 2, fp) != 2) return set_error("%s: no header", fname);
    *m = h[0];
    *n = h[1];
  } else {
    int64_t h[2];
    if (fread(h, sizeof(int64_t), 2, fp) != 2) return set_error("%s: no header", fname);
    *m = h[0];
    *n = h[1];
  }
  if (*m < 0 || *n < 0) return set_error("%s: bad header (%lld x %lld)", fname, (long long)*m,
                                         (long long)*n);
  return 0;
}

int write_header(FILE* fp, int format, int64_t m, int64_t n) {
  if (format == 0) {
    if (m > INT32_MAX || n > INT32_MAX) return set_error("format 0 holds 32-bit dimensions");
    const int32_t h[2] = {(int32_t)m, (int32_t)n};
    if (fwrite(h, sizeof(int32_t), 2, fp) != 2) return set_error("binary file: write failed");
  } else {
    const int64_t h[2] = {m, n};
    if (fwrite(h, sizeof(int64_t), 2, fp) != 2) return set_error("binary file: write failed");
  }
  return 0;
}

int64_t rows_per_block(int64_t m, int64_t n, size_t eb) {
  const size_t row = (size_t)(n > 0 ? n : 1) * eb;
  int64_t r = (int64_t)(kBlockBytes / row);
  if (r < 1) r = 1;
  if (r > (int64_t(1) << 20)) r = int64_t(1) << 20;  // grid.y of the block transpose
  return r < m ? r : (m > 0 ? m : 1);
}

// out(i0 + i, j) <- in(i, j) for a row-major block in (rows x n, row stride
// n), out column-major with leading dimension ld; 32 x 32 tiles through LDS so
// that both the reads (along a row) and the writes (along a column) coalesce.
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256)
rows_to_cols_kernel(int64_t rows, int64_t n, const Tin* __restrict__ in, Tout* __restrict__ out,
                    int64_t ld) {
  __shared__ Tout tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.y * 32, j0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = 0; k < 32; k += 8) {
    const int64_t i = i0 + ty + k, j = j0 + tx;
    if (i < rows && j < n) tile[ty + k][tx] = (Tout)in[i * n + j];
  }
  __syncthreads();
  for (int k = 0; k < 32; k += 8) {
    const int64_t i = i0 + tx, j = j0 + ty + k;
    if (i < rows && j < n) out[i + j * ld] = tile[tx][ty + k];
  }
}

// The reverse: a row-major block out (rows x n) from column-major in(i0 + i, j).
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256)
cols_to_rows_kernel(int64_t rows, int64_t n, const Tin* __restrict__ in, int64_t ld,
                    Tout* __restrict__ out) {
  __shared__ Tout tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.y * 32, j0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = 0; k < 32; k += 8) {
    const int64_t i = i0 + tx, j = j0 + ty + k;
    if (i < rows && j < n) tile[tx][ty + k] = (Tout)in[i + j * ld];
  }
  __syncthreads();
  for (int k = 0; k < 32; k += 8) {
    const int64_t i = i0 + ty + k, j = j0 + tx;
    if (i < rows && j < n) out[i * n + j] = tile[ty + k][tx];
  }
}

dim3 tile_grid(int64_t rows, int64_t n) {
  return dim3((unsigned)((n + 31) / 32), (unsigned)((rows + 31) / 32));
}

int launch_rows_to_cols(Context* c, bool in_f32, bool out_f32, int64_t rows, int64_t n,
                        const void* in, void* out, int64_t ld) {
  const dim3 g = tile_grid(rows, n);
  if (g.y > 65535u) return set_error("binary load: block of %lld rows too tall", (long long)rows);
  if (in_f32 && out_f32)
    hipLaunchKernelGGL((rows_to_cols_kernel<float, float>), g, dim3(256), 0, c->stream, rows, n,
                       (const float*)in, (float*)out, ld);
  else if (in_f32)
    hipLaunchKernelGGL((rows_to_cols_kernel<float, double>), g, dim3(256), 0, c->stream, rows, n,
                       (const float*)in, (double*)out, ld);
  else if (out_f32)
    hipLaunchKernelGGL((rows_to_cols_kernel<double, float>), g, dim3(256), 0, c->stream, rows, n,
                       (const double*)in, (float*)out, ld);
  else
    hipLaunchKernelGGL((rows_to_cols_kernel<double, double>), g, dim3(256), 0, c->stream, rows,
                       n, (const double*)in, (double*)out, ld);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int launch_cols_to_rows(Context* c, bool in_f32, bool out_f32, int64_t rows, int64_t n,
                        const void* in, int64_t ld, void* out) {
  const dim3 g = tile_grid(rows, n);
  if (g.y > 65535u) return set_error("binary write: block of %lld rows too tall", (long long)rows);
  if (in_f32 && out_f32)
    hipLaunchKernelGGL((cols_to_rows_kernel<float, float>), g, dim3(256), 0, c->stream, rows, n,
                       (const float*)in, ld, (float*)out);
  else if (in_f32)
    hipLaunchKernelGGL((cols_to_rows_kernel<float, double>), g, dim3(256), 0, c->stream, rows, n,
                       (const float*)in, ld, (double*)out);
  else if (out_f32)
    hipLaunchKernelGGL((cols_to_rows_kernel<double, float>), g, dim3(256), 0, c->stream, rows, n,
                       (const double*)in, ld, (float*)out);
  else
    hipLaunchKernelGGL((cols_to_rows_kernel<double, double>), g, dim3(256), 0, c->stream, rows,
                       n, (const double*)in, ld, (double*)out);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// Pinned host staging of the device loader and writer: two buffers, each
// holding one block of rows; freed after the stream's copies out of them end.
struct PinnedPair {
  hipStream_t stream;
  void* p[2] = {nullptr, nullptr};
  explicit PinnedPair(hipStream_t s) : stream(s) {}
  ~PinnedPair() {
    if (p[0] || p[1]) (void)hipStreamSynchronize(stream);
    for (void* q : p)
      if (q) (void)hipHostFree(q);
  }
  int alloc(size_t bytes) {
    for (int i = 0; i < 2; ++i) RSVD_HIP(hipHostMalloc(&p[i], bytes, hipHostMallocDefault));
    return 0;
  }
};

struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~EventPair() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
  int create() {
    for (int i = 0; i < 2; ++i) RSVD_HIP(hipEventCreateWithFlags(&e[i], hipEventDisableTiming));
    return 0;
  }
};

}  // namespace
}  // namespace rsvd

using namespace rsvd;

extern "C" {

// MKL/mvf.c:77-99 (format 0), read by blocks of rows.
mat* matrix_load_from_binary_file(char* fname) {
  clear_error();
  FileCloser fc{fopen(fname, "rb")};
  if (!fc.f) {
    set_error("cannot open %s", fname);
    return nullptr;
  }
  int64_t m = 0, n = 0;
  if (read_header(fc.f, fname, 0, &m, &n)) return nullptr;
  mat* M = matrix_new((int)m, (int)n);
  if (!M) {
    set_error("%s: out of host memory", fname);
    return nullptr;
  }
  const int64_t R = rows_per_block(m, n, 8);
  std::vector<double> blk((size_t)(R * (n > 0 ? n : 1)));
  for (int64_t r0 = 0; r0 < m; r0 += R) {
    const int64_t rows = (m - r0 < R) ? m - r0 : R;
    if (fread(blk.data(), sizeof(double), (size_t)(rows * n), fc.f) != (size_t)(rows * n)) {
      matrix_delete(M);
      set_error("%s: truncated in rows %lld..%lld", fname, (long long)r0,
                (long long)(r0 + rows - 1));
      return nullptr;
    }
    // transpose the block by 64-column strips (the strip of each row stays
    // in cache while its column segments are written)
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
      const int64_t j1 = (n - j0 < 64) ? n : j0 + 64;
      for (int64_t i = 0; i < rows; ++i) {
        const double* src = blk.data() + (size_t)(i * n);
        for (int64_t j = j0; j < j1; ++j) M->d[(size_t)j * (size_t)m + (size_t)(r0 + i)] = src[j];
      }
    }
  }
  return M;
}

// MKL/mvf.c:110-133 (format 0), written by blocks of rows.
void matrix_write_to_binary_file(mat* M, char* fname) {
  clear_error();
  FileCloser fc{fopen(fname, "wb")};
  if (!fc.f) {
    set_error("cannot open %s", fname);
    return;
  }
  const int64_t m = M->nrows, n = M->ncols;
  if (write_header(fc.f, 0, m, n)) return;
  const int64_t R = rows_per_block(m, n, 8);
  std::vector<double> blk((size_t)(R * (n > 0 ? n : 1)));
  for (int64_t r0 = 0; r0 < m; r0 += R) {
    const int64_t rows = (m - r0 < R) ? m - r0 : R;
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
      const int64_t j1 = (n - j0 < 64) ? n : j0 + 64;
      for (int64_t i = 0; i < rows; ++i) {
        double* dst = blk.data() + (size_t)(i * n);
        for (int64_t j = j0; j < j1; ++j) dst[j] = M->d[(size_t)j * (size_t)m + (size_t)(r0 + i)];
      }
    }
    if (fwrite(blk.data(), sizeof(double), (size_t)(rows * n), fc.f) != (size_t)(rows * n)) {
      set_error("%s: write failed", fname);
      return;
    }
  }
}

int rsvd_binary_file_shape(const char* fname, int format, long long* m, long long* n) {
  clear_error();
  if (!fname || !m || !n) return set_error("rsvd_binary_file_shape: NULL argument");
  FileCloser fc{fopen(fname, "rb")};
  if (!fc.f) return set_error("cannot open %s", fname);
  int64_t mm = 0, nn = 0;
  RSVD_TRY(read_header(fc.f, fname, format, &mm, &nn));
  *m = mm;
  *n = nn;
  return 0;
}

int rsvd_dev_load_binary(const char* fname, int format, int dtype, void* dA, long long lda) {
  clear_error();
  Context* c = nullptr;
  RSVD_TRY(ctx_get(&c));
  std::unique_lock<std::recursive_mutex> lk(c->mu);
  RSVD_TRY(order_after_null_stream(c));
  if (!fname || !dA) return set_error("rsvd_dev_load_binary: NULL argument");
  if (dtype != 0 && dtype != 1) return set_error("rsvd_dev_load_binary: dtype %d", dtype);
  FileCloser fc{fopen(fname, "rb")};
  if (!fc.f) return set_error("cannot open %s", fname);
  int64_t m = 0, n = 0;
  RSVD_TRY(read_header(fc.f, fname, format, &m, &n));
  if (lda < (m > 1 ? m : 1)) return set_error("rsvd_dev_load_binary: lda %lld < rows %lld", lda,
                                              (long long)m);
  if (m == 0 || n == 0) return 0;
  const size_t eb = entry_bytes(format);
  const int64_t R = rows_per_block(m, n, eb);
  const size_t blk = (size_t)(R * n) * eb;
  PinnedPair host(c->stream);
  RSVD_TRY(host.alloc(blk));
  DBuf<char> stage;
  RSVD_TRY(stage.alloc(2 * blk));
  EventPair ev;
  RSVD_TRY(ev.create());
  const bool f32in = format == 2, f32out = dtype == 1;
  int b = 0;
  for (int64_t r0 = 0; r0 < m; r0 += R, b ^= 1) {
    const int64_t rows = (m - r0 < R) ? m - r0 : R;
    const size_t bytes = (size_t)(rows * n) * eb;
    // host buffer b is free again once the copy issued two blocks ago ended
    if (r0 >= 2 * R) RSVD_HIP(hipEventSynchronize(ev.e[b]));
    if (fread(host.p[b], 1, bytes, fc.f) != bytes) {
      (void)hipStreamSynchronize(c->stream);
      return set_error("%s: truncated in rows %lld..%lld", fname, (long long)r0,
                       (long long)(r0 + rows - 1));
    }
    char* st = stage.p + (size_t)b * blk;
    RSVD_HIP(hipMemcpyAsync(st, host.p[b], bytes, hipMemcpyHostToDevice, c->stream));
    RSVD_HIP(hipEventRecord(ev.e[b], c->stream));
    char* dst = (char*)dA + (size_t)r0 * (f32out ? 4 : 8);
    RSVD_TRY(launch_rows_to_cols(c, f32in, f32out, rows, n, st, dst, lda));
  }
  RSVD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int rsvd_dev_write_binary(const char* fname, int format, int dtype, const void* dA, long long m,
                          long long n, long long lda) {
  clear_error();
  Context* c = nullptr;
  RSVD_TRY(ctx_get(&c));
  std::unique_lock<std::recursive_mutex> lk(c->mu);
  RSVD_TRY(order_after_null_stream(c));
  if (!fname || (!dA && m && n)) return set_error("rsvd_dev_write_binary: NULL argument");
  if (dtype != 0 && dtype != 1) return set_error("rsvd_dev_write_binary: dtype %d", dtype);
  if (format < 0 || format > 2) return set_error("binary file: unknown format %d", format);
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1))
    return set_error("rsvd_dev_write_binary: bad shape %lld x %lld, lda %lld", m, n, lda);
  FileCloser fc{fopen(fname, "wb")};
  if (!fc.f) return set_error("cannot open %s", fname);
  RSVD_TRY(write_header(fc.f, format, m, n));
  if (m == 0 || n == 0) return 0;
  const size_t eb = entry_bytes(format);
  const int64_t R = rows_per_block(m, n, eb);
  const size_t blk = (size_t)(R * n) * eb;
  PinnedPair host(c->stream);
  RSVD_TRY(host.alloc(blk));
  DBuf<char> stage;
  RSVD_TRY(stage.alloc(blk));
  const bool f32in = dtype == 1, f32out = format == 2;
  for (int64_t r0 = 0; r0 < m; r0 += R) {
    const int64_t rows = (m - r0 < R) ? m - r0 : R;
    const size_t bytes = (size_t)(rows * n) * eb;
    const char* src = (const char*)dA + (size_t)r0 * (f32in ? 4 : 8);
    RSVD_TRY(launch_cols_to_rows(c, f32in, f32out, rows, n, src, lda, stage.p));
    RSVD_HIP(hipMemcpyAsync(host.p[0], stage.p, bytes, hipMemcpyDeviceToHost, c->stream));
    RSVD_HIP(hipStreamSynchronize(c->stream));
    if (fwrite(host.p[0], 1, bytes, fc.f) != bytes) return set_error("%s: write failed", fname);
  }
  return 0;
}

}  // extern "C"
