// rsvd_internal.h -- device context, error plumbing, workspace pool and the
// kernel-level entry points shared by the librsvdpack translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace rsvd {

// ---------------------------------------------------------------- errors
// Every internal routine returns int: 0 = success, nonzero = failure, with the
// message stored in a thread-local buffer that rsvd_last_error() exposes.
int set_error(const char* fmt, ...);
void clear_error();
const char* last_error();

#define RSVD_HIP(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      return ::rsvd::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,     \
                               hipGetErrorString(e_));                          \
  } while (0)

#define RSVD_TRY(call)        \
  do {                        \
    int r_ = (call);          \
    if (r_) return r_;        \
  } while (0)

#define RSVD_LAUNCH_CHECK() RSVD_HIP(hipGetLastError())

// ---------------------------------------------------------------- context
struct KernelStat {
  std::string name;
  double total_ms = 0;
  long long count = 0;
};

struct Context {
  int device = -1;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;
  // cached device buffers, reused across calls (size-bucketed free list)
  struct Block { void* p; size_t bytes; };
  std::vector<Block> free_blocks;
  size_t pooled_bytes = 0;
  // per-kernel timing (enabled by rsvd_kstats_enable)
  bool kstats = false;
  std::vector<std::string> kstats_only;  // empty: every timer records
  std::vector<KernelStat> stats;
  // pending event pairs, resolved on rsvd_kstats_get / sync
  struct Pending { int idx; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  // pinned host slot for the small device -> host reads of the control flow
  // (convergence and breakdown flags, norms); allocated on first use, kept
  void* pinned = nullptr;
  // recorded on the null stream at the entry of device-pointer routines
  hipEvent_t entry_ev = nullptr;
  // Allocation-free mode: while scratch_mode is set, dmalloc hands out
  // first-fit ranges of the library's static device scratch array (device
  // address `scratch`, kScratchBytes long) instead of pooled hipMalloc
  // memory, and dfree returns them.  Live ranges: (offset, bytes), sorted.
  char* scratch = nullptr;
  bool scratch_mode = false;
  std::vector<std::pair<size_t, size_t>> scratch_live;
  // rsvd_set_fp64_vendor_gemm: the large FP64 products through rocBLAS
  // (handle on `stream`, rocblas_handle, created on first use)
  bool fp64_vendor = false;
  void* blas = nullptr;
};

// Size of the static device scratch array behind the allocation-free entry
// point (rsvd_dev_fixed_rank_noalloc).
constexpr size_t kScratchBytes = size_t(256) << 20;

// Scope of an allocation-free call: every device buffer the library takes
// inside it comes from the scratch array; none is left live at its end.
struct ScratchScope {
  Context* c;
  explicit ScratchScope(Context* ctx) : c(ctx) {
    c->scratch_live.clear();
    c->scratch_mode = true;
  }
  ~ScratchScope() {
    c->scratch_mode = false;
    c->scratch_live.clear();
  }
};

// Lazily initialises the current device and the library stream.
int ctx_get(Context** out);

// Makes the library stream wait for the work queued on the null stream so far.
int order_after_null_stream(Context* c);

// dst <- bytes (<= 64) at device address src, after the work queued on the
// library stream: one direct copy into pinned host memory, then a wait.
int read_small(Context* c, void* dst, const void* src, size_t bytes);

// Device buffer from the pool (contents undefined).  free() returns it.
int dmalloc(void** p, size_t bytes);
void dfree(void* p);
// Release every pooled buffer back to HIP.
void pool_trim();

// RAII buffer that returns to the pool.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { reset(); }
  int alloc(size_t count) {
    reset();
    void* q = nullptr;
    int r = dmalloc(&q, (count ? count : 1) * sizeof(T));
    if (r) return r;
    p = static_cast<T*>(q);
    n = count;
    return 0;
  }
  void reset() {
    if (p) dfree(p);
    p = nullptr;
    n = 0;
  }
};

// Kernel timing scope: records HIP events around launches on the library
// stream when kernel statistics are enabled.
struct KTimer {
  Context* c;
  int idx = -1;
  hipEvent_t a = nullptr, b = nullptr;
  KTimer(Context* ctx, const char* name);
  ~KTimer();
};

// ---------------------------------------------------------------- views
// Column-major device matrix view: element (i,j) at p[i + j*ld].
template <typename T>
struct DMat {
  T* p;
  int64_t rows, cols, ld;
  __host__ __device__ T* col(int64_t j) const { return p + j * ld; }
};

// k slices for a grid of `tiles` workgroups when `slots` of them run at once:
// the s in [1, smax] that minimises ceil(tiles s / slots) / s (waves per unit
// of work), the fewest slices among equals.
static inline int pick_k_slices(int64_t tiles, int64_t slots, int64_t smax) {
  int best = 1;
  double best_cost = 1e300;
  for (int64_t t = 1; t <= smax; ++t) {
    const double cost = (double)((tiles * t + slots - 1) / slots) / (double)t;
    if (cost < best_cost * (1.0 - 1e-9)) {
      best_cost = cost;
      best = (int)t;
    }
  }
  return best;
}

// ---------------------------------------------------------------- kernels
// GEMM: C = alpha * op(A) * op(B) + beta * C, column-major, double or float.
// ta/tb: 'N' or 'T'.  sumsq (optional, double[grid]) receives per-workgroup
// partial sums of squares of the stored C (fused Frobenius norm); pass nullptr
// to skip.
int gemm_f64(Context* c, char ta, char tb, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
             int64_t ldc);
// Same as gemm_f64, timed under the kernel-statistics family `tag`.
int gemm_f64_tag(Context* c, const char* tag, char ta, char tb, int64_t M, int64_t N, int64_t K,
                 double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                 double beta, double* C, int64_t ldc);
int gemm_f32(Context* c, char ta, char tb, int64_t M, int64_t N, int64_t K, float alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C,
             int64_t ldc);
// fp32 C = alpha*op(A)*op(B) + beta*C on the bf16 matrix cores, fp32 operands
// rounded to bf16 while staged (split = 1) or split into bf16 hi + lo with
// three products (split = 3, ~fp32 accuracy).  `tag` names the launch for
// kernel statistics.
// C = alpha op(A) S + beta C for a skinny fp32 S (K x N, k-contiguous), A
// read once from HBM; bf16 (split 1) or split-bf16 (split 3) matrix cores.
int apass_bf16(Context* c, const char* tag, bool ta, int64_t M, int64_t N, int64_t K, float alpha,
               const float* A, int64_t lda, const float* S, int64_t lds, float beta, float* C,
               int64_t ldc, int split);
// a_c = max |A| over the 128-column chunk c of A (M x N, ld lda) / 448, into
// scale[0 .. ceil(N / 128)) (1 for an all-zero chunk): the fp8 e4m3 scales.
int fp8_column_scales(Context* c, int64_t M, int64_t N, const float* A, int64_t lda, float* scale);
// C = alpha op(A) S + beta C with A quantised to e4m3 per 128-column chunk of A
// (ascale from fp8_column_scales of A) and S to e4m3 per output column, on
// the fp8 matrix cores with fp32 accumulation; a K % 32 remainder in
// split-bf16.
int apass_fp8(Context* c, const char* tag, bool ta, int64_t M, int64_t N, int64_t K, float alpha,
              const float* A, int64_t lda, const float* ascale, const float* S, int64_t lds,
              float beta, float* C, int64_t ldc);
// a_c = max |A| over the 128-column chunk c / 32768: the fp16 scales.
int fp16_column_scales(Context* c, int64_t M, int64_t N, const float* A, int64_t lda,
                       float* scale);
// C = alpha op(A) S + beta C with A / a_c rounded to fp16 (ascale from
// fp16_column_scales of A) and S in two fp16 planes (hi + lo, per-column
// scales), two products on the fp16 matrix cores, fp32 accumulation; shapes
// the streaming kernel does not fill, and a K % 32 remainder, in split-bf16.
int apass_fp16(Context* c, const char* tag, bool ta, int64_t M, int64_t N, int64_t K, float alpha,
               const float* A, int64_t lda, const float* ascale, const float* S, int64_t lds,
               float beta, float* C, int64_t ldc);
// true when apass_bf16 has enough row blocks to fill the GPU.
bool apass_suitable(Context* c, int64_t M, int64_t N);
// The tiled bf16 GEMM without the A-streaming dispatch (any shape).
int gemm_bf16_general(Context* c, const char* tag, char ta, char tb, int64_t M, int64_t N,
                      int64_t K, float alpha, const float* A, int64_t lda, const float* B,
                      int64_t ldb, float beta, float* C, int64_t ldc, int split);
int gemm_bf16(Context* c, const char* tag, char ta, char tb, int64_t M, int64_t N, int64_t K,
              float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
              float* C, int64_t ldc, int split);
// fp32 C = C - Q Bt^T (Q M x K, Bt N x K) with fp32-exact products (three
// bf16 planes per operand, six products, fp32 accumulation) for k >= klo and
// the bf16 product alone for k < klo (a multiple of 32), and the FP64 sum of
// squares of the updated C into *sumsq (device scalar, overwritten).
int gemm_bf16_deflate(Context* c, int64_t M, int64_t N, int64_t K, const float* Q, int64_t ldq,
                      const float* Bt, int64_t ldbt, float* C, int64_t ldc, double* sumsq,
                      int64_t klo);
// C = C - Qp*Bp with fused FP64 sum of squares of the updated C into *sumsq
// (device scalar, overwritten).
int gemm_f64_deflate(Context* c, int64_t M, int64_t N, int64_t K, const double* Q, int64_t ldq,
                     const double* B, int64_t ldb, double* C, int64_t ldc, double* sumsq);

// Gaussian fill: out[i + j*ld] = element (col0+j)*rows_total + i of the stream.
int gaussian_f64(Context* c, uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                 int64_t rows_total, double* out, int64_t ld);
int gaussian_f32(Context* c, uint64_t seed, int64_t rows, int64_t cols, int64_t col0,
                 int64_t rows_total, float* out, int64_t ld);

// A(i,j) += s * element (j*rows + i) of the Gaussian stream for `seed`.
int add_gaussian_f64(Context* c, uint64_t seed, int64_t rows, int64_t cols, double s, double* A,
                     int64_t lda);

// Elementwise / reductions (double)
int copy2d_f64(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds, double* dst,
               int64_t ldd);
int fill2d_f64(Context* c, int64_t rows, int64_t cols, double v, double* dst, int64_t ldd);
int sumsq_f64(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds,
              double* out_scalar);
int sumsq_diff_f64(Context* c, int64_t rows, int64_t cols, const double* a, int64_t lda,
                   const double* b, int64_t ldb, double* out_scalar);
int transpose_f64(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds,
                  double* dst, int64_t ldd);
int scale_cols_f64(Context* c, int64_t rows, int64_t cols, double* a, int64_t lda,
                   const double* colscale, int invert);
int copy2d_f32(Context* c, int64_t rows, int64_t cols, const float* src, int64_t lds, float* dst,
               int64_t ldd);
// Error (set_error, "<what>: non-finite singular values ...") when any of
// x[0 .. n) (device) is NaN or infinite; synchronises.
int check_finite_f64(Context* c, const double* x, int64_t n, const char* what);
// out[j] = sqrt(in[j])
int sqrt_f64(Context* c, int64_t n, const double* in, double* out);
int f32_to_f64(Context* c, int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst,
               int64_t ldd);
int f64_to_f32(Context* c, int64_t rows, int64_t cols, const double* src, int64_t lds, float* dst,
               int64_t ldd);
// Plain device copy of `bytes` (multiple of 16, 16-byte aligned pointers).
int stream_copy(Context* c, const void* src, void* dst, int64_t bytes);
// G(i,j) = G(j,i) for i > j (upper triangle mirrored).
int symmetrize_upper_f64(Context* c, int64_t n, double* G);
int add_diag_f64(Context* c, int64_t n, double v, double* a, int64_t lda);
int add_scalar_f64(Context* c, int64_t n, double v, double* a);
// out[j] = X(:,j) . Y(:,j)
int coldot_f64(Context* c, int64_t L, int64_t n, const double* X, int64_t ldx, const double* Y,
               int64_t ldy, double* out);

// Orthonormalisation.  In place: Y (m x n) -> Q with Q^T Q = I and Y = Q R.
// R (n x n, upper, ld n) optional.  Uses CholeskyQR2; falls back to
// Householder when the Gram matrix is numerically singular.  *used_fallback
// (optional) reports which path ran.
int orth_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* R,
             int* used_fallback);
// Householder QR in place (blocked geqrf + orgqr, LAPACK's reflectors and
// sign convention), R (n x n, ld n) optional.  vwork (optional, m * min(n, 32)
// doubles, device) holds the panel reflectors; allocated when null.
int householder_qr_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* R,
                       double* vwork = nullptr);
// Blocked Householder QR of Y (m x n, m >= n) in place: R in the upper
// triangle, reflector j below the diagonal of column j (v_j(j) = 1 implied),
// tau (n, device).  vwork as above.  Tall (optional, ceil(n / 32) * 32 * 32
// doubles, device) receives each 32-column panel's block-reflector factor T
// (H_j0 ... H_j0+31 = I - V T V^T), for orgqr_f64.
int geqrf_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, double* tau,
              double* vwork = nullptr, double* Tall = nullptr);
// The first n columns of Q = H_0 ... H_{n-1} in place of the reflectors
// stored below the diagonal of Y (m x n, v_j(j) = 1 implied) with tau
// (device): blocked, 32 columns a panel; the panels' T from geqrf_f64 when
// Tall is given, formed from V and tau otherwise.
int orgqr_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, const double* tau,
              double* vwork = nullptr, const double* Tall = nullptr);
// orgqr_f64 with its own work buffer.
int org2r_f64(Context* c, int64_t m, int64_t n, double* Y, int64_t ldy, const double* tau);
// R = chol(G) (upper) and Rinv = R^-1 in one blocked pass (n x n, ld n, zero
// below the diagonal).  G's trailing part is overwritten.  *ok = 0 when a
// pivot is not safely positive.
int chol_inv_upper_f64(Context* c, int64_t n, double* G, double* R, double* Rinv, int* ok);

// One-sided Jacobi.  W (L x n, ld L) is rotated in place until its columns
// are mutually orthogonal to rel_tol (0: the rounding floor eps*sqrt(L));
// V (n x n, ld n, optional) accumulates the rotations (start from identity).
// *sweeps_out = sweeps used.
constexpr double kJacobiTol = 1e-12;
int jacobi_f64(Context* c, int64_t L, int64_t n, double* W, int64_t ldw, double* V, int64_t ldv,
               int max_sweeps, int* sweeps_out, double rel_tol = kJacobiTol);
// Batched fixed-rank rSVD (batched.hip); arguments as rsvd_dev_fixed_rank_batched.
int fixed_rank_batched(Context* c, int count, const int* m, const int* n, const int* k, int p,
                       int q, uint64_t seed, const float* dA, const long long* a_off, float* dU,
                       const long long* u_off, double* dS, const long long* s_off, float* dV,
                       const long long* v_off, int* status);
const char* batched_last_config();
void batched_release_plans();
// Learned (q, rank) front end (learned.hip).
int mlp_num_params();
int mlp_init(Context* c, uint64_t seed, float* theta);
int mlp_forward(Context* c, const float* theta, const float* X, int B, float* out);
int mlp_grad(Context* c, const float* theta, const float* X, const int* idx, const int* q,
             const float* r, float lam, int B, float* grad, float* loss);
int mlp_train(Context* c, float* theta, const float* X, const int* q, const float* r, int n,
              float lam, int epochs, int batch, float lr, uint64_t seed, float* loss_hist);
// The 64 learned-path features of A (FP64) from a sketch of its float copy.
int sketch_features(Context* c, const double* A, const float* A32, int64_t m, int64_t n,
                    int64_t lda, uint64_t seed, float* feat);
// Symmetric eigensolver (two-sided block Jacobi).  G (n x n, ld ldg, both
// triangles, symmetric) is diagonalised in place: on return G ~ diag(lambda)
// and V (optional, n x n) holds the eigenvectors (V^T G_in V = diag);
// evals (optional) = diag(G), unsorted.  *sweeps_out = sweeps used.  The
// iteration ends after the first sweep whose rotations were all below 1e-10
// of sqrt(g_pp g_qq), or 1e-4 with loose = 1 (eigenvectors to fp32 level).
int symeig_jacobi_f64(Context* c, int64_t n, double* G, int64_t ldg, double* V, int64_t ldv,
                      double* evals, int max_sweeps, int* sweeps_out, int loose = 0);
// Column norms of W into nrm[n]; optional normalisation of the columns.
int colnorms_f64(Context* c, int64_t L, int64_t n, double* W, int64_t ldw, double* nrm,
                 int normalise);
// Sorts (values, columns of X[, columns of Y]) by value, ascending or
// descending; perm scratch handled internally.
int sort_pairs_f64(Context* c, int64_t n, double* vals, int descending, double* X, int64_t rowsX,
                   int64_t ldx, double* Y, int64_t rowsY, int64_t ldy);

}  // namespace rsvd
