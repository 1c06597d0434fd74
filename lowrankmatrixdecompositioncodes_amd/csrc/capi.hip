// capi.hip -- the RSVDPACK C API (rsvdpack.h) and the additive extensions
// (rsvdpack_ext.h) on top of the device pipeline.
//
// Host `mat` / `vec` storage, element access, file I/O, gathers, resizes and
// other pure data movement stay host code (the structs are host memory by
// contract).  Every arithmetic routine -- GEMMs, QR, eigen / SVD, norms,
// Gaussian draws and the L2 decompositions -- runs on the GPU; when no gfx950
// device is present the call fails (outputs NULL, rsvd_last_error() set) and
// nothing is computed on the CPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <vector>

#include "pipeline.h"
#include "idcur.h"
#include "pipeline_mp.h"
#include "rsvdpack.h"
#include "rsvdpack_ext.h"

using namespace rsvd;

namespace {

std::mutex g_cfg_mu;
bool g_seed_fixed = false;
unsigned long long g_seed = 0;
std::vector<double> g_omega;  // injected Omega (host copy)

unsigned long long draw_seed() {
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  return g_seed_fixed ? g_seed : (unsigned long long)time(NULL);
}

// Omega for a draw of `count` elements: injected host data uploaded to `buf`,
// or the seeded stream.
int host_omega(Context* c, int64_t n, int64_t cols, DBuf<double>& buf, OmegaSrc* om) {
  std::vector<double> inj;
  {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    if (!g_omega.empty()) inj = g_omega;
  }
  if (!inj.empty()) {
    const size_t need = (size_t)n * (size_t)cols;
    if (inj.size() < need)
      return set_error("injected Omega holds %zu values, %zu needed", inj.size(), need);
    RSVD_TRY(buf.alloc(need));
    RSVD_HIP(hipMemcpyAsync(buf.p, inj.data(), need * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    RSVD_HIP(hipStreamSynchronize(c->stream));
    om->dev = buf.p;
    om->ld = n;
    return 0;
  }
  om->seed = draw_seed();
  om->dev = nullptr;
  return 0;
}

int upload(Context* c, const mat* M, DBuf<double>& d) {
  const size_t N = (size_t)M->nrows * (size_t)M->ncols;
  RSVD_TRY(d.alloc(N));
  if (N)
    RSVD_HIP(hipMemcpyAsync(d.p, M->d, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return 0;
}

int download(Context* c, const double* src, int64_t rows, int64_t cols, int64_t ld, double* dst) {
  if (rows * cols == 0) return 0;
  if (ld == rows) {
    RSVD_HIP(hipMemcpyAsync(dst, src, sizeof(double) * rows * cols, hipMemcpyDeviceToHost,
                            c->stream));
  } else {
    RSVD_HIP(hipMemcpy2DAsync(dst, sizeof(double) * rows, src, sizeof(double) * ld,
                              sizeof(double) * rows, cols, hipMemcpyDeviceToHost, c->stream));
  }
  return 0;
}

int stream_sync(Context* c) {
  RSVD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

bool valid(const mat* M) { return M && M->nrows >= 0 && M->ncols >= 0 && (M->d || !M->nrows || !M->ncols); }

// The reference returns whatever a NaN or Inf in M propagates to; so does
// this host API, and it also leaves the reason in rsvd_last_error.
static void flag_nonfinite(const std::vector<double>& sv, const char* what) {
  for (double x : sv)
    if (!std::isfinite(x)) {
      set_error("%s: non-finite singular values (the input holds NaN or Inf)", what);
      return;
    }
}

mat* diag_mat(const std::vector<double>& s, int k) {
  mat* S = matrix_new(k, k);
  for (int i = 0; i < k; ++i) S->d[(size_t)i * k + i] = s[i];
  return S;
}

struct Api {
  Context* c = nullptr;
  std::unique_lock<std::recursive_mutex> lk;
  int rc = 0;
  Api() {
    clear_error();
    rc = ctx_get(&c);
    if (!rc) lk = std::unique_lock<std::recursive_mutex>(c->mu);
  }
};

// Entry of a routine that reads or writes caller device memory.  The library
// stream is non-blocking, so the work the caller queued on the null stream
// before the call (and, through it, on every blocking stream) is made to
// precede the call's own work; every such routine also ends with a stream
// synchronisation, so its outputs are complete when it returns.
struct DevApi : Api {
  DevApi() {
    if (!rc) rc = order_after_null_stream(c);
  }
};

// Device outputs of a decomposition that the library allocates for the
// caller (U, S, V): all three or none; freed again unless release() hands
// them over.
struct OutBufs {
  void* p[3] = {nullptr, nullptr, nullptr};
  int alloc(size_t b0, size_t b1, size_t b2) {
    const size_t b[3] = {b0, b1, b2};
    for (int i = 0; i < 3; ++i) {
      hipError_t e = hipMalloc(&p[i], b[i] ? b[i] : 1);
      if (e != hipSuccess) {
        p[i] = nullptr;
        return set_error("hipMalloc(%zu) for an output failed: %s", b[i], hipGetErrorString(e));
      }
    }
    return 0;
  }
  template <typename TU, typename TV>
  void release(TU** u, double** s, TV** v) {
    *u = (TU*)p[0];
    *s = (double*)p[1];
    *v = (TV*)p[2];
    p[0] = p[1] = p[2] = nullptr;
  }
  ~OutBufs() {
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

}  // namespace

extern "C" {

// ======================================================================
// L1 storage (host)
// ======================================================================
mat* matrix_new(int nrows, int ncols) {
  mat* M = (mat*)malloc(sizeof(mat));
  if (!M) return nullptr;
  size_t N = (size_t)(nrows > 0 ? nrows : 0) * (size_t)(ncols > 0 ? ncols : 0);
  M->d = (double*)calloc(N ? N : 1, sizeof(double));
  M->nrows = nrows;
  M->ncols = ncols;
  return M;
}

vec* vector_new(int nrows) {
  vec* v = (vec*)malloc(sizeof(vec));
  if (!v) return nullptr;
  v->d = (double*)calloc(nrows > 0 ? (size_t)nrows : 1, sizeof(double));
  v->nrows = nrows;
  return v;
}

void matrix_delete(mat* M) {
  if (!M) return;
  free(M->d);
  free(M);
}

void vector_delete(vec* v) {
  if (!v) return;
  free(v->d);
  free(v);
}

void matrix_set_element(mat* M, int row_num, int col_num, double val) {
  M->d[(size_t)col_num * M->nrows + row_num] = val;
}
double matrix_get_element(mat* M, int row_num, int col_num) {
  return M->d[(size_t)col_num * M->nrows + row_num];
}
void vector_set_element(vec* v, int row_num, double val) { v->d[row_num] = val; }
double vector_get_element(vec* v, int row_num) { return v->d[row_num]; }

// matrix_load_from_binary_file / matrix_write_to_binary_file: binio.hip

void matrix_print(mat* M) {
  for (int i = 0; i < M->nrows; ++i) {
    for (int j = 0; j < M->ncols; ++j) printf("%f  ", matrix_get_element(M, i, j));
    printf("\n");
  }
}

void vector_print(vec* v) {
  for (int i = 0; i < v->nrows; ++i) printf("%f\n", v->d[i]);
}

void matrix_copy(mat* D, mat* S) {
  memcpy(D->d, S->d, sizeof(double) * (size_t)S->nrows * S->ncols);
}

void matrix_build_transpose(mat* Mt, mat* M) {
  for (int j = 0; j < M->ncols; ++j)
    for (int i = 0; i < M->nrows; ++i) Mt->d[(size_t)i * Mt->nrows + j] = M->d[(size_t)j * M->nrows + i];
}

void matrix_sub(mat* A, mat* B) {
  const size_t N = (size_t)A->nrows * A->ncols;
  for (size_t i = 0; i < N; ++i) A->d[i] -= B->d[i];
}

void matrix_get_selected_columns(mat* M, int* inds, mat* Mc) {
  for (int j = 0; j < Mc->ncols; ++j)
    memcpy(Mc->d + (size_t)j * M->nrows, M->d + (size_t)inds[j] * M->nrows,
           sizeof(double) * M->nrows);
}
void matrix_set_selected_columns(mat* M, int* inds, mat* Mc) {
  for (int j = 0; j < Mc->ncols; ++j)
    memcpy(M->d + (size_t)inds[j] * M->nrows, Mc->d + (size_t)j * M->nrows,
           sizeof(double) * M->nrows);
}
void matrix_get_selected_rows(mat* M, int* inds, mat* Mr) {
  for (int j = 0; j < M->ncols; ++j)
    for (int i = 0; i < Mr->nrows; ++i)
      Mr->d[(size_t)j * Mr->nrows + i] = M->d[(size_t)j * M->nrows + inds[i]];
}
void matrix_set_selected_rows(mat* M, int* inds, mat* Mr) {
  for (int j = 0; j < M->ncols; ++j)
    for (int i = 0; i < Mr->nrows; ++i)
      M->d[(size_t)j * M->nrows + inds[i]] = Mr->d[(size_t)j * Mr->nrows + i];
}

void matrix_copy_symmetric(mat* S, mat* M) {
  for (int j = 0; j < M->ncols; ++j)
    for (int i = 0; i <= j && i < M->nrows; ++i)
      S->d[(size_t)j * S->nrows + i] = M->d[(size_t)j * M->nrows + i];
}

void initialize_diagonal_matrix(mat* D, vec* data) {
  for (int i = 0; i < D->nrows; ++i) D->d[(size_t)i * D->nrows + i] = data->d[i];
}

void invert_diagonal_matrix(mat* Dinv, mat* D) {
  for (int i = 0; i < D->nrows; ++i)
    Dinv->d[(size_t)i * Dinv->nrows + i] = 1.0 / D->d[(size_t)i * D->nrows + i];
}

static void resize_cols(mat** M, int k, int from_end) {
  mat* R = matrix_new((*M)->nrows, k);
  const int off = from_end ? (*M)->ncols - k : 0;
  memcpy(R->d, (*M)->d + (size_t)off * (*M)->nrows, sizeof(double) * (size_t)k * (*M)->nrows);
  matrix_delete(*M);
  *M = R;
}
static void resize_rows(mat** M, int k, int from_end) {
  mat* R = matrix_new(k, (*M)->ncols);
  const int off = from_end ? (*M)->nrows - k : 0;
  for (int j = 0; j < R->ncols; ++j)
    memcpy(R->d + (size_t)j * k, (*M)->d + (size_t)j * (*M)->nrows + off, sizeof(double) * k);
  matrix_delete(*M);
  *M = R;
}
void resize_matrix_by_columns(mat** M, int k) { resize_cols(M, k, 0); }
void resize_matrix_by_columns_from_end(mat** M, int k) { resize_cols(M, k, 1); }
void resize_matrix_by_rows(mat** M, int k) { resize_rows(M, k, 0); }
void resize_matrix_by_rows_from_end(mat** M, int k) { resize_rows(M, k, 1); }

double get_seconds_frac(struct timeval start_timeval, struct timeval end_timeval) {
  long secs_used = (end_timeval.tv_sec - start_timeval.tv_sec);
  long micros_used = ((secs_used * 1000000) + end_timeval.tv_usec) - (start_timeval.tv_usec);
  return (micros_used / 1e6);
}

// ======================================================================
// L1 compute (GPU)
// ======================================================================
void initialize_random_matrix(mat* M) {
  Api api;
  if (api.rc || !valid(M)) { if (!api.rc) set_error("initialize_random_matrix: bad matrix"); return; }
  const int64_t m = M->nrows, n = M->ncols;
  DBuf<double> d, ob;
  OmegaSrc om;
  if (host_omega(api.c, m, n, ob, &om)) return;
  if (d.alloc((size_t)(m * n))) return;
  if (om.dev) {
    if (copy2d_f64(api.c, m, n, om.dev, om.ld, d.p, m)) return;
  } else if (gaussian_f64(api.c, om.seed, m, n, 0, m, d.p, m)) {
    return;
  }
  if (download(api.c, d.p, m, n, m, M->d)) return;
  stream_sync(api.c);
}

static void gemm_host(char ta, char tb, mat* A, mat* B, mat* C, int64_t M, int64_t N, int64_t K) {
  Api api;
  if (api.rc) return;
  if (!valid(A) || !valid(B) || !valid(C)) { set_error("gemm: bad matrix"); return; }
  if (C->nrows != M || C->ncols != N) {
    set_error("gemm: output is %dx%d, product is %lldx%lld", C->nrows, C->ncols, (long long)M,
              (long long)N);
    return;
  }
  DBuf<double> dA, dB, dC;
  if (upload(api.c, A, dA) || upload(api.c, B, dB) || dC.alloc((size_t)(M * N))) return;
  if (gemm_f64(api.c, ta, tb, M, N, K, 1.0, dA.p, A->nrows > 0 ? A->nrows : 1, dB.p,
               B->nrows > 0 ? B->nrows : 1, 0.0, dC.p, M > 0 ? M : 1))
    return;
  if (download(api.c, dC.p, M, N, M, C->d)) return;
  stream_sync(api.c);
}

void matrix_matrix_mult(mat* A, mat* B, mat* C) {
  if (A && B && A->ncols != B->nrows) { set_error("matrix_matrix_mult: inner dims differ"); return; }
  gemm_host('N', 'N', A, B, C, A->nrows, B->ncols, A->ncols);
}
void matrix_transpose_matrix_mult(mat* A, mat* B, mat* C) {
  if (A && B && A->nrows != B->nrows) { set_error("matrix_transpose_matrix_mult: inner dims differ"); return; }
  gemm_host('T', 'N', A, B, C, A->ncols, B->ncols, A->nrows);
}
void matrix_matrix_transpose_mult(mat* A, mat* B, mat* C) {
  if (A && B && A->ncols != B->ncols) { set_error("matrix_matrix_transpose_mult: inner dims differ"); return; }
  gemm_host('N', 'T', A, B, C, A->nrows, B->nrows, A->ncols);
}

static double sumsq_host(mat* A, mat* B) {
  Api api;
  if (api.rc) return NAN;
  DBuf<double> dA, dB, out;
  double h = NAN;
  if (upload(api.c, A, dA) || out.alloc(1)) return NAN;
  if (B) {
    if (upload(api.c, B, dB)) return NAN;
    if (sumsq_diff_f64(api.c, A->nrows, A->ncols, dA.p, A->nrows > 0 ? A->nrows : 1, dB.p,
                       B->nrows > 0 ? B->nrows : 1, out.p))
      return NAN;
  } else if (sumsq_f64(api.c, A->nrows, A->ncols, dA.p, A->nrows > 0 ? A->nrows : 1, out.p)) {
    return NAN;
  }
  if (hipMemcpyAsync(&h, out.p, sizeof h, hipMemcpyDeviceToHost, api.c->stream) != hipSuccess)
    return NAN;
  if (stream_sync(api.c)) return NAN;
  return h;
}

double get_matrix_frobenius_norm(mat* M) {
  if (!valid(M)) { set_error("get_matrix_frobenius_norm: bad matrix"); return NAN; }
  return sqrt(sumsq_host(M, nullptr));
}

double get_percent_error_between_two_mats(mat* A, mat* B) {
  if (!valid(A) || !valid(B) || A->nrows != B->nrows || A->ncols != B->ncols) {
    set_error("get_percent_error_between_two_mats: shape mismatch");
    return NAN;
  }
  const double d = sqrt(sumsq_host(A, B));
  const double a = sqrt(sumsq_host(A, nullptr));
  return 100.0 * d / a;
}

void compute_evals_and_evecs_of_symm_matrix(mat* S, vec* evals) {
  Api api;
  if (api.rc) return;
  if (!valid(S) || S->nrows != S->ncols || !evals || evals->nrows < S->nrows) {
    set_error("compute_evals_and_evecs_of_symm_matrix: bad arguments");
    return;
  }
  Context* c = api.c;
  const int64_t n = S->nrows;
  if (n == 0) return;
  DBuf<double> G, G0, V, lam;
  if (upload(c, S, G) || G0.alloc((size_t)(n * n)) || V.alloc((size_t)(n * n)) ||
      lam.alloc((size_t)n))
    return;
  // symmetric from the upper triangle (dsyev 'U' reads only that triangle)
  if (symmetrize_upper_f64(c, n, G.p) || copy2d_f64(c, n, n, G.p, n, G0.p, n)) return;
  int sw = 0;
  if (symeig_jacobi_f64(c, n, G.p, n, V.p, n, nullptr, 60, &sw)) return;
  // eigenvalues as Rayleigh quotients v^T G v of the input
  if (gemm_f64(c, 'N', 'N', n, n, n, 1.0, G0.p, n, V.p, n, 0.0, G.p, n)) return;
  if (coldot_f64(c, n, n, V.p, n, G.p, n, lam.p)) return;
  if (sort_pairs_f64(c, n, lam.p, 0, V.p, n, n, nullptr, 0, 0)) return;
  if (download(c, V.p, n, n, n, S->d) || download(c, lam.p, n, 1, n, evals->d)) return;
  stream_sync(c);
}

static void qr_host(mat* M, mat* Q, mat* R) {
  Api api;
  if (api.rc) return;
  if (!valid(M) || !valid(Q) || Q->nrows != M->nrows || Q->ncols != M->ncols) {
    set_error("QR: Q must have the shape of M");
    return;
  }
  const int64_t m = M->nrows, n = M->ncols;
  if (m < n) { set_error("QR: m=%lld < n=%lld is not supported", (long long)m, (long long)n); return; }
  if (R && (R->nrows < n || R->ncols < n)) { set_error("QR: R must be at least %lldx%lld", (long long)n, (long long)n); return; }
  DBuf<double> Y, dR;
  if (upload(api.c, M, Y)) return;
  if (R && dR.alloc((size_t)(n * n))) return;
  int fb = 0;
  if (orth_f64(api.c, m, n, Y.p, m, R ? dR.p : nullptr, &fb)) return;
  if (download(api.c, Y.p, m, n, m, Q->d)) return;
  if (R) {
    if (R->nrows == n) {
      if (download(api.c, dR.p, n, n, n, R->d)) return;
    } else {
      if (hipMemcpy2DAsync(R->d, sizeof(double) * R->nrows, dR.p, sizeof(double) * n,
                           sizeof(double) * n, n, hipMemcpyDeviceToHost,
                           api.c->stream) != hipSuccess) {
        set_error("QR: R download failed");
        return;
      }
    }
  }
  stream_sync(api.c);
}

void compact_QR_factorization(mat* M, mat* Q, mat* R) { qr_host(M, Q, R); }
void QR_factorization_getQ(mat* M, mat* Q) { qr_host(M, Q, nullptr); }

void singular_value_decomposition(mat* M, mat* U, mat* S, mat* Vt) {
  Api api;
  if (api.rc) return;
  if (!valid(M) || !valid(U) || !valid(S) || !valid(Vt)) { set_error("SVD: bad matrix"); return; }
  Context* c = api.c;
  const int64_t m = M->nrows, n = M->ncols, k = m < n ? m : n;
  if (k == 0) return;
  const bool tall = m >= n;
  // one-sided Jacobi on X (L x k) with L = max(m,n): X = M (tall) or M^T (wide)
  const int64_t L = tall ? m : n;
  DBuf<double> dM, X, V, sig;
  if (upload(c, M, dM) || X.alloc((size_t)(L * k)) || V.alloc((size_t)(k * k)) || sig.alloc((size_t)k))
    return;
  if (tall) {
    if (copy2d_f64(c, m, n, dM.p, m, X.p, m)) return;
  } else if (transpose_f64(c, m, n, dM.p, m, X.p, n)) {
    return;
  }
  int sw = 0;
  if (jacobi_f64(c, L, k, X.p, L, V.p, k, 60, &sw)) return;
  if (colnorms_f64(c, L, k, X.p, L, sig.p, 1)) return;
  if (sort_pairs_f64(c, k, sig.p, 1, X.p, L, L, V.p, k, k)) return;
  // tall: M = X_n Sigma V^T -> U = X_n (m x k), Vt = V^T (k x n)
  // wide: M^T = X_n Sigma V^T -> U = V (m x k), Vt = X_n^T (k x n)
  DBuf<double> T;
  if (T.alloc((size_t)(k * (tall ? n : L)))) return;
  if (tall) {
    if (U->nrows != m || U->ncols < k || Vt->nrows < k || Vt->ncols != n) { set_error("SVD: output shapes"); return; }
    if (download(c, X.p, m, k, m, U->d)) return;
    if (transpose_f64(c, k, k, V.p, k, T.p, k)) return;
    if (hipMemcpy2DAsync(Vt->d, sizeof(double) * Vt->nrows, T.p, sizeof(double) * k,
                         sizeof(double) * k, n, hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
      set_error("SVD: Vt download failed");
      return;
    }
  } else {
    if (U->nrows != m || U->ncols < k || Vt->nrows < k || Vt->ncols != n) { set_error("SVD: output shapes"); return; }
    if (download(c, V.p, m, k, k, U->d)) return;
    if (transpose_f64(c, n, k, X.p, n, T.p, k)) return;
    if (hipMemcpy2DAsync(Vt->d, sizeof(double) * Vt->nrows, T.p, sizeof(double) * k,
                         sizeof(double) * k, n, hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
      set_error("SVD: Vt download failed");
      return;
    }
  }
  std::vector<double> s((size_t)k);
  if (download(c, sig.p, k, 1, k, s.data()) || stream_sync(c)) return;
  for (int64_t i = 0; i < k && i < S->nrows && i < S->ncols; ++i) S->d[(size_t)i * S->nrows + i] = s[i];
}

void form_svd_product_matrix(mat* U, mat* S, mat* V, mat* P) {
  Api api;
  if (api.rc) return;
  if (!valid(U) || !valid(S) || !valid(V) || !valid(P)) { set_error("form_svd_product_matrix: bad matrix"); return; }
  Context* c = api.c;
  const int64_t m = P->nrows, n = P->ncols, k = S->nrows;
  if (U->nrows != m || U->ncols != k || V->nrows != n || V->ncols != S->ncols || S->ncols != k) {
    set_error("form_svd_product_matrix: shape mismatch");
    return;
  }
  DBuf<double> dU, dS, dV, SVt, dP;
  if (upload(c, U, dU) || upload(c, S, dS) || upload(c, V, dV) || SVt.alloc((size_t)(k * n)) ||
      dP.alloc((size_t)(m * n)))
    return;
  if (gemm_f64(c, 'N', 'T', k, n, k, 1.0, dS.p, k > 0 ? k : 1, dV.p, n > 0 ? n : 1, 0.0, SVt.p, k > 0 ? k : 1)) return;
  if (gemm_f64(c, 'N', 'N', m, n, k, 1.0, dU.p, m > 0 ? m : 1, SVt.p, k > 0 ? k : 1, 0.0, dP.p, m > 0 ? m : 1)) return;
  if (download(c, dP.p, m, n, m, P->d)) return;
  stream_sync(c);
}

// ======================================================================
// L2 decompositions
// ======================================================================
void low_rank_svd_rand_decomp_fixed_rank(mat* M, int k, int p, int vnum, int q, int s, int* frank,
                                         mat** U, mat** S, mat** V) {
  (void)frank;  // not written by the reference either
  if (U) *U = nullptr;
  if (S) *S = nullptr;
  if (V) *V = nullptr;
  Api api;
  if (api.rc) return;
  if (!valid(M) || !U || !S || !V) { set_error("low_rank_svd_rand_decomp_fixed_rank: bad arguments"); return; }
  Context* c = api.c;
  const int64_t m = M->nrows, n = M->ncols;
  if (k <= 0 || p < 0 || (int64_t)k + p > (m < n ? m : n)) {
    set_error("low_rank_svd_rand_decomp_fixed_rank: need 0 < k, 0 <= p, k+p <= min(m,n) "
              "(k=%d p=%d m=%lld n=%lld)", k, p, (long long)m, (long long)n);
    return;
  }
  DBuf<double> dA, ob, dU, dS, dV;
  OmegaSrc om;
  if (upload(c, M, dA) || host_omega(c, n, (int64_t)k + p, ob, &om)) return;
  if (dU.alloc((size_t)(m * k)) || dS.alloc((size_t)k) || dV.alloc((size_t)(n * k))) return;
  SvdInfo info{};
  if (fixed_rank_f64(c, dA.p, m, n, m, k, p, vnum, q, s, om, dU.p, m, dS.p, dV.p, n, &info)) return;
  mat* Uh = matrix_new((int)m, k);
  mat* Vh = matrix_new((int)n, k);
  std::vector<double> sv((size_t)k);
  if (download(c, dU.p, m, k, m, Uh->d) || download(c, dV.p, n, k, n, Vh->d) ||
      download(c, dS.p, k, 1, k, sv.data()) || stream_sync(c)) {
    matrix_delete(Uh);
    matrix_delete(Vh);
    return;
  }
  *U = Uh;
  *S = diag_mat(sv, k);
  *V = Vh;
  flag_nonfinite(sv, "low_rank_svd_rand_decomp_fixed_rank");
}

void low_rank_svd_blockrand_decomp_fixed_rank_or_prec(mat* M, int k, int p, double TOL, int vnum,
                                                      int kstep, int q, int s, int* frank,
                                                      mat** U, mat** S, mat** V) {
  if (U) *U = nullptr;
  if (S) *S = nullptr;
  if (V) *V = nullptr;
  Api api;
  if (api.rc) return;
  if (!valid(M) || !U || !S || !V || !frank) {
    set_error("low_rank_svd_blockrand_decomp_fixed_rank_or_prec: bad arguments");
    return;
  }
  Context* c = api.c;
  const int64_t m = M->nrows, n = M->ncols;
  int p_eff = 0, nstep = 0, rank_mode = 0;
  if (blockrand_plan(m, n, k, p, kstep, &p_eff, &nstep, &rank_mode)) return;
  int ks = kstep, ns = nstep, tm = 0;
  if (qb_adjust(m, n, &ks, &ns, 1, &tm)) return;
  DBuf<double> dA, ob;
  OmegaSrc om;
  if (upload(c, M, dA) || host_omega(c, n, (int64_t)ks * ns, ob, &om)) return;
  BlockrandOut out;
  if (blockrand_f64(c, dA.p, m, n, m, k, p, TOL, vnum, kstep, q, s, om, &out)) return;
  const int fr = out.frank;
  mat* Uh = matrix_new((int)m, fr);
  mat* Vh = matrix_new((int)n, fr);
  std::vector<double> sv((size_t)fr);
  if (download(c, out.U.p, m, fr, m, Uh->d) || download(c, out.V.p, n, fr, n, Vh->d) ||
      download(c, out.S.p, fr, 1, fr, sv.data()) || stream_sync(c)) {
    matrix_delete(Uh);
    matrix_delete(Vh);
    return;
  }
  *frank = fr;
  *U = Uh;
  *S = diag_mat(sv, fr);
  *V = Vh;
  flag_nonfinite(sv, "low_rank_svd_blockrand_decomp_fixed_rank_or_prec");
}

static void qb_host(mat* M, int kstep, int nstep, double TOL, int q, int s, int* frank, mat** Q,
                    mat** B, int is_new) {
  if (Q) *Q = nullptr;
  if (B) *B = nullptr;
  Api api;
  if (api.rc) return;
  if (!valid(M) || !Q || !B) { set_error("randQB: bad arguments"); return; }
  Context* c = api.c;
  const int64_t m = M->nrows, n = M->ncols;
  int ks = kstep, ns = nstep, tm = 0;
  if (qb_adjust(m, n, &ks, &ns, is_new, &tm)) return;
  const int64_t l = (int64_t)ks * ns;
  DBuf<double> dA, ob, dQ, dB;
  OmegaSrc om;
  if (upload(c, M, dA) || host_omega(c, n, l, ob, &om)) return;
  if (dQ.alloc((size_t)(m * l)) || dB.alloc((size_t)(l * n))) return;
  int fr = 0, la = 0;
  QbInfo qi{};
  if (qb_blocked_f64(c, dA.p, m, n, m, kstep, nstep, TOL, q, s, is_new, om, dQ.p, m, dB.p, l, &fr,
                     &la, &qi))
    return;
  const int64_t cols = tm ? fr : l;  // TOL mode resizes Q, B to frank (rra.c:1781-1788)
  mat* Qh = matrix_new((int)m, (int)cols);
  mat* Bh = matrix_new((int)cols, (int)n);
  if (download(c, dQ.p, m, cols, m, Qh->d) || download(c, dB.p, cols, n, l, Bh->d) || stream_sync(c)) {
    matrix_delete(Qh);
    matrix_delete(Bh);
    return;
  }
  if (frank) *frank = fr;
  *Q = Qh;
  *B = Bh;
}

void randQB_pb(mat* M, int kstep, int nstep, int q, int s, mat** Q, mat** B) {
  qb_host(M, kstep, nstep, 0.0, q, s, nullptr, Q, B, 0);
}

void randQB_pb_new(mat* M, int kstep, int nstep, double TOL, int q, int s, int* frank, mat** Q,
                   mat** B) {
  qb_host(M, kstep, nstep, TOL, q, s, frank, Q, B, 1);
}

static void report(mat* M, mat* P) {
  printf("norm(M,fro) = %f\n", get_matrix_frobenius_norm(M));
  printf("norm(P,fro) = %f\n", get_matrix_frobenius_norm(P));
  printf("percent error = %f\n", get_percent_error_between_two_mats(M, P));
}

void use_low_rank_svd_for_approximation(mat* M, mat* U, mat* S, mat* V) {
  mat* P = matrix_new(M->nrows, M->ncols);
  form_svd_product_matrix(U, S, V, P);
  report(M, P);
  matrix_delete(P);
}

void use_QB_decomp_for_approximation(mat* M, mat* Q, mat* B) {
  mat* P = matrix_new(M->nrows, M->ncols);
  matrix_matrix_mult(Q, B, P);
  report(M, P);
  matrix_delete(P);
}

// ======================================================================
// interpolative and CUR decompositions (rra.c:1969-2332)
// ======================================================================
static vec* index_vec(const std::vector<int>& I) {
  vec* v = vector_new((int)I.size());
  for (size_t i = 0; i < I.size(); ++i) v->d[i] = (double)I[i];
  return v;
}

// The QB inside draws RN of n x (kstep * nstep) columns after randQB_pb_new's
// kstep clamp and TOL-mode nstep, with nstep = (k + p) / kstep (0 in TOL mode).
static int idcur_setup(Api& api, mat* M, int k, int p, int kstep, DBuf<double>& dA,
                       DBuf<double>& ob, OmegaSrc* om) {
  Context* c = api.c;
  const int64_t m = M->nrows, n = M->ncols;
  if (kstep <= 0) return set_error("kstep must be positive (kstep=%d)", kstep);
  int nstep = k > 0 ? (k + p) / kstep : 0, ks = kstep, tm = 0;
  RSVD_TRY(qb_adjust(m, n, &ks, &nstep, 1, &tm));
  RSVD_TRY(upload(c, M, dA));
  return host_omega(c, n, (int64_t)ks * nstep, ob, om);
}

static mat* download_mat(Context* c, const double* src, int64_t rows, int64_t cols, int64_t ld) {
  mat* H = matrix_new((int)rows, (int)cols);
  if (download(c, src, rows, cols, ld, H->d) || stream_sync(c)) {
    matrix_delete(H);
    return nullptr;
  }
  return H;
}

// rra.c:924-975 pivotedQR_mkl: M(:, I) = Q R with column pivoting (every
// column free), k = min(m, n); R is k x n (m <= n) or k x k, Q m x k.
void pivotedQR_mkl(mat* M, mat** Q, mat** R, vec** I) {
  if (Q) *Q = nullptr;
  if (R) *R = nullptr;
  if (I) *I = nullptr;
  Api api;
  if (api.rc) return;
  if (!valid(M) || !Q || !R || !I) {
    set_error("pivotedQR_mkl: bad arguments");
    return;
  }
  Context* c = api.c;
  const int64_t m = M->nrows, n = M->ncols, k = m < n ? m : n;
  const int64_t Rc = m <= n ? n : k;
  DBuf<double> dA, tau, Rd;
  DBuf<int> perm;
  if (upload(c, M, dA) || tau.alloc((size_t)(k > 0 ? k : 1)) || perm.alloc((size_t)(n > 0 ? n : 1)))
    return;
  if (pivoted_qr_f64(c, m, n, dA.p, m, perm.p, tau.p)) return;
  // R: the upper trapezoid of the first k rows, before Q overwrites them
  if (Rd.alloc((size_t)(k * Rc > 0 ? k * Rc : 1)) || copy2d_f64(c, k, Rc, dA.p, m, Rd.p, k) ||
      zero_lower_f64(c, k, Rc, Rd.p, k) || org2r_f64(c, m, k, dA.p, m, tau.p))
    return;
  std::vector<int> ip((size_t)n);
  mat* Qh = matrix_new((int)m, (int)k);
  mat* Rh = matrix_new((int)k, (int)Rc);
  if (download(c, dA.p, m, k, m, Qh->d) || download(c, Rd.p, k, Rc, k, Rh->d) ||
      hipMemcpyAsync(ip.data(), perm.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost,
                     c->stream) != hipSuccess ||
      stream_sync(c)) {
    set_error("pivotedQR_mkl: copy back failed");
    matrix_delete(Qh);
    matrix_delete(Rh);
    return;
  }
  *Q = Qh;
  *R = Rh;
  *I = index_vec(ip);
}

void id_blockrand_decomp_fixed_rank_or_prec(mat* M, int k, int p, double TOL, int kstep, int q,
                                            int s, int* frank, vec** I, mat** T) {
  if (I) *I = nullptr;
  if (T) *T = nullptr;
  Api api;
  if (api.rc) return;
  if (!valid(M) || !I || !T || !frank) {
    set_error("id_blockrand_decomp_fixed_rank_or_prec: bad arguments");
    return;
  }
  DBuf<double> dA, ob;
  OmegaSrc om;
  if (idcur_setup(api, M, k, p, kstep, dA, ob, &om)) return;
  const int64_t m = M->nrows, n = M->ncols;
  IdOut out;
  if (id_blockrand_f64(api.c, dA.p, m, n, m, k, p, TOL, kstep, q, s, om, &out)) return;
  mat* Th = download_mat(api.c, out.T.p, out.frank, n - out.frank, out.frank);
  if (!Th) return;
  *frank = out.frank;
  *I = index_vec(out.I);
  *T = Th;
}

void id_two_sided_blockrand_decomp_fixed_rank_or_prec(mat* M, int k, int p, double TOL, int kstep,
                                                      int q, int s, int* frank, vec** Icol,
                                                      vec** Irow, mat** T, mat** S) {
  if (Icol) *Icol = nullptr;
  if (Irow) *Irow = nullptr;
  if (T) *T = nullptr;
  if (S) *S = nullptr;
  Api api;
  if (api.rc) return;
  if (!valid(M) || !Icol || !Irow || !T || !S || !frank) {
    set_error("id_two_sided_blockrand_decomp_fixed_rank_or_prec: bad arguments");
    return;
  }
  DBuf<double> dA, ob;
  OmegaSrc om;
  if (idcur_setup(api, M, k, p, kstep, dA, ob, &om)) return;
  const int64_t m = M->nrows, n = M->ncols;
  IdOut col, row;
  if (id_two_sided_blockrand_f64(api.c, dA.p, m, n, m, k, p, TOL, kstep, q, s, om, &col, &row))
    return;
  const int f = col.frank;
  mat* Th = download_mat(api.c, col.T.p, f, n - f, f);
  mat* Sh = Th ? download_mat(api.c, row.T.p, f, m - f, f) : nullptr;
  if (!Sh) {
    matrix_delete(Th);
    return;
  }
  *frank = f;
  *Icol = index_vec(col.I);
  *Irow = index_vec(row.I);
  *T = Th;
  *S = Sh;
}

void cur_blockrand_decomp_fixed_rank_or_prec(mat* M, int k, int p, double TOL, int kstep, int q,
                                             int s, int* frank, mat** C, mat** U, mat** R) {
  if (C) *C = nullptr;
  if (U) *U = nullptr;
  if (R) *R = nullptr;
  Api api;
  if (api.rc) return;
  if (!valid(M) || !C || !U || !R || !frank) {
    set_error("cur_blockrand_decomp_fixed_rank_or_prec: bad arguments");
    return;
  }
  DBuf<double> dA, ob;
  OmegaSrc om;
  if (idcur_setup(api, M, k, p, kstep, dA, ob, &om)) return;
  const int64_t m = M->nrows, n = M->ncols;
  CurOut out;
  if (cur_blockrand_f64(api.c, dA.p, m, n, m, k, p, TOL, kstep, q, s, om, &out)) return;
  const int f = out.frank;
  mat* Ch = download_mat(api.c, out.C.p, m, f, m);
  mat* Uh = Ch ? download_mat(api.c, out.U.p, f, f, f) : nullptr;
  mat* Rh = Uh ? download_mat(api.c, out.R.p, f, n, f) : nullptr;
  if (!Rh) {
    matrix_delete(Ch);
    matrix_delete(Uh);
    return;
  }
  *frank = f;
  *C = Ch;
  *U = Uh;
  *R = Rh;
}

// MKL/mvf.c:1322-1336: P = (C U) R.
void form_cur_product_matrix(mat* C, mat* U, mat* R, mat* P) {
  clear_error();
  if (!valid(C) || !valid(U) || !valid(R) || !valid(P) || C->ncols != U->nrows ||
      U->ncols != R->nrows || P->nrows != C->nrows || P->ncols != R->ncols) {
    set_error("form_cur_product_matrix: shape mismatch");
    return;
  }
  mat* CU = matrix_new(C->nrows, U->ncols);
  matrix_matrix_mult(C, U, CU);
  matrix_matrix_mult(CU, R, P);
  matrix_delete(CU);
}

// [I_k; X^T](inv, :) for an ID's coefficients X (k x (n - k)) and pivots I.
static mat* interp_rows(mat* X, vec* I, int k) {
  const int n = I->nrows;
  std::vector<int> inv((size_t)n);
  for (int i = 0; i < n; ++i) inv[(size_t)(int)I->d[i]] = i;          // mvf.c:1195-1201
  mat* V = matrix_new(n, k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < n; ++i) {
      const int r = inv[(size_t)i];
      V->d[(size_t)j * n + i] = r < k ? (r == j ? 1.0 : 0.0) : X->d[(size_t)(r - k) * k + j];
    }
  return V;
}

// rra.c:2402-2478: MA = M(:, I(1:k)) V^T with V = [I_k; T^T](I^-1, :).
void use_id_decomp_for_approximation(mat* M, mat* T, vec* I, int k) {
  clear_error();
  if (!valid(M) || !valid(T) || !I || I->nrows != M->ncols || k <= 0 || k > M->ncols ||
      T->nrows != k || T->ncols != M->ncols - k) {
    set_error("use_id_decomp_for_approximation: shape mismatch");
    return;
  }
  const int m = M->nrows;
  mat* V = interp_rows(T, I, k);
  mat* MI = matrix_new(m, k);
  for (int j = 0; j < k; ++j)
    memcpy(MI->d + (size_t)j * m, M->d + (size_t)(int)I->d[j] * m, sizeof(double) * m);
  mat* MA = matrix_new(m, M->ncols);
  matrix_matrix_transpose_mult(MI, V, MA);
  printf("norm(M,fro) = %f\n", get_matrix_frobenius_norm(M));
  printf("norm(MA,fro) = %f\n", get_matrix_frobenius_norm(MA));
  printf("percent error = %f\n", get_percent_error_between_two_mats(M, MA));
  matrix_delete(V);
  matrix_delete(MI);
  matrix_delete(MA);
}

// rra.c:2484-2556: MA = U M(Irow(1:k), Icol(1:k)) V^T with V from (Icol, T) and
// U = [I_k; S^T](Irow^-1, :).
void use_id_two_sided_decomp_for_approximation(mat* M, mat* T, mat* S, vec* Icol, vec* Irow,
                                               int k) {
  clear_error();
  if (!valid(M) || !valid(T) || !valid(S) || !Icol || !Irow || Icol->nrows != M->ncols ||
      Irow->nrows != M->nrows || k <= 0 || T->nrows != k || S->nrows != k ||
      T->ncols != M->ncols - k || S->ncols != M->nrows - k) {
    set_error("use_id_two_sided_decomp_for_approximation: shape mismatch");
    return;
  }
  const int m = M->nrows, n = M->ncols;
  mat* V = interp_rows(T, Icol, k);
  mat* U = interp_rows(S, Irow, k);
  mat* MIJ = matrix_new(k, k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i)
      MIJ->d[(size_t)j * k + i] = M->d[(size_t)(int)Icol->d[j] * m + (size_t)(int)Irow->d[i]];
  mat* UMIJ = matrix_new(m, k);
  mat* MA = matrix_new(m, n);
  matrix_matrix_mult(U, MIJ, UMIJ);
  matrix_matrix_transpose_mult(UMIJ, V, MA);
  printf("norm(M,fro) = %f\n", get_matrix_frobenius_norm(M));
  printf("norm(MA,fro) = %f\n", get_matrix_frobenius_norm(MA));
  printf("percent error = %f\n", get_percent_error_between_two_mats(M, MA));
  matrix_delete(V);
  matrix_delete(U);
  matrix_delete(MIJ);
  matrix_delete(UMIJ);
  matrix_delete(MA);
}

// rra.c:2563-2573
void use_cur_decomp_for_approximation(mat* M, mat* C, mat* U, mat* R) {
  mat* P = matrix_new(M->nrows, M->ncols);
  form_cur_product_matrix(C, U, R, P);
  report(M, P);
  matrix_delete(P);
}

// ======================================================================
// extensions
// ======================================================================
const char* rsvd_last_error(void) { return last_error(); }
const char* rsvd_version(void) { return "rsvdpack-mi355x 0.1 (gfx950)"; }

int rsvd_init(void) {
  Api api;
  return api.rc;
}

void rsvd_set_seed(unsigned long long seed) {
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  g_seed = seed;
  g_seed_fixed = true;
}
void rsvd_use_time_seed(void) {
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  g_seed_fixed = false;
}
int rsvd_set_omega(const double* host, long long count) {
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  if (!host || count <= 0) {
    g_omega.clear();
    return 0;
  }
  g_omega.assign(host, host + count);
  return 0;
}

int rsvd_gaussian(unsigned long long seed, long long offset, long long count, double* host_out) {
  Api api;
  if (api.rc) return api.rc;
  if (count <= 0) return 0;
  DBuf<double> d;
  RSVD_TRY(d.alloc((size_t)count));
  // a 1 x count slice of the stream starting at `offset`: one row, the
  // element index of column j is offset + j
  RSVD_TRY(gaussian_f64(api.c, seed, 1, count, offset, 1, d.p, 1));
  RSVD_TRY(download(api.c, d.p, count, 1, count, host_out));
  return stream_sync(api.c);
}

void* rsvd_dev_alloc(size_t bytes) {
  Api api;
  if (api.rc) return nullptr;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  return p;
}
void rsvd_dev_free(void* p) {
  if (!p) return;
  Api api;
  if (api.rc) return;
  (void)hipStreamSynchronize(api.c->stream);
  (void)hipFree(p);
}
int rsvd_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, api.c->stream));
  return stream_sync(api.c);
}
int rsvd_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, api.c->stream));
  return stream_sync(api.c);
}
int rsvd_memset_dev(void* dst, int value, size_t bytes) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_HIP(hipMemsetAsync(dst, value, bytes, api.c->stream));
  return stream_sync(api.c);
}
int rsvd_dev_gaussian(unsigned long long seed, long long m, long long n, double* dst, long long ld) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(gaussian_f64(api.c, seed, m, n, 0, m, dst, ld));
  return stream_sync(api.c);
}
int rsvd_sync(void) {
  Api api;
  if (api.rc) return api.rc;
  return stream_sync(api.c);
}
void rsvd_release_cache(void) {
  Api api;
  if (api.rc) return;
  batched_release_plans();
  pool_trim();
}

int rsvd_dev_fixed_rank(const double* dA, long long m, long long n, long long lda, int k, int p,
                        int vnum, int q, int s, unsigned long long seed, const double* dOmega,
                        double* dU, long long ldu, double* dS, double* dV, long long ldv) {
  DevApi api;
  if (api.rc) return api.rc;
  OmegaSrc om;
  om.seed = seed;
  om.dev = dOmega;
  om.ld = n;
  RSVD_TRY(fixed_rank_f64(api.c, dA, m, n, lda, k, p, vnum, q, s, om, dU, ldu, dS, dV, ldv,
                          nullptr));
  return check_finite_f64(api.c, dS, k, "rsvd_dev_fixed_rank");
}

int rsvd_dev_fixed_rank_noalloc(const double* dA, long long m, long long n, long long lda, int k,
                                int p, int vnum, int q, int s, unsigned long long seed,
                                const double* dOmega, double* dU, long long ldu, double* dS,
                                double* dV, long long ldv) {
  DevApi api;
  if (api.rc) return api.rc;
  // The panels the path holds at once (Omega, Y, Z, the orthonormalisation's
  // copy, B, in doubles) must fit the scratch array; the remaining buffers
  // are checked by the scratch allocator before the kernel that needs them.
  const double l = (double)k + (double)p;
  const double panels = 8.0 * l * ((dOmega ? 0.0 : (double)n) + 2.0 * (double)m + 2.0 * (double)n);
  if (k > 0 && p >= 0 && panels > (double)kScratchBytes)
    return set_error("rsvd_dev_fixed_rank_noalloc: %lld x %lld with k + p = %.0f needs at least "
                     "%.0f MiB of device buffers, more than its %zu MiB scratch; use "
                     "rsvd_dev_fixed_rank", m, n, l, panels / 1048576.0, kScratchBytes >> 20);
  OmegaSrc om;
  om.seed = seed;
  om.dev = dOmega;
  om.ld = n;
  ScratchScope scope(api.c);
  RSVD_TRY(fixed_rank_f64(api.c, dA, m, n, lda, k, p, vnum, q, s, om, dU, ldu, dS, dV, ldv,
                          nullptr));
  return check_finite_f64(api.c, dS, k, "rsvd_dev_fixed_rank_noalloc");
}

long long rsvd_qb_max_cols(long long m, long long n, int kstep, int nstep, int is_new) {
  int ks = kstep, ns = nstep, tm = 0;
  if (qb_adjust(m, n, &ks, &ns, is_new, &tm)) return -1;
  return (long long)ks * ns;
}

int rsvd_dev_randqb(const double* dA, long long m, long long n, long long lda, int kstep, int nstep,
                    double TOL, int q, int s, int is_new, unsigned long long seed, const double* dRN,
                    double* dQ, long long ldq, double* dB, long long ldb, int* frank,
                    double* final_norm) {
  DevApi api;
  if (api.rc) return api.rc;
  OmegaSrc om;
  om.seed = seed;
  om.dev = dRN;
  om.ld = n;
  int fr = 0, la = 0;
  QbInfo qi{};
  RSVD_TRY(qb_blocked_f64(api.c, dA, m, n, lda, kstep, nstep, TOL, q, s, is_new, om, dQ, ldq, dB,
                          ldb, &fr, &la, &qi));
  if (frank) *frank = fr;
  if (final_norm) *final_norm = qi.last_norm;
  return stream_sync(api.c);
}

int rsvd_dev_blockrand(const double* dA, long long m, long long n, long long lda, int k, int p,
                       double TOL, int vnum, int kstep, int q, int s, unsigned long long seed,
                       const double* dRN, int* frank, double** dU, double** dS, double** dV) {
  DevApi api;
  if (api.rc) return api.rc;
  OmegaSrc om;
  om.seed = seed;
  om.dev = dRN;
  om.ld = n;
  if (!frank || !dU || !dS || !dV) return set_error("rsvd_dev_blockrand: null output pointer");
  *dU = *dS = *dV = nullptr;
  BlockrandOut out;
  RSVD_TRY(blockrand_f64(api.c, dA, m, n, lda, k, p, TOL, vnum, kstep, q, s, om, &out));
  const int fr = out.frank;
  RSVD_TRY(check_finite_f64(api.c, out.S.p, fr, "rsvd_dev_blockrand"));
  OutBufs ob;
  RSVD_TRY(ob.alloc(sizeof(double) * (size_t)m * fr, sizeof(double) * (size_t)fr,
                    sizeof(double) * (size_t)n * fr));
  RSVD_HIP(hipMemcpyAsync(ob.p[0], out.U.p, sizeof(double) * (size_t)m * fr,
                          hipMemcpyDeviceToDevice, api.c->stream));
  RSVD_HIP(hipMemcpyAsync(ob.p[1], out.S.p, sizeof(double) * (size_t)fr, hipMemcpyDeviceToDevice,
                          api.c->stream));
  RSVD_HIP(hipMemcpyAsync(ob.p[2], out.V.p, sizeof(double) * (size_t)n * fr,
                          hipMemcpyDeviceToDevice, api.c->stream));
  RSVD_TRY(stream_sync(api.c));
  *frank = fr;
  ob.release(dU, dS, dV);
  return 0;
}

static int dev_gemm(Context* c, int dtype, char ta, char tb, long long M, long long N, long long K,
                    double alpha, const void* A, long long lda, const void* B, long long ldb,
                    double beta, void* C, long long ldc) {
  if (dtype == 0)
    return gemm_f64(c, ta, tb, M, N, K, alpha, (const double*)A, lda, (const double*)B, ldb,
                    beta, (double*)C, ldc);
  if (dtype == 1)
    return gemm_f32(c, ta, tb, M, N, K, (float)alpha, (const float*)A, lda, (const float*)B,
                    ldb, (float)beta, (float*)C, ldc);
  if (dtype == 2 || dtype == 3)
    return gemm_bf16(c, "gemm_bf16", ta, tb, M, N, K, (float)alpha, (const float*)A, lda,
                     (const float*)B, ldb, (float)beta, (float*)C, ldc, dtype == 2 ? 1 : 3);
  if (dtype == 4) {
    if (tb != 'N' && tb != 'n') return set_error("rsvd_dev_gemm: fp8 needs tb = 'N'");
    const bool t = (ta == 'T' || ta == 't');
    const int64_t ar = t ? K : M, ac = t ? M : K;  // A is ar x ac
    DBuf<float> sc;
    RSVD_TRY(sc.alloc((size_t)((ac + 127) / 128 + 1)));
    RSVD_TRY(fp8_column_scales(c, ar, ac, (const float*)A, lda, sc.p));
    return apass_fp8(c, "gemm_fp8", t, M, N, K, (float)alpha, (const float*)A, lda, sc.p,
                     (const float*)B, ldb, (float)beta, (float*)C, ldc);
  }
  if (dtype == 5)
    return gemm_bf16_general(c, "gemm_bf16x6", ta, tb, M, N, K, (float)alpha, (const float*)A,
                             lda, (const float*)B, ldb, (float)beta, (float*)C, ldc, 6);
  if (dtype == 6) {
    if (tb != 'N' && tb != 'n') return set_error("rsvd_dev_gemm: fp16 needs tb = 'N'");
    const bool t = (ta == 'T' || ta == 't');
    const int64_t ar = t ? K : M, ac = t ? M : K;  // A is ar x ac
    DBuf<float> sc;
    RSVD_TRY(sc.alloc((size_t)((ac + 127) / 128 + 1)));
    RSVD_TRY(fp16_column_scales(c, ar, ac, (const float*)A, lda, sc.p));
    return apass_fp16(c, "gemm_fp16", t, M, N, K, (float)alpha, (const float*)A, lda, sc.p,
                      (const float*)B, ldb, (float)beta, (float*)C, ldc);
  }
  return set_error("rsvd_dev_gemm: dtype %d not supported", dtype);
}

int rsvd_dev_gemm(int dtype, char ta, char tb, long long M, long long N, long long K, double alpha,
                  const void* A, long long lda, const void* B, long long ldb, double beta, void* C,
                  long long ldc) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(dev_gemm(api.c, dtype, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc));
  return stream_sync(api.c);
}

int rsvd_dev_fixed_rank_batched(int count, const int* m, const int* n, const int* k, int p, int q,
                                unsigned long long seed, const float* dA, const long long* a_off,
                                float* dU, const long long* u_off, double* dS,
                                const long long* s_off, float* dV, const long long* v_off,
                                int* status) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(fixed_rank_batched(api.c, count, m, n, k, p, q, seed, dA, a_off, dU, u_off, dS, s_off,
                              dV, v_off, status));
  return stream_sync(api.c);
}

const char* rsvd_batched_last_config(void) { return batched_last_config(); }

int rsvd_dev_batch_test_matrices(int count, const int* m, const int* n, const double* tau,
                                 unsigned long long seed, float* dA, const long long* a_off) {
  DevApi api;
  if (api.rc) return api.rc;
  if (count < 0 || (count > 0 && (!m || !n || !tau || !a_off || !dA)))
    return set_error("rsvd_dev_batch_test_matrices: bad arguments");
  std::vector<double> sig;
  for (int i = 0; i < count; ++i) {
    if (m[i] < 1 || n[i] < 1 || !(tau[i] > 0))
      return set_error("rsvd_dev_batch_test_matrices: matrix %d: bad shape or tau", i);
    long long r = m[i] < n[i] ? m[i] : n[i];
    const long long cap = (long long)ceil(19.0 * tau[i]);
    if (r > cap) r = cap;
    sig.resize((size_t)r);
    for (long long j = 0; j < r; ++j) sig[(size_t)j] = exp(-(double)j / tau[i]);
    RSVD_TRY(lowrank_test_matrix(api.c, m[i], n[i], (int)r, sig.data(), 0.0, seed + (uint64_t)i,
                                 nullptr, 0, dA + a_off[i], m[i]));
  }
  return stream_sync(api.c);
}

int rsvd_mlp_num_params(void) { return mlp_num_params(); }

int rsvd_dev_mlp_init(unsigned long long seed, float* dTheta) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(mlp_init(api.c, seed, dTheta));
  return stream_sync(api.c);
}

int rsvd_dev_mlp_forward(const float* dTheta, const float* dX, int batch, float* dOut) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(mlp_forward(api.c, dTheta, dX, batch, dOut));
  return stream_sync(api.c);
}

int rsvd_dev_mlp_grad(const float* dTheta, const float* dX, const int* dq, const float* dr,
                      float lam, int batch, float* dGrad, float* dLoss) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(mlp_grad(api.c, dTheta, dX, nullptr, dq, dr, lam, batch, dGrad, dLoss));
  return stream_sync(api.c);
}

int rsvd_dev_mlp_train(float* dTheta, const float* dX, const int* dq, const float* dr, int n,
                       float lam, int epochs, int batch, float lr, unsigned long long seed,
                       float* loss_hist) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(mlp_train(api.c, dTheta, dX, dq, dr, n, lam, epochs, batch, lr, seed, loss_hist));
  return stream_sync(api.c);
}

int rsvd_dev_sketch_features(const double* dA, const float* dA32, long long m, long long n,
                             long long lda, unsigned long long seed, float* dFeat) {
  DevApi api;
  if (api.rc) return api.rc;
  if (!dA || !dA32 || !dFeat) return set_error("rsvd_dev_sketch_features: NULL argument");
  RSVD_TRY(sketch_features(api.c, dA, dA32, m, n, lda, seed, dFeat));
  return stream_sync(api.c);
}

int rsvd_dev_learned_lowrank(const float* dTheta, const double* dA, const float* dA32,
                             long long m, long long n, long long lda, double tol_rel, int p,
                             double margin, int kstep, int q_fallback, unsigned long long seed,
                             int* frank, double** dU, double** dS, double** dV, int* q_pred,
                             int* k_pred, int* fell_back, double* resid_rel) {
  DevApi api;
  if (api.rc) return api.rc;
  Context* c = api.c;
  if (!dTheta || !dA || !dA32 || !frank || !dU || !dS || !dV)
    return set_error("rsvd_dev_learned_lowrank: NULL argument");
  if (!(tol_rel > 0) || p < 0 || margin < 0)
    return set_error("rsvd_dev_learned_lowrank: need tol_rel > 0, p >= 0, margin >= 0");
  const long long mn = m < n ? m : n;
  if (mn < 64 || mn - p < 1) return set_error("rsvd_dev_learned_lowrank: matrix too small");
  *dU = *dS = *dV = nullptr;
  // prediction
  DBuf<float> feat, out;
  RSVD_TRY(feat.alloc(64));
  RSVD_TRY(out.alloc(6));
  RSVD_TRY(sketch_features(c, dA, dA32, m, n, lda, seed ^ 0x5eed5eedULL, feat.p));
  RSVD_TRY(mlp_forward(c, dTheta, feat.p, 1, out.p));
  float o[6];
  RSVD_HIP(hipMemcpyAsync(o, out.p, sizeof o, hipMemcpyDeviceToHost, c->stream));
  DBuf<double> nrm;
  RSVD_TRY(nrm.alloc(1));
  RSVD_TRY(sumsq_f64(c, m, n, dA, lda, nrm.p));
  double a2 = 0;
  RSVD_HIP(hipMemcpyAsync(&a2, nrm.p, sizeof a2, hipMemcpyDeviceToHost, c->stream));
  RSVD_TRY(stream_sync(c));
  if (!std::isfinite(a2))
    return set_error("rsvd_dev_learned_lowrank: non-finite ||A||_F (the input holds NaN or Inf)");
  int q = 0;
  for (int i = 1; i < 5; ++i)
    if (o[i] > o[q]) q = i;
  // output 5 predicts log(rank) / log(min(m, n))
  double rk = std::isfinite(o[5]) ? o[5] : 1.0;
  rk = rk < 0.0 ? 0.0 : (rk > 1.0 ? 1.0 : rk);
  long long k0 = (long long)ceil(exp(rk * log((double)mn)) * (1.0 + margin)) + 2;
  if (k0 < 1) k0 = 1;
  if (q_pred) *q_pred = q;
  OmegaSrc om;
  om.seed = seed;
  DBuf<double> eye;
  // Up to two fixed-rank attempts, each judged by its FP64 residual: the
  // predicted k, then (if that misses TOL) ceil(1.5 k) + 8 -- an under-
  // predicted rank costs one more fixed-rank solve instead of the whole
  // adaptive method.  A k at the sketch limit (k + p >= min(m, n)) is a
  // prediction of full rank: for m >= n it is solved as such -- k = n, no
  // oversampling, Omega = I, so Y = A and the pipeline is an exact thin QR of
  // A followed by the small eigenproblem (q is moot: Y already spans
  // range(A)).
  for (int attempt = 0; attempt < 2; ++attempt) {
    long long k = attempt == 0 ? k0 : (long long)ceil(1.5 * (double)k0) + 8;
    const bool full = k >= mn - p && m >= n;
    if (full) k = mn;
    else if (k > mn - p) k = mn - p;
    const int pp = full ? 0 : p, qq = full ? 1 : q + 1;
    if (k_pred && attempt == 0) *k_pred = (int)k;
    if (!(full || k < mn - p)) break;  // m < n at the sketch limit
    if (full && !eye.p) {
      RSVD_TRY(eye.alloc((size_t)(n * n)));
      RSVD_TRY(fill2d_f64(c, n, n, 0.0, eye.p, n));
      RSVD_TRY(fill2d_f64(c, 1, n, 1.0, eye.p, n + 1));  // the diagonal
    }
    om.dev = full ? eye.p : nullptr;
    om.ld = full ? n : 0;
    OutBufs ob;
    RSVD_TRY(ob.alloc(sizeof(double) * (size_t)(m * k), sizeof(double) * (size_t)k,
                      sizeof(double) * (size_t)(n * k)));
    double* u = (double*)ob.p[0];
    double* sv = (double*)ob.p[1];
    double* v = (double*)ob.p[2];
    RSVD_TRY(fixed_rank_f64(c, dA, m, n, lda, (int)k, pp, 2, qq, 1, om, u, m, sv, v, n, nullptr));
    double s2 = 0, res = 0;
    if (full) {
      // ||A||^2 - sum s^2 cancels for an exact decomposition: form the
      // residual explicitly (one m x n x n product)
      RSVD_TRY(lowrank_residual(c, dA, m, n, lda, k, u, m, sv, v, n, 0, &res));
    } else {
      RSVD_TRY(sumsq_f64(c, k, 1, sv, k, nrm.p));
      RSVD_TRY(read_small(c, &s2, nrm.p, sizeof s2));
      res = sqrt(a2 > s2 ? a2 - s2 : 0.0);
    }
    if (res <= tol_rel * sqrt(a2)) {
      *frank = (int)k;
      ob.release(dU, dS, dV);
      if (fell_back) *fell_back = attempt == 0 ? 0 : 3;
      if (resid_rel) *resid_rel = res / sqrt(a2);
      return 0;
    }
    if (full) break;  // an exact decomposition above TOL: rounding; the fallback decides
  }
  // fallback: the adaptive blocked method in TOL mode
  om.dev = nullptr;
  om.ld = 0;
  double s2 = 0;
  BlockrandOut bo;
  RSVD_TRY(blockrand_f64(c, dA, m, n, lda, -1, p, tol_rel * sqrt(a2), 2, kstep, q_fallback, 1, om,
                         &bo));
  const int fr = bo.frank;
  OutBufs ob;
  RSVD_TRY(ob.alloc(sizeof(double) * (size_t)m * (fr > 0 ? fr : 1),
                    sizeof(double) * (size_t)(fr > 0 ? fr : 1),
                    sizeof(double) * (size_t)n * (fr > 0 ? fr : 1)));
  if (fr > 0) {
    RSVD_HIP(hipMemcpyAsync(ob.p[0], bo.U.p, sizeof(double) * (size_t)m * fr,
                            hipMemcpyDeviceToDevice, c->stream));
    RSVD_HIP(hipMemcpyAsync(ob.p[1], bo.S.p, sizeof(double) * (size_t)fr,
                            hipMemcpyDeviceToDevice, c->stream));
    RSVD_HIP(hipMemcpyAsync(ob.p[2], bo.V.p, sizeof(double) * (size_t)n * fr,
                            hipMemcpyDeviceToDevice, c->stream));
    RSVD_TRY(sumsq_f64(c, fr, 1, (const double*)ob.p[1], fr, nrm.p));
    RSVD_TRY(read_small(c, &s2, nrm.p, sizeof s2));
  }
  RSVD_TRY(stream_sync(c));
  const double res_fb = sqrt(a2 > s2 ? a2 - s2 : 0.0);
  if (res_fb <= tol_rel * sqrt(a2) || m < n) {
    *frank = fr;
    ob.release(dU, dS, dV);
    if (fell_back) *fell_back = 1;
    if (resid_rel) *resid_rel = res_fb / sqrt(a2);
    return 0;
  }
  // TOL is out of the adaptive method's reach below full rank: the exact
  // full-rank decomposition (k = n, p = 0, Omega = I: a thin QR of A and the
  // small eigenproblem), its residual formed explicitly
  DBuf<double> I;
  RSVD_TRY(I.alloc((size_t)(n * n)));
  RSVD_TRY(fill2d_f64(c, n, n, 0.0, I.p, n));
  RSVD_TRY(fill2d_f64(c, 1, n, 1.0, I.p, n + 1));
  OmegaSrc oi;
  oi.dev = I.p;
  oi.ld = n;
  OutBufs ox;
  RSVD_TRY(ox.alloc(sizeof(double) * (size_t)(m * n), sizeof(double) * (size_t)n,
                    sizeof(double) * (size_t)(n * n)));
  RSVD_TRY(fixed_rank_f64(c, dA, m, n, lda, (int)n, 0, 2, 1, 1, oi, (double*)ox.p[0], m,
                          (double*)ox.p[1], (double*)ox.p[2], n, nullptr));
  double res_x = 0;
  RSVD_TRY(lowrank_residual(c, dA, m, n, lda, n, ox.p[0], m, (const double*)ox.p[1], ox.p[2], n,
                            0, &res_x));
  *frank = (int)n;
  ox.release(dU, dS, dV);
  if (fell_back) *fell_back = 2;
  if (resid_rel) *resid_rel = res_x / sqrt(a2);
  return 0;
}

int rsvd_dev_copy(const void* src, void* dst, long long bytes) {
  DevApi api;
  if (api.rc) return api.rc;
  if (bytes < 0) return set_error("rsvd_dev_copy: negative size");
  RSVD_TRY(stream_copy(api.c, src, dst, bytes));
  return stream_sync(api.c);
}

int rsvd_dev_lowrank_test_matrix(long long m, long long n, int r, const double* sigma,
                                 double noise_std, unsigned long long seed, double* dA64,
                                 long long lda64, float* dA32, long long lda32) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(lowrank_test_matrix(api.c, m, n, r, sigma, noise_std, seed, dA64, lda64, dA32, lda32));
  return stream_sync(api.c);
}

int rsvd_dev_fixed_rank_mp(const float* dA, long long m, long long n, long long lda, int k, int p,
                           int vnum, int q, int s, unsigned long long seed, const float* dOmega,
                           int precision, float* dU, long long ldu, double* dS, float* dV,
                           long long ldv) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(fixed_rank_mp(api.c, dA, m, n, lda, k, p, vnum, q, s, seed, dOmega, precision, dU,
                         ldu, dS, dV, ldv, nullptr));
  return check_finite_f64(api.c, dS, k, "rsvd_dev_fixed_rank_mp");
}

int rsvd_dev_randqb_mp(const float* dA, long long m, long long n, long long lda, int kstep,
                       int nstep, double TOL, int q, int s, int is_new, unsigned long long seed,
                       const float* dRN, int precision, float* dQ, long long ldq, float* dBt,
                       long long ldbt, int* frank, double* final_norm) {
  DevApi api;
  if (api.rc) return api.rc;
  int fr = 0, la = 0;
  QbInfo qi{};
  RSVD_TRY(qb_blocked_mp(api.c, dA, m, n, lda, kstep, nstep, TOL, q, s, is_new, seed, dRN,
                         precision, dQ, ldq, dBt, ldbt, &fr, &la, &qi));
  if (frank) *frank = fr;
  if (final_norm) *final_norm = qi.last_norm;
  return stream_sync(api.c);
}

int rsvd_dev_blockrand_mp(const float* dA, long long m, long long n, long long lda, int k, int p,
                          double TOL, int vnum, int kstep, int q, int s, unsigned long long seed,
                          const float* dRN, int precision, int* frank, float** dU, double** dS,
                          float** dV) {
  DevApi api;
  if (api.rc) return api.rc;
  if (!frank || !dU || !dS || !dV) return set_error("rsvd_dev_blockrand_mp: null output pointer");
  *dU = nullptr;
  *dS = nullptr;
  *dV = nullptr;
  BlockrandMpOut out;
  RSVD_TRY(blockrand_mp(api.c, dA, m, n, lda, k, p, TOL, vnum, kstep, q, s, seed, dRN, precision,
                        &out));
  const int fr = out.frank;
  RSVD_TRY(check_finite_f64(api.c, out.S.p, fr, "rsvd_dev_blockrand_mp"));
  OutBufs ob;
  RSVD_TRY(ob.alloc(sizeof(float) * (size_t)m * fr, sizeof(double) * (size_t)fr,
                    sizeof(float) * (size_t)n * fr));
  RSVD_HIP(hipMemcpyAsync(ob.p[0], out.U.p, sizeof(float) * (size_t)m * fr,
                          hipMemcpyDeviceToDevice, api.c->stream));
  RSVD_HIP(hipMemcpyAsync(ob.p[1], out.S.p, sizeof(double) * (size_t)fr, hipMemcpyDeviceToDevice,
                          api.c->stream));
  RSVD_HIP(hipMemcpyAsync(ob.p[2], out.V.p, sizeof(float) * (size_t)n * fr,
                          hipMemcpyDeviceToDevice, api.c->stream));
  RSVD_TRY(stream_sync(api.c));
  *frank = fr;
  ob.release(dU, dS, dV);
  return 0;
}

int rsvd_dev_lowrank_residual(const double* dA, long long m, long long n, long long lda,
                              long long k, const void* dU, long long ldu, const double* dS,
                              const void* dV, long long ldv, int uv_f32, double* out) {
  DevApi api;
  if (api.rc) return api.rc;
  RSVD_TRY(lowrank_residual(api.c, dA, m, n, lda, k, dU, ldu, dS, dV, ldv, uv_f32, out));
  return stream_sync(api.c);
}

int rsvd_dev_householder_qr(double* dY, long long m, long long n, long long ldy, double* dR) {
  DevApi api;
  if (api.rc) return api.rc;
  if (!dY) return set_error("rsvd_dev_householder_qr: NULL Y");
  if (m < 0 || n < 0 || m < n || ldy < (m > 1 ? m : 1))
    return set_error("rsvd_dev_householder_qr: need m >= n >= 0 and ldy >= m (m=%lld n=%lld "
                     "ldy=%lld)", m, n, ldy);
  RSVD_TRY(householder_qr_f64(api.c, m, n, dY, ldy, dR));
  return stream_sync(api.c);
}

int rsvd_dev_sym_eig_psd(long long n, const double* dG, long long ldg, double* dEvals,
                         double* dEvecs, long long ldv, int* sweeps) {
  DevApi api;
  if (api.rc) return api.rc;
  Context* c = api.c;
  DBuf<double> W;
  RSVD_TRY(W.alloc((size_t)(n * n)));
  RSVD_TRY(copy2d_f64(c, n, n, dG, ldg, W.p, n));
  RSVD_TRY(symmetrize_upper_f64(c, n, W.p));
  int sw = 0;
  RSVD_TRY(symeig_jacobi_f64(c, n, W.p, n, dEvecs, ldv, dEvals, 60, &sw));
  RSVD_TRY(sort_pairs_f64(c, n, dEvals, 0, dEvecs, n, ldv, nullptr, 0, 0));
  if (sweeps) *sweeps = sw;
  return stream_sync(c);
}

// ---------------------------------------------------------------- timing
static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

int rsvd_timer_start(void) {
  Api api;
  if (api.rc) return api.rc;
  if (!g_t0) {
    RSVD_HIP(hipEventCreate(&g_t0));
    RSVD_HIP(hipEventCreate(&g_t1));
  }
  RSVD_HIP(hipEventRecord(g_t0, api.c->stream));
  return 0;
}

double rsvd_timer_stop_ms(void) {
  Api api;
  if (api.rc || !g_t0) return -1.0;
  float ms = 0;
  if (hipEventRecord(g_t1, api.c->stream) != hipSuccess) return -1.0;
  if (hipEventSynchronize(g_t1) != hipSuccess) return -1.0;
  if (hipEventElapsedTime(&ms, g_t0, g_t1) != hipSuccess) return -1.0;
  return ms;
}

static void kstats_resolve(Context* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& pe : c->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
      c->stats[pe.idx].total_ms += ms;
      c->stats[pe.idx].count += 1;
    }
    c->event_pool.push_back(pe.a);
    c->event_pool.push_back(pe.b);
  }
  c->pending.clear();
}

void rsvd_set_fp64_vendor_gemm(int on) {
  Api api;
  if (api.rc) return;
  api.c->fp64_vendor = on != 0;
}

void rsvd_kstats_enable(int on) {
  Api api;
  if (api.rc) return;
  kstats_resolve(api.c);
  api.c->kstats = on != 0;
}
void rsvd_kstats_select(const char* names) {
  Api api;
  if (api.rc) return;
  kstats_resolve(api.c);
  api.c->kstats_only.clear();
  if (!names) return;
  std::string cur;
  for (const char* q = names;; ++q) {
    if (*q == ',' || *q == 0) {
      if (!cur.empty()) api.c->kstats_only.push_back(cur);
      cur.clear();
      if (*q == 0) break;
    } else {
      cur.push_back(*q);
    }
  }
}
void rsvd_kstats_reset(void) {
  Api api;
  if (api.rc) return;
  kstats_resolve(api.c);
  api.c->stats.clear();
}
int rsvd_kstats_get(const char* name, double* total_ms, long long* count) {
  Api api;
  if (api.rc) return api.rc;
  kstats_resolve(api.c);
  for (auto& s : api.c->stats)
    if (s.name == name) {
      if (total_ms) *total_ms = s.total_ms;
      if (count) *count = s.count;
      return 0;
    }
  return 1;
}

}  // extern "C"
