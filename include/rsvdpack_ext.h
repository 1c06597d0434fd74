/*
 * rsvdpack_ext.h -- additive extensions of librsvdpack beyond the RSVDPACK API.
 *
 * None of these has a counterpart in the reference (which returns void,
 * seeds Omega from time(NULL) and keeps every matrix in host memory); they
 * exist for callers that need errors, reproducible Omega, data already
 * resident in HBM, lower-precision sketching, or timing.  All functions
 * returning int give 0 on success and nonzero on failure, with the reason in
 * rsvd_last_error().  Device pointers are HIP device allocations (e.g. from
 * rsvd_dev_alloc or hipMalloc) on the current device; matrices are column-major.
 * A NaN or Inf in the input matrix makes the device decompositions
 * (rsvd_dev_fixed_rank*, rsvd_dev_blockrand*) return nonzero ("non-finite
 * singular values"); the reference's host API still returns its outputs and
 * sets rsvd_last_error() with the same message.
 */
#ifndef RSVDPACK_EXT_H
#define RSVDPACK_EXT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failure on the calling thread ("" if none). */
const char *rsvd_last_error(void);
/* Optional explicit initialisation of the device context (lazy otherwise). */
int rsvd_init(void);
/* Library version string. */
const char *rsvd_version(void);

/* ---- Omega ---- */
/* Fix the seed of every later Gaussian draw (initialize_random_matrix and the
 * Omega / RN matrices inside the decompositions).  Each draw restarts the
 * stream at element 0, as the reference's time(NULL) reseeding does within one
 * second. */
void rsvd_set_seed(unsigned long long seed);
/* Return to the reference behaviour: seed = time(NULL) at each draw. */
void rsvd_use_time_seed(void);
/* Inject Omega: later draws of `count` or fewer elements copy the first
 * elements of `host` (column-major) instead of sampling.  NULL clears. */
int rsvd_set_omega(const double *host, long long count);
/* Elements [offset, offset+count) of the Gaussian stream for `seed`, generated
 * on the GPU and copied to host. */
int rsvd_gaussian(unsigned long long seed, long long offset, long long count, double *host_out);

/* ---- binary matrix files ----
 * The reference's file layout: two integers (rows, columns), then the
 * entries row after row.  format 0: int32 header, double entries (MKL/mvf.c:77-133,
 * make_matrix_binary.m:32-41; what matrix_load_from_binary_file reads);
 * format 1: int64 header, double entries
 * (multi_core_mkl_code_64bit/matrix_vector_functions_intel_mkl.c:78-101);
 * format 2: int64 header, float entries
 * (oneapi_code/matrix_vector_functions_one_api.c:71-95). */
/* Rows and columns from the header of a file in `format`. */
int rsvd_binary_file_shape(const char *fname, int format, long long *m, long long *n);
/* Loads a file straight into a device matrix (column-major, lda >= rows), as
 * double (dtype 0) or float (dtype 1): blocks of rows are read into two
 * pinned host buffers in turn, copied to HBM and transposed on the device, the
 * read of one block overlapping the copy and transpose of the previous one.
 * Float entries widen exactly to double; double entries round to float for
 * dtype 1. */
int rsvd_dev_load_binary(const char *fname, int format, int dtype, void *dA, long long lda);
/* Writes an m x n device matrix (double: dtype 0, float: 1; column-major,
 * lda) to a file in `format` (double entries round to float for format 2). */
int rsvd_dev_write_binary(const char *fname, int format, int dtype, const void *dA, long long m,
                          long long n, long long lda);

/* ---- device memory helpers ---- */
void *rsvd_dev_alloc(size_t bytes);
void rsvd_dev_free(void *p);
int rsvd_memcpy_h2d(void *dst_dev, const void *src_host, size_t bytes);
int rsvd_memcpy_d2h(void *dst_host, const void *src_dev, size_t bytes);
int rsvd_memset_dev(void *dst_dev, int value, size_t bytes);
/* Fill an m x n device matrix with the seeded Gaussian stream (double). */
int rsvd_dev_gaussian(unsigned long long seed, long long m, long long n, double *dst, long long ld);
/* Waits for all work the library queued. */
int rsvd_sync(void);
/* Releases the library's cached device buffers. */
void rsvd_release_cache(void);

/* ---- device-resident decompositions (FP64) ---- */
/* low_rank_svd_rand_decomp_fixed_rank on a device matrix A (m x n, lda).
 * Omega: dOmega (n x (k+p), ld n) if non-NULL, else the seeded stream.
 * Outputs (caller-allocated): U m x k (ldu), S k values in the reference order
 * (vnum == 2 ascending, else descending), V n x k (ldv). */
int rsvd_dev_fixed_rank(const double *dA, long long m, long long n, long long lda, int k, int p,
                        int vnum, int q, int s, unsigned long long seed, const double *dOmega,
                        double *dU, long long ldu, double *dS, double *dV, long long ldv);

/* rsvd_dev_fixed_rank without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory during a call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, arithmetic and results as rsvd_dev_fixed_rank;
 * every buffer the computation needs beyond the caller's inputs and outputs
 * comes from one static device array of 256 MiB that the library holds on the
 * device from its load until the process ends.  Problems whose buffers exceed
 * it (roughly 8 (k + p) (3 n + 2 m) bytes plus O((k + p)^2) for the small
 * solves) are rejected with an error.  Every call on a device shares that
 * device's array and finds the previous call's contents in it (nothing there
 * is assumed zero); two calls that overlapped on one device would race.
 * Within a process they cannot: calls are serialised by the library's lock,
 * run on its one stream and return only when finished. */
int rsvd_dev_fixed_rank_noalloc(const double *dA, long long m, long long n, long long lda, int k,
                                int p, int vnum, int q, int s, unsigned long long seed,
                                const double *dOmega, double *dU, long long ldu, double *dS,
                                double *dV, long long ldv);

/* Mixed-precision fixed-rank rSVD of a float32 device matrix A (m x n, lda).
 * precision 1: sketch and power-iteration passes on the bf16 matrix cores
 * (fp32 accumulation); 2: the same passes in split-bf16 (16-bit operands, about
 * 1e-7 relative error on long sums); 3:
 * the same passes in fp8 e4m3 (A scaled per 128 columns, the skinny operand
 * per column; fp32 accumulation); 4 / 5: every pass but the last in bf16 /
 * fp8, the last power-iteration pass (whose range becomes Q) in split-bf16;
 * 6: every pass over A, the projection B^T = A^T Q included, on the fp16
 * matrix cores with A divided by a per-128-column scale and rounded to fp16
 * and the skinny operand in two fp16 planes (hi + lo, per-column scales): two
 * products per product, fp32 accumulation (about 2e-6 relative error of
 * sigma at 65536 x 16384, k = 512).  Otherwise: CholeskyQR with FP64
 * Cholesky, split-bf16 projection B^T = A^T Q, FP64 BB^T and eigen/SVD solve.  Omega: dOmega (n x (k+p) float, ld n) or the
 * seeded stream rounded to float.  Outputs: dU m x k float (ldu), dS k doubles
 * (reference order), dV n x k float (ldv). */
int rsvd_dev_fixed_rank_mp(const float *dA, long long m, long long n, long long lda, int k, int p,
                           int vnum, int q, int s, unsigned long long seed, const float *dOmega,
                           int precision, float *dU, long long ldu, double *dS, float *dV,
                           long long ldv);

/* randQB_pb_new (is_new = 1, MKL/rra.c:1576-1801) / randQB_pb (is_new = 0,
 * rra.c:1425-1569) on a float32 device matrix A in mixed precision (the QB of
 * rsvd_dev_blockrand_mp below).  dQ: m x rsvd_qb_max_cols float (ldq); dBt:
 * B^T, n x rsvd_qb_max_cols float (ldbt), i.e. B stored by rows.  *frank =
 * columns produced; *final_norm = the FP64 Frobenius norm of the deflated
 * float32 copy of A after the last block. */
int rsvd_dev_randqb_mp(const float *dA, long long m, long long n, long long lda, int kstep,
                       int nstep, double TOL, int q, int s, int is_new, unsigned long long seed,
                       const float *dRN, int precision, float *dQ, long long ldq, float *dBt,
                       long long ldbt, int *frank, double *final_norm);

/* low_rank_svd_blockrand_decomp_fixed_rank_or_prec (MKL/rra.c:239-381, its QB
 * randQB_pb_new rra.c:1576-1801) on a float32 device matrix A, in mixed
 * precision: a float32 working copy of A is deflated block by block (R -= Qp
 * Bp with fp32-exact products, Bp = Qp^T R in split-bf16 with its error
 * cleared on the next block) with the FP64 sum of squares of R as the TOL
 * norm; the
 * sketch and power-iteration passes over R run at `precision` (1 bf16, 2
 * split-bf16, 3 fp8 e4m3, as rsvd_dev_fixed_rank_mp); each Qp is
 * orthonormalised in FP64 and re-orthogonalised against every earlier block;
 * the final SVD as rsvd_dev_fixed_rank_mp.  k, p, TOL, vnum, kstep, q, s as
 * rsvd_dev_blockrand.  dRN: n x (kstep*nstep) float or NULL (the seeded
 * stream rounded to float).  Outputs allocated by the library (free with
 * rsvd_dev_free): *dU m x frank float, *dS frank doubles, *dV n x frank
 * float. */
int rsvd_dev_blockrand_mp(const float *dA, long long m, long long n, long long lda, int k, int p,
                          double TOL, int vnum, int kstep, int q, int s, unsigned long long seed,
                          const float *dRN, int precision, int *frank, float **dU, double **dS,
                          float **dV);

/* ||A - U diag(S) V^T||_F for a double device matrix A and device factors U
 * (m x k), S (k), V (n x k) stored as double (uv_f32 = 0) or float (1). */
int rsvd_dev_lowrank_residual(const double *dA, long long m, long long n, long long lda,
                              long long k, const void *dU, long long ldu, const double *dS,
                              const void *dV, long long ldv, int uv_f32, double *out);

/* Columns that randQB_pb (is_new = 0) / randQB_pb_new (is_new = 1) may
 * produce for these arguments (after the kstep clamp / TOL-mode nstep). */
long long rsvd_qb_max_cols(long long m, long long n, int kstep, int nstep, int is_new);

/* randQB_pb / randQB_pb_new on a device matrix.  dQ: m x rsvd_qb_max_cols
 * (ldq), dB: rsvd_qb_max_cols x n (ldb).  *frank = columns produced. */
int rsvd_dev_randqb(const double *dA, long long m, long long n, long long lda, int kstep,
                    int nstep, double TOL, int q, int s, int is_new, unsigned long long seed,
                    const double *dRN, double *dQ, long long ldq, double *dB, long long ldb,
                    int *frank, double *final_norm);

/* low_rank_svd_blockrand_decomp_fixed_rank_or_prec on a device matrix.
 * Outputs are allocated by the library (free with rsvd_dev_free):
 * *dU m x frank, *dS frank values, *dV n x frank. */
int rsvd_dev_blockrand(const double *dA, long long m, long long n, long long lda, int k, int p,
                       double TOL, int vnum, int kstep, int q, int s, unsigned long long seed,
                       const double *dRN, int *frank, double **dU, double **dS, double **dV);

/* C = alpha*op(A)*op(B) + beta*C on device matrices; ta / tb = 'N' or 'T'.
 * dtype 0: double; 1: float (fp32 matrix cores); 2: float storage, products on
 * the bf16 matrix cores with fp32 accumulation; 3: float storage, split-bf16
 * (hi*hi + hi*lo + lo*hi of 16-bit operand splits, fp32 accumulation); 4: float
 * storage, op(A) in fp8
 * e4m3 with one scale per 128 columns of A and B in e4m3 with one scale per
 * column, fp8 matrix cores, fp32 accumulation (tb = 'N' only); 5: float
 * storage, fp32-exact products from three bf16 planes per operand (six bf16
 * products), fp32 accumulation; 6: float storage, op(A) / a_c in fp16 (a_c =
 * max |A| over each 128-column chunk of A / 32768) times B in two fp16 planes
 * (hi + lo of B a_c / b_n, b_n = the column's max / 32768), two fp16 products,
 * fp32 accumulation (tb = 'N' only). */
int rsvd_dev_gemm(int dtype, char ta, char tb, long long M, long long N, long long K, double alpha,
                  const void *A, long long lda, const void *B, long long ldb, double beta, void *C,
                  long long ldc);

/* Batched fixed-rank rSVD (vnum = 2 semantics) of `count` independent float32
 * device matrices.  Matrix i is m[i] x n[i], column-major with lda = m[i], at
 * dA + a_off[i] (elements); rank k[i], oversampling p, q as in
 * low_rank_svd_rand_decomp_fixed_rank, with 1 <= k[i], k[i] + p <= min(m[i],
 * n[i]) and k[i] + p <= 48.  Omega_i is the seeded Gaussian stream with seed
 * seed + i (the same Omega rsvd_dev_fixed_rank draws with that seed).
 * Outputs per matrix: dU + u_off[i] (m[i] x k[i] float, ld m[i]), dS + s_off[i]
 * (k[i] doubles, ascending as vnum = 2 returns them), dV + v_off[i] (n[i] x
 * k[i] float, ld n[i]).  The passes over A run on the split-bf16 matrix cores
 * (fp32 accuracy); Gram matrices, Cholesky factors and the small eigenproblems
 * in FP64.  The first call for a batch layout tunes the kernels on that batch
 * and captures the call as a HIP graph; later calls with the same layout and
 * pointers replay it.  *status (optional) is nonzero when a Cholesky pivot of
 * some matrix was not positive (its sketch was numerically rank deficient).
 * Host arrays; device buffers as elsewhere. */
int rsvd_dev_fixed_rank_batched(int count, const int *m, const int *n, const int *k, int p, int q,
                                unsigned long long seed, const float *dA, const long long *a_off,
                                float *dU, const long long *u_off, double *dS,
                                const long long *s_off, float *dV, const long long *v_off,
                                int *status);

/* Kernel variants the batched autotuner picked for the most recent new batch
 * layout, with the timings it measured (a static string). */
const char *rsvd_batched_last_config(void);

/* Test batch for the batched path: matrix i (m[i] x n[i], float, at dA +
 * a_off[i], lda m[i]) = U diag(sigma) V^T with Haar U, V from the seeded
 * stream (seed + i) and sigma_j = exp(-j / tau[i]) for j < min(m, n, 19 tau),
 * zero beyond (below float resolution relative to sigma_0 = 1). */
int rsvd_dev_batch_test_matrices(int count, const int *m, const int *n, const double *tau,
                                 unsigned long long seed, float *dA, const long long *a_off);

/* ---- learned (q, rank) front end (no counterpart in the reference) ----
 * A 64-128-128-128-6 perceptron (GELU; fp32 parameters packed W1 b1 W2 b2 W3 b3
 * W4 b4, weights [out][in] row-major) maps the 64 sketch features of a matrix
 * to logits of q in {0..4} (outputs 0..4) and rank / min(m, n) (output 5). */
int rsvd_mlp_num_params(void);
/* He initialisation from the seeded stream (biases zero). */
int rsvd_dev_mlp_init(unsigned long long seed, float *dTheta);
/* dOut (batch x 6, row-major) = network(dX) for dX batch x 64 (row-major). */
int rsvd_dev_mlp_forward(const float *dTheta, const float *dX, int batch, float *dOut);
/* Gradient of the mean loss logsumexp(o[0:5]) - o[q] + lam (o[5] - r)^2 over
 * the batch into dGrad (num_params); per-sample losses into dLoss (batch). */
int rsvd_dev_mlp_grad(const float *dTheta, const float *dX, const int *dq, const float *dr,
                      float lam, int batch, float *dGrad, float *dLoss);
/* Adam training over shuffled mini-batches of the n samples; loss_hist (host,
 * optional) receives the mean loss of each epoch. */
int rsvd_dev_mlp_train(float *dTheta, const float *dX, const int *dq, const float *dr, int n,
                       float lam, int epochs, int batch, float lr, unsigned long long seed,
                       float *loss_hist);
/* The 64 features of a matrix (m, n >= 64; dA FP64, dA32 its float copy,
 * same lda): Y = A32 Omega (64 seeded Gaussian columns, split-bf16 products),
 * Q its orthonormal basis, lam_0 >= ... >= lam_63 the squared singular values
 * of B = Q^T A in FP64, a2 = ||A||_F^2; x_j = log10(t_j / a2) / 8 with
 * t_j = a2 - (lam_0 + ... + lam_j) (floored at 1e-16 a2): the energy the
 * best rank-(j + 1) approximation inside the sketch leaves. */
int rsvd_dev_sketch_features(const double *dA, const float *dA32, long long m, long long n,
                             long long lda, unsigned long long seed, float *dFeat);
/* TOL-driven low-rank approximation with the learned front end: predicts
 * (q, rank) from the sketch features of dA (dA32 its float copy), runs one
 * FP64 fixed-rank rSVD of dA (vnum = 2, oversampling p, q_pred + 1, i.e. q_pred
 * power iterations, k = ceil(rank (1 + margin)) + 2, rank = min(m, n)^o_5 from
 * the network's output o_5 clamped to [0, 1]), and accepts it when its
 * FP64 residual ||A - U S V^T||_F = sqrt(||A||_F^2 - sum sigma^2) is at most
 * tol_rel ||A||_F.  A predicted k at the sketch limit (k + p >= min(m, n)) is a
 * prediction of full rank; for m >= n it is solved as one: k = n, p = 0,
 * Omega = I (an exact thin QR of A and the small eigenproblem), its residual
 * formed explicitly.  A result above TOL is retried once with
 * k = ceil(1.5 k) + 8 (same rules); a second miss falls back to
 * low_rank_svd_blockrand_decomp_fixed_rank_or_prec in TOL mode (TOL = tol_rel
 * ||A||_F, kstep, q_fallback) and returns that method's result; when that too
 * misses TOL (TOL not reachable below full rank) and m >= n, the exact
 * full-rank decomposition (k = n, Omega = I) is returned instead, so no result
 * exceeds TOL.  Outputs as rsvd_dev_blockrand; *q_pred, *k_pred the
 * prediction, *fell_back 0 when the first attempt is returned, 3 when the
 * retry is, 1 when the adaptive fallback's result is, 2 when the exact one
 * is; *resid_rel the relative residual of the result. */
int rsvd_dev_learned_lowrank(const float *dTheta, const double *dA, const float *dA32,
                             long long m, long long n, long long lda, double tol_rel, int p,
                             double margin, int kstep, int q_fallback, unsigned long long seed,
                             int *frank, double **dU, double **dS, double **dV, int *q_pred,
                             int *k_pred, int *fell_back, double *resid_rel);

/* Plain device-to-device copy of `bytes` bytes (a multiple of 16; both
 * pointers 16-byte aligned) by a streaming kernel: exactly `bytes` read and
 * written.  Used to calibrate HBM byte counters against the library's kernels. */
int rsvd_dev_copy(const void *src, void *dst, long long bytes);

/* Synthetic test matrix in HBM: A = U diag(sigma) V^T + noise_std * N(0,1)
 * with Haar U (m x r), V (n x r) drawn from the seeded stream; sigma is a host
 * array of r values.  dA64 (double) and / or dA32 (float) may be NULL; when
 * both are given dA64 holds exactly the float-rounded values. */
int rsvd_dev_lowrank_test_matrix(long long m, long long n, int r, const double *sigma,
                                 double noise_std, unsigned long long seed, double *dA64,
                                 long long lda64, float *dA32, long long lda32);

/* Eigendecomposition of a symmetric positive semidefinite device matrix G
 * (n x n, upper triangle used): eigenvalues ascending into dEvals, orthonormal
 * eigenvectors into dEvecs (n x n, ldv); *sweeps = Jacobi sweeps used.  This is
 * the solver behind the vnum = 2 path. */
int rsvd_dev_sym_eig_psd(long long n, const double *dG, long long ldg, double *dEvals,
                         double *dEvecs, long long ldv, int *sweeps);

/* QR_factorization_getQ / compact_QR_factorization (MKL/mvf.c:1214-1263) with
 * the reference's own arithmetic, LAPACK dgeqrf + dorgqr: a blocked
 * Householder QR of the device matrix Y (m x n, ldy, m >= n), Y overwritten
 * by Q (LAPACK's reflectors and sign convention: R(i, i) = -sign(alpha_i)
 * times the norm, so Q and R equal LAPACK's to rounding), R (n x n, ld n,
 * upper; optional).  The default orthonormalisation of the library is
 * CholeskyQR2 (positive diagonal of R) and falls back to this routine. */
int rsvd_dev_householder_qr(double *dY, long long m, long long n, long long ldy, double *dR);

/* Routes the large FP64 products of the plain path (the passes over A: 10+
 * GFLOP, K >= 4096) through rocBLAS's dgemm (on = 1) instead of the library's
 * FP64 matrix-core GEMM (on = 0, the default): a vendor reference to time the
 * plain method against.  Results agree to rounding. */
void rsvd_set_fp64_vendor_gemm(int on);

/* ---- timing of the library's stream ---- */
int rsvd_timer_start(void);
/* Milliseconds since rsvd_timer_start, measured with HIP events on the
 * library stream (synchronises). */
double rsvd_timer_stop_ms(void);
/* Per-kernel event timing (off by default). */
void rsvd_kstats_enable(int on);
void rsvd_kstats_reset(void);
/* Restrict the per-kernel timing to the comma-separated families in `names`
 * (NULL or "": all), so a timed region carries events only around the
 * launches it reports. */
void rsvd_kstats_select(const char *names);
/* Total milliseconds and launch count of kernel family `name` (e.g.
 * "gemm_f64", "jacobi_round"); returns nonzero if never launched. */
int rsvd_kstats_get(const char *name, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif

#endif /* RSVDPACK_EXT_H */
