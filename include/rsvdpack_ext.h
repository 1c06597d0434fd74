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

/* Mixed-precision fixed-rank rSVD of a float32 device matrix A (m x n, lda).
 * precision 1: sketch and power-iteration passes on the bf16 matrix cores
 * (fp32 accumulation); 2: the same passes in split-bf16 (fp32 accuracy).
 * Always: CholeskyQR with FP64 Cholesky, split-bf16 projection B^T = A^T Q,
 * FP64 BB^T and eigen/SVD solve.  Omega: dOmega (n x (k+p) float, ld n) or the
 * seeded stream rounded to float.  Outputs: dU m x k float (ldu), dS k doubles
 * (reference order), dV n x k float (ldv). */
int rsvd_dev_fixed_rank_mp(const float *dA, long long m, long long n, long long lda, int k, int p,
                           int vnum, int q, int s, unsigned long long seed, const float *dOmega,
                           int precision, float *dU, long long ldu, double *dS, float *dV,
                           long long ldv);

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
 * (hi*hi + hi*lo + lo*hi, about fp32 accuracy). */
int rsvd_dev_gemm(int dtype, char ta, char tb, long long M, long long N, long long K, double alpha,
                  const void *A, long long lda, const void *B, long long ldb, double beta, void *C,
                  long long ldc);

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

/* ---- timing of the library's stream ---- */
int rsvd_timer_start(void);
/* Milliseconds since rsvd_timer_start, measured with HIP events on the
 * library stream (synchronises). */
double rsvd_timer_stop_ms(void);
/* Per-kernel event timing (off by default). */
void rsvd_kstats_enable(int on);
void rsvd_kstats_reset(void);
/* Total milliseconds and launch count of kernel family `name` (e.g.
 * "gemm_f64", "jacobi_round"); returns nonzero if never launched. */
int rsvd_kstats_get(const char *name, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif

#endif /* RSVDPACK_EXT_H */
