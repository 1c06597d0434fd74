/*
 * cpu_ref.c -- ORACLE (test infrastructure only, never shipped in the product).
 *
 * A clean-room CPU restatement of the RSVDPACK multi-core MKL path, used by the
 * tests, the smoke check and the CPU-baseline leg of the benchmark as the
 * checker.  Nothing in librsvdpack links, loads or calls this file.
 *
 * It follows the reference algorithm step by step and makes the same LAPACK /
 * BLAS calls (dgemm for every product, dgeqrf+dorgqr for every thin QR, dsyev
 * for the BB^T eigenproblem, dgesvd for the R-hat SVD).  The BLAS/LAPACK
 * library is resolved at run time: Intel MKL (libmkl_rt, the library the
 * reference is built against) when present, else the OpenBLAS bundled with
 * scipy.  Citations are relative to the RSVDPACK tree: rra.c =
 * multi_core_mkl_code/rank_revealing_algorithms_intel_mkl.c, mvf.c =
 * multi_core_mkl_code/matrix_vector_functions_intel_mkl.c.
 *
 * The reference seeds its Gaussian test matrix from time(NULL) (mvf.c:470), so
 * runs are not reproducible.  The oracle instead draws Omega from a seeded
 * counter-based generator (Philox4x64-10, Salmon et al. SC'11) mapped through
 * the inverse normal CDF (Wichura, AS241 PPND16), rounded to float and widened
 * to double as the reference's float VSL output is; or takes an injected Omega.
 *
 * Pinning (see tests/test_oracle_*.py): Philox outputs against numpy's
 * Philox4x64 bit generator; the inverse CDF against scipy.special.ndtri; the
 * decompositions against closed forms (exact recovery of a rank-r matrix,
 * k+p = min(m,n) giving the full SVD, the projector identity U S V^T =
 * P_K M with K = (M M^T)^(q-1) M Omega, Q^T Q = I, Q B = P_Q M), against
 * Eckart-Young lower bounds, and hand-derived control-flow quirks (blockrand
 * nstep/frank arithmetic, kstep clamp, TOL-mode stop step).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#define OMIN(x, y) (((x) < (y)) ? (x) : (y))
#define OMAX(x, y) (((x) > (y)) ? (x) : (y))

typedef struct {
    int nrows, ncols;
    double *d;
} mat;

typedef struct {
    int nrows;
    double *d;
} vec;

/* ------------------------------------------------------------------ */
/* BLAS / LAPACK resolved at run time                                  */
/* ------------------------------------------------------------------ */
typedef void (*dgemm_t)(const char *, const char *, const int *, const int *, const int *,
                        const double *, const double *, const int *, const double *,
                        const int *, const double *, double *, const int *, size_t, size_t);
typedef void (*dgeqrf_t)(const int *, const int *, double *, const int *, double *, double *,
                         const int *, int *);
typedef void (*dorgqr_t)(const int *, const int *, const int *, double *, const int *,
                         const double *, double *, const int *, int *);
typedef void (*dsyev_t)(const char *, const char *, const int *, double *, const int *,
                        double *, double *, const int *, int *, size_t, size_t);
typedef void (*dgesvd_t)(const char *, const char *, const int *, const int *, double *,
                         const int *, double *, double *, const int *, double *, const int *,
                         double *, const int *, int *, size_t, size_t);
typedef void (*setthreads_t)(int);
typedef void (*dgeqp3_t)(const int *, const int *, double *, const int *, int *, double *,
                         double *, const int *, int *);
typedef void (*dtrsm_t)(const char *, const char *, const char *, const char *, const int *,
                        const int *, const double *, const double *, const int *, double *,
                        const int *, size_t, size_t, size_t, size_t);
typedef void (*dgesv_t)(const int *, const int *, double *, const int *, int *, double *,
                        const int *, int *);
/* single precision (the oneAPI variant, see the FP32 section below) */
typedef void (*sgemm_t)(const char *, const char *, const int *, const int *, const int *,
                        const float *, const float *, const int *, const float *, const int *,
                        const float *, float *, const int *, size_t, size_t);
typedef void (*sgeqrf_t)(const int *, const int *, float *, const int *, float *, float *,
                         const int *, int *);
typedef void (*sorgqr_t)(const int *, const int *, const int *, float *, const int *,
                         const float *, float *, const int *, int *);
typedef void (*ssyev_t)(const char *, const char *, const int *, float *, const int *, float *,
                        float *, const int *, int *, size_t, size_t);
typedef void (*sgesvd_t)(const char *, const char *, const int *, const int *, float *,
                         const int *, float *, float *, const int *, float *, const int *,
                         float *, const int *, int *, size_t, size_t);

static dgemm_t p_dgemm;
static dgeqrf_t p_dgeqrf;
static dorgqr_t p_dorgqr;
static dsyev_t p_dsyev;
static dgesvd_t p_dgesvd;
static setthreads_t p_setthreads;
static dgeqp3_t p_dgeqp3;
static dtrsm_t p_dtrsm;
static dgesv_t p_dgesv;
static sgemm_t p_sgemm;
static sgeqrf_t p_sgeqrf;
static sorgqr_t p_sorgqr;
static sgesvd_t p_sgesvd;
static ssyev_t p_ssyev;
static int g_backend = 0; /* 1 = MKL, 2 = OpenBLAS */
static char g_backend_name[512];

static void *try_open(const char *path) { return dlopen(path, RTLD_NOW | RTLD_LOCAL); }

static int bind_prefix(void *h, const char *pre) {
    char nm[64];
#define BIND(var, type, base)                                     \
    snprintf(nm, sizeof nm, "%s%s", pre, base);                   \
    var = (type)dlsym(h, nm);                                     \
    if (!var) return 0;
    BIND(p_dgemm, dgemm_t, "dgemm_");
    BIND(p_dgeqrf, dgeqrf_t, "dgeqrf_");
    BIND(p_dorgqr, dorgqr_t, "dorgqr_");
    BIND(p_dsyev, dsyev_t, "dsyev_");
    BIND(p_dgesvd, dgesvd_t, "dgesvd_");
    BIND(p_dgeqp3, dgeqp3_t, "dgeqp3_");
    BIND(p_dtrsm, dtrsm_t, "dtrsm_");
    BIND(p_dgesv, dgesv_t, "dgesv_");
    BIND(p_sgemm, sgemm_t, "sgemm_");
    BIND(p_sgeqrf, sgeqrf_t, "sgeqrf_");
    BIND(p_sorgqr, sorgqr_t, "sorgqr_");
    BIND(p_sgesvd, sgesvd_t, "sgesvd_");
    BIND(p_ssyev, ssyev_t, "ssyev_");
#undef BIND
    return 1;
}

/* Returns 1 for MKL, 2 for OpenBLAS, 0 when no library could be bound.
 * extra_dirs: optional ':'-separated list of directories holding
 * libscipy_openblas*.so (passed by the Python loader). */
int oref_init(int nthreads, const char *openblas_path) {
    if (g_backend) {
        if (p_setthreads && nthreads > 0) p_setthreads(nthreads);
        return g_backend;
    }
    const char *mkl_paths[] = {"/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
                               "libmkl_rt.so.1", "libmkl_rt.so", NULL};
    for (int i = 0; mkl_paths[i]; ++i) {
        void *h = try_open(mkl_paths[i]);
        if (h && bind_prefix(h, "")) {
            /* LP64 interface, GNU OpenMP threading layer. */
            typedef int (*seti_t)(int);
            seti_t set_if = (seti_t)dlsym(h, "MKL_Set_Interface_Layer");
            seti_t set_tl = (seti_t)dlsym(h, "MKL_Set_Threading_Layer");
            if (set_if) set_if(0);
            if (set_tl) set_tl(1);
            p_setthreads = (setthreads_t)dlsym(h, "MKL_Set_Num_Threads");
            g_backend = 1;
            snprintf(g_backend_name, sizeof g_backend_name, "mkl:%s", mkl_paths[i]);
            break;
        }
    }
    if (!g_backend && openblas_path && openblas_path[0]) {
        void *h = try_open(openblas_path);
        if (h && bind_prefix(h, "scipy_")) {
            p_setthreads = (setthreads_t)dlsym(h, "scipy_openblas_set_num_threads");
            g_backend = 2;
            snprintf(g_backend_name, sizeof g_backend_name, "openblas:%s", openblas_path);
        } else if (h && bind_prefix(h, "")) {
            p_setthreads = (setthreads_t)dlsym(h, "openblas_set_num_threads");
            g_backend = 2;
            snprintf(g_backend_name, sizeof g_backend_name, "openblas:%s", openblas_path);
        }
    }
    if (g_backend && p_setthreads && nthreads > 0) p_setthreads(nthreads);
    return g_backend;
}

const char *oref_backend(void) { return g_backend_name; }

/* ------------------------------------------------------------------ */
/* Gaussian test matrices                                              */
/* ------------------------------------------------------------------ */

/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11): multipliers 0xD2E7470EE14C6C93, 0xCA5A826395121157,
 * Weyl key increments 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B. */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo) {
    unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
    *lo = (uint64_t)p;
    *hi = (uint64_t)(p >> 64);
}

void oref_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n1 = lo1;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        uint64_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Natural log for x > 0, x finite: x = f * 2^e with f in [sqrt(1/2), sqrt(2)),
 * log f = 2 atanh(t), t = (f-1)/(f+1), |t| <= 0.1716, series to t^27. */
static double ref_log(double x) {
    int e;
    double f = frexp(x, &e); /* f in [0.5, 1) */
    if (f < 0.70710678118654752440) { f = f * 2.0; e -= 1; }
    double t = (f - 1.0) / (f + 1.0);
    double t2 = t * t;
    double s = 1.0 / 27.0;
    for (int j = 25; j >= 1; j -= 2) s = s * t2 + 1.0 / (double)j;
    double lf = 2.0 * t * s;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return ((double)e * ln2_lo + lf) + (double)e * ln2_hi;
}

/* Inverse standard normal CDF, Wichura (1988) AS241 PPND16. */
double oref_ndtri(double p) {
    double q = p - 0.5, r, x;
    if (fabs(q) <= 0.425) {
        r = 0.180625 - q * q;
        x = q * (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
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
    r = sqrt(-ref_log(r));
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

/* Element e of the stream for `seed` is lane e%4 of Philox block e/4 under key
 * (seed, 0) and counter (e/4, 0, 0, 0); the 64-bit word u becomes
 * ((u >> 11) + 0.5) * 2^-53 in (0,1), then ndtri, then float, then double. */
void oref_gaussian(uint64_t seed, int64_t offset, int64_t count, double *out) {
    uint64_t key[2] = {seed, 0};
    int64_t e = offset, end = offset + count;
    while (e < end) {
        uint64_t ctr[4] = {(uint64_t)(e >> 2), 0, 0, 0}, w[4];
        oref_philox4x64(ctr, key, w);
        for (int lane = (int)(e & 3); lane < 4 && e < end; ++lane, ++e) {
            double u = ((double)(w[lane] >> 11) + 0.5) * 0x1.0p-53;
            out[e - offset] = (double)(float)oref_ndtri(u);
        }
    }
}

/* Omega source: either an injected buffer or the seeded stream. */
static uint64_t g_seed = 777;
static const double *g_inj = NULL;
static int64_t g_inj_count = 0;

void oref_set_seed(uint64_t seed) { g_seed = seed; }
void oref_set_omega(const double *buf, int64_t count) { g_inj = buf; g_inj_count = count; }

/* Stands in for mvf.c:458-486 initialize_random_matrix. */
static int initialize_random_matrix(mat *M) {
    int64_t N = (int64_t)M->nrows * M->ncols;
    if (g_inj) {
        if (g_inj_count < N) return -1;
        memcpy(M->d, g_inj, (size_t)N * sizeof(double));
        return 0;
    }
    oref_gaussian(g_seed, 0, N, M->d);
    return 0;
}

/* ------------------------------------------------------------------ */
/* L1 restatement (mvf.c)                                              */
/* ------------------------------------------------------------------ */
mat *oref_matrix_new(int nrows, int ncols) { /* mvf.c:8-16 */
    mat *M = (mat *)malloc(sizeof(mat));
    M->d = (double *)calloc((size_t)nrows * (size_t)ncols + 1, sizeof(double));
    M->nrows = nrows;
    M->ncols = ncols;
    return M;
}
void oref_matrix_delete(mat *M) { /* mvf.c:30-35 */
    if (!M) return;
    free(M->d);
    free(M);
}
static vec *vector_new(int n) {
    vec *v = (vec *)malloc(sizeof(vec));
    v->d = (double *)calloc((size_t)n + 1, sizeof(double));
    v->nrows = n;
    return v;
}
static void vector_delete(vec *v) { free(v->d); free(v); }
#define MEL(M, i, j) ((M)->d[(size_t)(j) * (size_t)(M)->nrows + (size_t)(i)])

static void matrix_copy(mat *D, mat *S) { /* mvf.c:216-226 */
    memcpy(D->d, S->d, (size_t)S->nrows * S->ncols * sizeof(double));
}
static void matrix_sub(mat *A, mat *B) { /* mvf.c:274-283 */
    size_t N = (size_t)A->nrows * A->ncols;
    for (size_t i = 0; i < N; ++i) A->d[i] -= B->d[i];
}
double oref_frobenius(const double *d, int64_t N) { /* mvf.c:360-372 */
    double s = 0;
    for (int64_t i = 0; i < N; ++i) s += d[i] * d[i];
    return sqrt(s);
}
static double get_matrix_frobenius_norm(mat *M) {
    return oref_frobenius(M->d, (int64_t)M->nrows * M->ncols);
}
static void matrix_build_transpose(mat *Mt, mat *M) { /* mvf.c:246-253 */
    for (int i = 0; i < M->nrows; ++i)
        for (int j = 0; j < M->ncols; ++j) MEL(Mt, j, i) = MEL(M, i, j);
}

/* mvf.c:538-561 -- the three GEMM shapes, leading dimensions = nrows. */
static void gemm(char ta, char tb, int m, int n, int k, mat *A, mat *B, mat *C) {
    const double one = 1.0, zero = 0.0;
    int lda = OMAX(1, A->nrows), ldb = OMAX(1, B->nrows), ldc = OMAX(1, C->nrows);
    if (m == 0 || n == 0) return;
    p_dgemm(&ta, &tb, &m, &n, &k, &one, A->d, &lda, B->d, &ldb, &zero, C->d, &ldc, 1, 1);
}
static void matrix_matrix_mult(mat *A, mat *B, mat *C) {
    gemm('N', 'N', A->nrows, B->ncols, A->ncols, A, B, C);
}
static void matrix_transpose_matrix_mult(mat *A, mat *B, mat *C) {
    gemm('T', 'N', A->ncols, B->ncols, A->nrows, A, B, C);
}
static void matrix_matrix_transpose_mult(mat *A, mat *B, mat *C) {
    gemm('N', 'T', A->nrows, B->nrows, A->ncols, A, B, C);
}

/* column gather / scatter, row gather / scatter -- mvf.c:647-753 */
static void get_cols(mat *M, const int *inds, mat *Mc) {
    for (int j = 0; j < Mc->ncols; ++j)
        memcpy(&MEL(Mc, 0, j), &MEL(M, 0, inds[j]), (size_t)M->nrows * sizeof(double));
}
static void set_cols(mat *M, const int *inds, mat *Mc) {
    for (int j = 0; j < Mc->ncols; ++j)
        memcpy(&MEL(M, 0, inds[j]), &MEL(Mc, 0, j), (size_t)M->nrows * sizeof(double));
}
static void set_rows(mat *M, const int *inds, mat *Mr) {
    for (int i = 0; i < Mr->nrows; ++i)
        for (int j = 0; j < M->ncols; ++j) MEL(M, inds[i], j) = MEL(Mr, i, j);
}

/* mvf.c:757-768 upper-triangle copy */
static void matrix_copy_symmetric(mat *S, mat *M) {
    for (int i = 0; i < M->nrows; ++i)
        for (int j = i; j < M->ncols; ++j) MEL(S, i, j) = MEL(M, i, j);
}
/* mvf.c:490-499, 519-528 */
static void initialize_diagonal_matrix(mat *D, vec *data) {
    for (int i = 0; i < D->nrows; ++i) MEL(D, i, i) = data->d[i];
}
static void invert_diagonal_matrix(mat *Dinv, mat *D) {
    for (int i = 0; i < D->nrows; ++i) MEL(Dinv, i, i) = 1.0 / MEL(D, i, i);
}

/* mvf.c:1206-1209 dsyev('V','U') */
static void compute_evals_and_evecs_of_symm_matrix(mat *S, vec *evals) {
    int n = S->nrows, lda = OMAX(1, S->ncols), lwork = -1, info = 0;
    double wq;
    p_dsyev("V", "U", &n, S->d, &lda, evals->d, &wq, &lwork, &info, 1, 1);
    lwork = (int)wq + 1;
    double *work = (double *)malloc(sizeof(double) * (size_t)lwork);
    p_dsyev("V", "U", &n, S->d, &lda, evals->d, work, &lwork, &info, 1, 1);
    free(work);
}

static void geqrf(int m, int n, double *a, int lda, double *tau) {
    int lwork = -1, info = 0;
    double wq;
    p_dgeqrf(&m, &n, a, &lda, tau, &wq, &lwork, &info);
    lwork = (int)wq + 1;
    double *work = (double *)malloc(sizeof(double) * (size_t)lwork);
    p_dgeqrf(&m, &n, a, &lda, tau, work, &lwork, &info);
    free(work);
}
static void orgqr(int m, int n, int k, double *a, int lda, double *tau) {
    int lwork = -1, info = 0;
    double wq;
    p_dorgqr(&m, &n, &k, a, &lda, tau, &wq, &lwork, &info);
    lwork = (int)wq + 1;
    double *work = (double *)malloc(sizeof(double) * (size_t)lwork);
    p_dorgqr(&m, &n, &k, a, &lda, tau, work, &lwork, &info);
    free(work);
}

/* mvf.c:1214-1245 compact QR: R = triu(R_full(1:k,1:k)), Q = orgqr */
static void compact_QR_factorization(mat *M, mat *Q, mat *R) {
    int m = M->nrows, n = M->ncols, k = OMIN(m, n);
    mat *R_full = oref_matrix_new(m, n);
    matrix_copy(R_full, M);
    vec *tau = vector_new(k);
    geqrf(m, n, R_full->d, m, tau->d);
    for (int i = 0; i < k; ++i)
        for (int j = i; j < k; ++j) MEL(R, i, j) = MEL(R_full, i, j);
    matrix_copy(Q, R_full);
    orgqr(Q->nrows, Q->ncols, OMIN(Q->ncols, Q->nrows), Q->d, Q->nrows, tau->d);
    oref_matrix_delete(R_full);
    vector_delete(tau);
}

/* mvf.c:1251-1263 */
static void QR_factorization_getQ(mat *M, mat *Q) {
    int m = M->nrows, n = M->ncols, k = OMIN(m, n);
    matrix_copy(Q, M);
    vec *tau = vector_new(k);
    geqrf(m, n, Q->d, m, tau->d);
    orgqr(m, n, n, Q->d, m, tau->d);
    vector_delete(tau);
}

/* mvf.c:1270-1284 dgesvd('S','S') */
static void singular_value_decomposition(mat *M, mat *U, mat *S, mat *Vt) {
    int m = M->nrows, n = M->ncols, k = OMIN(m, n), lwork = -1, info = 0;
    int ldu = OMAX(1, m), ldvt = OMAX(1, k);
    vec *svals = vector_new(k);
    double wq;
    p_dgesvd("S", "S", &m, &n, M->d, &m, svals->d, U->d, &ldu, Vt->d, &ldvt, &wq, &lwork,
             &info, 1, 1);
    lwork = (int)wq + 1;
    double *work = (double *)malloc(sizeof(double) * (size_t)lwork);
    p_dgesvd("S", "S", &m, &n, M->d, &m, svals->d, U->d, &ldu, Vt->d, &ldvt, work, &lwork,
             &info, 1, 1);
    free(work);
    initialize_diagonal_matrix(S, svals);
    vector_delete(svals);
}

/* mvf.c:1078-1128 resizes (first / last columns or rows) */
static void resize_cols(mat **M, int k, int from_end) {
    mat *R = oref_matrix_new((*M)->nrows, k);
    int off = from_end ? (*M)->ncols - k : 0;
    for (int j = 0; j < k; ++j)
        memcpy(&MEL(R, 0, j), &MEL(*M, 0, off + j), (size_t)R->nrows * sizeof(double));
    oref_matrix_delete(*M);
    *M = R;
}
static void resize_rows(mat **M, int k, int from_end) {
    mat *R = oref_matrix_new(k, (*M)->ncols);
    int off = from_end ? (*M)->nrows - k : 0;
    for (int j = 0; j < R->ncols; ++j)
        for (int i = 0; i < k; ++i) MEL(R, i, j) = MEL(*M, off + i, j);
    oref_matrix_delete(*M);
    *M = R;
}

/* ------------------------------------------------------------------ */
/* L2 restatement (rra.c)                                              */
/* ------------------------------------------------------------------ */

/* Shared tail of rra.c:133-225 (fixed rank) and rra.c:288-381 (blockrand):
 * given Q (m x l), finish with QR-of-B^T (vnum==1 || vnum>2) or eig(BB^T). */
static void svd_from_Q(mat *M, mat *Q, int k, int vnum, mat **U, mat **S, mat **V) {
    int m = M->nrows, n = M->ncols, l = Q->ncols;
    *U = oref_matrix_new(m, l);
    *S = oref_matrix_new(l, l);
    *V = oref_matrix_new(n, l);
    if (vnum == 1 || vnum > 2) {
        mat *Bt = oref_matrix_new(n, l);
        matrix_transpose_matrix_mult(M, Q, Bt);           /* rra.c:139 */
        mat *Qhat = oref_matrix_new(n, l), *Rhat = oref_matrix_new(l, l);
        compact_QR_factorization(Bt, Qhat, Rhat);         /* rra.c:146 */
        mat *Uhat = oref_matrix_new(l, l), *Vhat_trans = oref_matrix_new(l, l);
        singular_value_decomposition(Rhat, Uhat, *S, Vhat_trans); /* rra.c:152 */
        matrix_matrix_transpose_mult(Q, Vhat_trans, *U);  /* rra.c:156 U = Q*Vhat */
        matrix_matrix_mult(Qhat, Uhat, *V);               /* rra.c:160 V = Qhat*Uhat */
        oref_matrix_delete(Rhat); oref_matrix_delete(Qhat); oref_matrix_delete(Uhat);
        oref_matrix_delete(Vhat_trans); oref_matrix_delete(Bt);
        resize_cols(U, k, 0); resize_cols(V, k, 0);       /* rra.c:170-173 */
        resize_cols(S, k, 0); resize_rows(S, k, 0);
    } else {
        mat *B = oref_matrix_new(l, n);
        matrix_transpose_matrix_mult(Q, M, B);            /* rra.c:181 B = Q^T M */
        mat *BBt = oref_matrix_new(l, l);
        matrix_matrix_transpose_mult(B, B, BBt);          /* rra.c:184 */
        vec *evals = vector_new(l);
        mat *Uhat = oref_matrix_new(l, l);
        matrix_copy_symmetric(Uhat, BBt);                 /* rra.c:190 */
        compute_evals_and_evecs_of_symm_matrix(Uhat, evals);
        vec *sv = vector_new(l);
        for (int i = 0; i < l; ++i) sv->d[i] = sqrt(evals->d[i]); /* rra.c:196-199, no clamp */
        initialize_diagonal_matrix(*S, sv);
        matrix_matrix_mult(Q, Uhat, *U);                  /* rra.c:204 */
        mat *Sinv = oref_matrix_new(l, l), *UhatSinv = oref_matrix_new(l, l);
        invert_diagonal_matrix(Sinv, *S);                 /* rra.c:211 */
        matrix_matrix_mult(Uhat, Sinv, UhatSinv);         /* rra.c:212 */
        matrix_transpose_matrix_mult(B, UhatSinv, *V);    /* rra.c:213 */
        oref_matrix_delete(BBt); oref_matrix_delete(Sinv); oref_matrix_delete(UhatSinv);
        oref_matrix_delete(Uhat); oref_matrix_delete(B);
        vector_delete(evals); vector_delete(sv);
        resize_cols(U, k, 1); resize_cols(V, k, 1);       /* rra.c:221-224 */
        resize_cols(S, k, 1); resize_rows(S, k, 1);
    }
}

/* rra.c:73-234 low_rank_svd_rand_decomp_fixed_rank */
int oref_fixed_rank(mat *M, int k, int p, int vnum, int q, int s, int *frank, mat **U, mat **S,
                    mat **V) {
    (void)frank; /* never written by the reference */
    int m = M->nrows, n = M->ncols, l = k + p;
    if (!g_backend) return -2;
    if (k <= 0 || l > OMIN(m, n) || s <= 0) return -1;
    mat *RN = oref_matrix_new(n, l);
    if (initialize_random_matrix(RN)) { oref_matrix_delete(RN); return -3; } /* rra.c:90-91 */
    mat *Y = oref_matrix_new(m, l);
    matrix_matrix_mult(M, RN, Y);                         /* rra.c:95 */
    mat *Z = oref_matrix_new(n, l), *Yorth = oref_matrix_new(m, l), *Zorth = oref_matrix_new(n, l);
    for (int j = 1; j < q; ++j) {                         /* rra.c:101: q-1 iterations */
        if ((2 * j - 2) % s == 0) {
            QR_factorization_getQ(Y, Yorth);
            matrix_transpose_matrix_mult(M, Yorth, Z);
        } else {
            matrix_transpose_matrix_mult(M, Y, Z);
        }
        if ((2 * j - 1) % s == 0) {
            QR_factorization_getQ(Z, Zorth);
            matrix_matrix_mult(M, Zorth, Y);
        } else {
            matrix_matrix_mult(M, Z, Y);
        }
    }
    mat *Q = oref_matrix_new(m, l);
    QR_factorization_getQ(Y, Q);                          /* rra.c:129-130 */
    svd_from_Q(M, Q, k, vnum, U, S, V);
    oref_matrix_delete(RN); oref_matrix_delete(Y); oref_matrix_delete(Q);
    oref_matrix_delete(Z); oref_matrix_delete(Yorth); oref_matrix_delete(Zorth);
    return 0;
}

/* Shared body of randQB_pb (rra.c:1425-1569, reortho_every = 1, no TOL) and
 * randQB_pb_new (rra.c:1576-1801, reortho_every = 2, TOL mode). */
static int qb_blocked(mat *M, int kstep, int nstep, double TOL, int q, int s, int *frank,
                      mat **Qo, mat **Bo, int is_new) {
    int m = M->nrows, n = M->ncols, tolMode = 0;
    if (is_new) {
        if (kstep > (int)(OMIN(m, n) / 2)) kstep = (int)OMIN(m, n) / 10;   /* rra.c:1589-1592 */
        if (nstep <= 0) { tolMode = 1; nstep = (int)(OMIN(m, n) / kstep); } /* rra.c:1595-1602 */
    }
    if (kstep <= 0 || nstep <= 0 || s <= 0) return -1;
    int l = kstep * nstep;
    mat *RN = oref_matrix_new(n, l);
    if (initialize_random_matrix(RN)) { oref_matrix_delete(RN); return -3; }
    mat *A = oref_matrix_new(m, n);
    *Qo = oref_matrix_new(m, l);
    *Bo = oref_matrix_new(l, n);
    mat *RNp = oref_matrix_new(n, kstep), *Yp = oref_matrix_new(m, kstep);
    mat *Qp = oref_matrix_new(m, kstep), *Bp = oref_matrix_new(kstep, n);
    mat *QpBp = oref_matrix_new(m, n), *AtQp = oref_matrix_new(n, kstep);
    mat *AtQp2 = oref_matrix_new(n, kstep), *QptA = oref_matrix_new(kstep, n);
    int *inds_local = (int *)malloc(sizeof(int) * (size_t)kstep);
    matrix_copy(A, M);
    int fr = 0;
    for (int step = 0; step < nstep; ++step) {
        for (int i = 0; i < kstep; ++i) inds_local[i] = kstep * step + i;
        get_cols(RN, inds_local, RNp);
        matrix_matrix_mult(A, RNp, Yp);                   /* rra.c:1644 */
        for (int j = 1; j <= q; ++j) {                    /* rra.c:1652: q iterations */
            if ((2 * j - 2) % s == 0) {
                QR_factorization_getQ(Yp, Qp);
                if (is_new) {                             /* rra.c:1657-1658 */
                    matrix_transpose_matrix_mult(Qp, A, QptA);
                    matrix_build_transpose(AtQp, QptA);
                } else {                                  /* rra.c:1477 */
                    matrix_transpose_matrix_mult(A, Qp, AtQp);
                }
            } else {
                matrix_transpose_matrix_mult(A, Yp, AtQp);
            }
            if ((2 * j - 1) % s == 0) {
                QR_factorization_getQ(AtQp, AtQp2);
                matrix_matrix_mult(A, AtQp2, Yp);
            } else {
                matrix_matrix_mult(A, AtQp, Yp);
            }
        }
        QR_factorization_getQ(Yp, Qp);                    /* rra.c:1690 */
        int reortho = is_new ? (step > 0 && step % 2 == 0) : (step > 0); /* rra.c:1703 / :1500 */
        if (reortho) {
            int nj = step * kstep;
            int *inds_global = (int *)malloc(sizeof(int) * (size_t)nj);
            for (int i = 0; i < nj; ++i) inds_global[i] = i;
            mat *Qj = oref_matrix_new(m, nj), *QjtQp = oref_matrix_new(nj, kstep);
            mat *QjQjtQp = oref_matrix_new(m, kstep);
            get_cols(*Qo, inds_global, Qj);
            matrix_transpose_matrix_mult(Qj, Qp, QjtQp);
            matrix_matrix_mult(Qj, QjtQp, QjQjtQp);
            matrix_copy(Yp, Qp);
            matrix_sub(Yp, QjQjtQp);
            QR_factorization_getQ(Yp, Qp);
            oref_matrix_delete(Qj); oref_matrix_delete(QjtQp); oref_matrix_delete(QjQjtQp);
            free(inds_global);
        }
        matrix_transpose_matrix_mult(Qp, A, Bp);          /* rra.c:1741 */
        matrix_matrix_mult(Qp, Bp, QpBp);                 /* rra.c:1750-1751 */
        matrix_sub(A, QpBp);
        set_cols(*Qo, inds_local, Qp);                    /* rra.c:1760-1761 */
        set_rows(*Bo, inds_local, Bp);
        fr = (step + 1) * kstep;                          /* rra.c:1770 */
        if (tolMode) {
            double normval = get_matrix_frobenius_norm(A);
            if (normval < TOL) break;
        }
    }
    if (tolMode) { resize_cols(Qo, fr, 0); resize_rows(Bo, fr, 0); } /* rra.c:1781-1788 */
    if (frank && is_new) *frank = fr;
    free(inds_local);
    oref_matrix_delete(A); oref_matrix_delete(RN); oref_matrix_delete(RNp);
    oref_matrix_delete(Yp); oref_matrix_delete(Qp); oref_matrix_delete(Bp);
    oref_matrix_delete(QpBp); oref_matrix_delete(AtQp); oref_matrix_delete(AtQp2);
    oref_matrix_delete(QptA);
    return 0;
}

int oref_randQB_pb(mat *M, int kstep, int nstep, int q, int s, mat **Q, mat **B) {
    if (!g_backend) return -2;
    return qb_blocked(M, kstep, nstep, 0.0, q, s, NULL, Q, B, 0);
}

int oref_randQB_pb_new(mat *M, int kstep, int nstep, double TOL, int q, int s, int *frank,
                       mat **Q, mat **B) {
    if (!g_backend) return -2;
    return qb_blocked(M, kstep, nstep, TOL, q, s, frank, Q, B, 1);
}

/* rra.c:239-381 low_rank_svd_blockrand_decomp_fixed_rank_or_prec */
int oref_blockrand(mat *M, int k, int p, double TOL, int vnum, int kstep, int q, int s,
                   int *frank, mat **U, mat **S, mat **V, int *l_out) {
    int m = M->nrows, n = M->ncols, rankMode, nstep;
    mat *Q = NULL, *B = NULL;
    if (!g_backend) return -2;
    if (kstep <= 0) return -1;
    if (k <= 0) { rankMode = 0; nstep = 0; }              /* rra.c:251-253 */
    else { rankMode = 1; }
    (void)nstep;
    if (p < kstep && (p + kstep) < OMIN(m, n)) p = kstep;  /* rra.c:260-262 */
    nstep = (int)ceil((double)((k + p) / kstep));         /* rra.c:263, integer division */
    int rc = oref_randQB_pb_new(M, kstep, nstep, TOL, q, s, frank, &Q, &B);
    if (rc) return rc;
    int l = B->nrows;
    if (rankMode == 1) *frank = k;                         /* rra.c:273-279 */
    else *frank = (int)round((double)(*frank / (*frank + p + 1e-6)) * (*frank));
    k = *frank;
    if (l_out) *l_out = l;
    if (k <= 0 || k > l) { oref_matrix_delete(Q); oref_matrix_delete(B); return -4; }
    svd_from_Q(M, Q, k, vnum, U, S, V);                   /* rra.c:288-380, re-reads M */
    oref_matrix_delete(Q);
    oref_matrix_delete(B);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Interpolative and CUR decompositions on top of the QB (rra.c)        */
/* ------------------------------------------------------------------ */

/* rra.c:924-975 pivotedQR_mkl: dgeqp3 on a copy of M (every column free: the
 * reference hands dgeqp3 an uninitialised pivot array, whose intended and
 * usual -- zero-filled -- content means exactly that), R = the upper
 * trapezoid of the first min(m, n) rows, Q = dorgqr, I = 0-based pivots. */
static void pivoted_qr(mat *M, mat **Q, mat **R, int *I) {
    int m = M->nrows, n = M->ncols, k = OMIN(m, n);
    mat *W = oref_matrix_new(m, n);
    matrix_copy(W, M);
    int *jpvt = (int *)calloc((size_t)n + 1, sizeof(int));
    double *tau = (double *)calloc((size_t)k + 1, sizeof(double));
    int lwork = -1, info = 0, lda = OMAX(1, m);
    double wq;
    p_dgeqp3(&m, &n, W->d, &lda, jpvt, tau, &wq, &lwork, &info);
    lwork = (int)wq + 1;
    double *work = (double *)malloc(sizeof(double) * (size_t)lwork);
    p_dgeqp3(&m, &n, W->d, &lda, jpvt, tau, work, &lwork, &info);
    free(work);
    int Rr = k, Rc = (m <= n) ? n : k, Qr = (m <= n) ? k : m, Qc = k;
    *R = oref_matrix_new(Rr, Rc);
    for (int j = 0; j < Rc; ++j)
        for (int i = 0; i <= OMIN(j, Rr - 1); ++i) MEL(*R, i, j) = MEL(W, i, j);
    if (Q) {
        orgqr(Qr, Qc, k, W->d, lda, tau);
        *Q = oref_matrix_new(Qr, Qc);
        for (int j = 0; j < Qc; ++j)
            for (int i = 0; i < Qr; ++i) MEL(*Q, i, j) = MEL(W, i, j);
    }
    for (int i = 0; i < n; ++i) I[i] = jpvt[i] - 1;
    free(jpvt); free(tau);
    oref_matrix_delete(W);
}

/* mvf.c:1477-1494 upper_triangular_system_solve, solve_type 1: X = A^-1 B by
 * dtrsm (left, upper, no transpose, non-unit). */
static void trsm_upper(mat *A, mat *B, mat *X) {
    const double one = 1.0;
    int m = B->nrows, n = B->ncols, lda = OMAX(1, A->nrows), ldb = OMAX(1, B->nrows);
    memcpy(X->d, B->d, (size_t)m * n * sizeof(double));
    if (m && n) p_dtrsm("L", "U", "N", "N", &m, &n, &one, A->d, &lda, X->d, &ldb, 1, 1, 1, 1);
}

/* T = Rk1^-1 Rk2 from the first f rows of R: Rk1 = R(0:f, 0:f) made upper
 * triangular, Rk2 = R(0:f, f:n) (rra.c:2005-2020, mvf.c:959-1010, 773-784). */
static mat *id_coefficients(mat *R, int f) {
    int n = R->ncols;
    mat *Rk1 = oref_matrix_new(f, f), *Rk2 = oref_matrix_new(f, n - f);
    for (int j = 0; j < f; ++j)
        for (int i = 0; i <= j; ++i) MEL(Rk1, i, j) = MEL(R, i, j);
    for (int j = f; j < n; ++j)
        for (int i = 0; i < f; ++i) MEL(Rk2, i, j - f) = MEL(R, i, j);
    mat *T = oref_matrix_new(f, n - f);
    trsm_upper(Rk1, Rk2, T);
    oref_matrix_delete(Rk1); oref_matrix_delete(Rk2);
    return T;
}

/* rra.c:1969-2027 id_blockrand_decomp_fixed_rank_or_prec: the QB of
 * randQB_pb_new with nstep = (k + p) / kstep (integer division, :1974; 0 in
 * TOL mode), the full pivoted QR of B, frank = k (rank mode) or the shrunk
 * round(frank / (frank + p + 1e-6) * frank) (TOL mode, :1998-2003), T from
 * the first frank rows of R.  I: n 0-based pivots. */
int oref_id_blockrand(mat *M, int k, int p, double TOL, int kstep, int q, int s, int *frank,
                      int *I, mat **T) {
    if (!g_backend) return -2;
    if (kstep <= 0) return -1;
    int nstep = (int)ceil((double)((k + p) / kstep)), rankMode = 1;
    if (k <= 0) { rankMode = 0; nstep = 0; }
    mat *Q = NULL, *B = NULL, *R = NULL;
    int rc = oref_randQB_pb_new(M, kstep, nstep, TOL, q, s, frank, &Q, &B);
    if (rc) return rc;
    pivoted_qr(B, NULL, &R, I);
    if (rankMode) *frank = k;
    else *frank = (int)round((double)(*frank / (*frank + p + 1e-6)) * (*frank));
    int f = *frank;
    if (f <= 0 || f > R->nrows || f > R->ncols) {
        oref_matrix_delete(Q); oref_matrix_delete(B); oref_matrix_delete(R);
        return -4;
    }
    *T = id_coefficients(R, f);
    oref_matrix_delete(Q); oref_matrix_delete(B); oref_matrix_delete(R);
    return 0;
}

/* rra.c:1807-1853 id_decomp_fixed_rank_or_prec on its full-rank branch
 * (k >= min(m, n), the case the two-sided ID uses): the full pivoted QR of M,
 * T = Rk1^-1 Rk2 from the first k rows. */
int oref_id_full(mat *M, int k, int *I, mat **T) {
    if (!g_backend) return -2;
    mat *R = NULL;
    pivoted_qr(M, NULL, &R, I);
    if (k <= 0 || k > R->nrows) { oref_matrix_delete(R); return -4; }
    *T = id_coefficients(R, k);
    oref_matrix_delete(R);
    return 0;
}

/* rra.c:2086-2112 id_two_sided_blockrand_decomp_fixed_rank_or_prec: the
 * column ID above, then a full ID of MI^T = M(:, Icol(1:frank))^T (frank x m),
 * whose pivots are the row indices Irow (m of them) and coefficients S. */
int oref_id_two_sided_blockrand(mat *M, int k, int p, double TOL, int kstep, int q, int s,
                                int *frank, int *Icol, int *Irow, mat **T, mat **S) {
    int rc = oref_id_blockrand(M, k, p, TOL, kstep, q, s, frank, Icol, T);
    if (rc) return rc;
    int m = M->nrows, f = *frank;
    mat *MIt = oref_matrix_new(f, m);
    for (int j = 0; j < f; ++j)
        for (int i = 0; i < m; ++i) MEL(MIt, j, i) = MEL(M, i, Icol[j]);
    rc = oref_id_full(MIt, f, Irow, S);
    oref_matrix_delete(MIt);
    if (rc) oref_matrix_delete(*T);
    return rc;
}

/* rra.c:2262-2332 cur_blockrand_decomp_fixed_rank_or_prec: the two-sided ID;
 * V = [I_k; T^T](Icol^-1, :) (n x k), R = M(Irow(1:k), :), C = M(:, Icol(1:k)),
 * U^T the solution of (R R^T) U^T = R V by dgesv (mvf.c:1525-1531). */
int oref_cur_blockrand(mat *M, int k, int p, double TOL, int kstep, int q, int s, int *frank,
                       mat **Cm, mat **Um, mat **Rm) {
    int m = M->nrows, n = M->ncols;
    int *Icol = (int *)malloc(sizeof(int) * (size_t)n), *Irow = (int *)malloc(sizeof(int) * (size_t)m);
    mat *T = NULL, *S = NULL;
    int rc = oref_id_two_sided_blockrand(M, k, p, TOL, kstep, q, s, frank, Icol, Irow, &T, &S);
    if (rc) { free(Icol); free(Irow); return rc; }
    k = *frank;
    int *Icolinv = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) Icolinv[Icol[i]] = i;           /* mvf.c:1195-1201 */
    mat *V = oref_matrix_new(n, k);                              /* V1(Icolinv, :) */
    for (int i = 0; i < n; ++i) {
        int r = Icolinv[i];
        for (int j = 0; j < k; ++j) MEL(V, i, j) = (r < k) ? (r == j ? 1.0 : 0.0) : MEL(T, j, r - k);
    }
    *Rm = oref_matrix_new(k, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < k; ++i) MEL(*Rm, i, j) = MEL(M, Irow[i], j);
    *Cm = oref_matrix_new(m, k);
    for (int j = 0; j < k; ++j)
        memcpy(&MEL(*Cm, 0, j), &MEL(M, 0, Icol[j]), sizeof(double) * (size_t)m);
    mat *RRt = oref_matrix_new(k, k), *RV = oref_matrix_new(k, k);
    matrix_matrix_transpose_mult(*Rm, *Rm, RRt);
    matrix_matrix_mult(*Rm, V, RV);
    int *ipiv = (int *)malloc(sizeof(int) * (size_t)(k + 1)), info = 0, kk = k;
    p_dgesv(&kk, &kk, RRt->d, &kk, ipiv, RV->d, &kk, &info);    /* RV <- Ut */
    *Um = oref_matrix_new(k, k);
    matrix_build_transpose(*Um, RV);
    free(ipiv); free(Icol); free(Irow); free(Icolinv);
    oref_matrix_delete(V); oref_matrix_delete(RRt); oref_matrix_delete(RV);
    oref_matrix_delete(T); oref_matrix_delete(S);
    return 0;
}

/* rra.c:924-975 exposed for the tests: R (min(m,n) x n for m <= n) and pivots. */
int oref_pivoted_qr(mat *M, mat **Q, mat **R, int *I) {
    if (!g_backend) return -2;
    pivoted_qr(M, Q, R, I);
    return 0;
}

/* mvf.c:1304-1318 form_svd_product_matrix and mvf.c:391-405 percent error. */
double oref_svd_residual(mat *M, mat *U, mat *S, mat *V) {
    int m = M->nrows, n = M->ncols, k = S->nrows;
    mat *SVt = oref_matrix_new(k, n), *P = oref_matrix_new(m, n);
    matrix_matrix_transpose_mult(S, V, SVt);
    matrix_matrix_mult(U, SVt, P);
    size_t N = (size_t)m * n;
    double acc = 0;
    for (size_t i = 0; i < N; ++i) { double d = M->d[i] - P->d[i]; acc += d * d; }
    oref_matrix_delete(SVt); oref_matrix_delete(P);
    return sqrt(acc);
}

double oref_wall_seconds(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

/* ------------------------------------------------------------------ */
/* FP32 restatement of the oneAPI variant (second oracle)              */
/* ------------------------------------------------------------------ */
/* ONEAPI = oneapi_code/; its mat holds int64 dimensions and float data
 * (ONEAPI/matrix_vector_functions_one_api.h:24-27) and every product, QR and
 * SVD is the single-precision LAPACK / BLAS routine (sgemm, sgeqrf+sorgqr,
 * sgesvd).  Restated here: randQB_pb2 (ONEAPI/rank_revealing_algorithms_one_api.c:605-813)
 * and low_rank_svd_rand_decomp_fromQB (same file :244-304).  Omega is the
 * seeded stream above (already float-valued) or an injected buffer, as for the
 * double oracle.  The Frobenius norm of the TOL test follows
 * ONEAPI/matrix_vector_functions_one_api.c:356-368: squares summed in float by
 * an OpenMP reduction, restated as `g_norm_chunks` contiguous chunks (the
 * static schedule of that many threads) each summed in float in index order,
 * the chunk sums added in float in chunk order. */
typedef struct {
    int64_t nrows, ncols;
    float *d;
} mat32;

static int g_norm_chunks = 16;
void oref32_set_norm_chunks(int c) { g_norm_chunks = c > 0 ? c : 1; }

mat32 *oref32_matrix_new(int64_t nrows, int64_t ncols) {
    mat32 *M = (mat32 *)malloc(sizeof(mat32));
    M->d = (float *)calloc((size_t)(nrows * ncols) + 1, sizeof(float));
    M->nrows = nrows;
    M->ncols = ncols;
    return M;
}
void oref32_matrix_delete(mat32 *M) {
    if (!M) return;
    free(M->d);
    free(M);
}
#define MEL32(M, i, j) ((M)->d[(size_t)(j) * (size_t)(M)->nrows + (size_t)(i)])

static void gemm32(char ta, char tb, int m, int n, int k, mat32 *A, mat32 *B, mat32 *C) {
    const float one = 1.0f, zero = 0.0f;
    int lda = (int)OMAX(1, A->nrows), ldb = (int)OMAX(1, B->nrows), ldc = (int)OMAX(1, C->nrows);
    if (m == 0 || n == 0) return;
    p_sgemm(&ta, &tb, &m, &n, &k, &one, A->d, &lda, B->d, &ldb, &zero, C->d, &ldc, 1, 1);
}
static void mm32(mat32 *A, mat32 *B, mat32 *C) {  /* ONEAPI mvf matrix_matrix_mult */
    gemm32('N', 'N', (int)A->nrows, (int)B->ncols, (int)A->ncols, A, B, C);
}
static void mtm32(mat32 *A, mat32 *B, mat32 *C) { /* matrix_transpose_matrix_mult */
    gemm32('T', 'N', (int)A->ncols, (int)B->ncols, (int)A->nrows, A, B, C);
}
static void mmt32(mat32 *A, mat32 *B, mat32 *C) { /* matrix_matrix_transpose_mult */
    gemm32('N', 'T', (int)A->nrows, (int)B->nrows, (int)A->ncols, A, B, C);
}
static void copy32(mat32 *D, mat32 *S) {
    memcpy(D->d, S->d, (size_t)(S->nrows * S->ncols) * sizeof(float));
}
static void sub32(mat32 *A, mat32 *B) {
    size_t N = (size_t)(A->nrows * A->ncols);
    for (size_t i = 0; i < N; ++i) A->d[i] -= B->d[i];
}
float oref32_frobenius(const float *d, int64_t N) {
    const int64_t T = g_norm_chunks;
    float total = 0.0f;
    for (int64_t t = 0; t < T; ++t) {
        /* static schedule: chunk t covers [t*N/T, (t+1)*N/T) */
        const int64_t lo = t * N / T, hi = (t + 1) * N / T;
        float part = 0.0f;
        for (int64_t i = lo; i < hi; ++i) part += d[i] * d[i];
        total += part;
    }
    return sqrtf(total);
}
static void build_transpose32(mat32 *Mt, mat32 *M) {
    for (int64_t i = 0; i < M->nrows; ++i)
        for (int64_t j = 0; j < M->ncols; ++j) MEL32(Mt, j, i) = MEL32(M, i, j);
}
static void qr_getQ32(mat32 *M, mat32 *Q) {       /* ONEAPI mvf QR_factorization_getQ */
    int m = (int)M->nrows, n = (int)M->ncols, k = (int)OMIN(m, n), lwork = -1, info = 0;
    float wq;
    copy32(Q, M);
    float *tau = (float *)calloc((size_t)k + 1, sizeof(float));
    p_sgeqrf(&m, &n, Q->d, &m, tau, &wq, &lwork, &info);
    lwork = (int)wq + 1;
    float *work = (float *)malloc(sizeof(float) * (size_t)lwork);
    p_sgeqrf(&m, &n, Q->d, &m, tau, work, &lwork, &info);
    free(work);
    lwork = -1;
    p_sorgqr(&m, &n, &n, Q->d, &m, tau, &wq, &lwork, &info);
    lwork = (int)wq + 1;
    work = (float *)malloc(sizeof(float) * (size_t)lwork);
    p_sorgqr(&m, &n, &n, Q->d, &m, tau, work, &lwork, &info);
    free(work);
    free(tau);
}
static void svd32(mat32 *M, mat32 *U, float *svals, mat32 *Vt) { /* sgesvd('S','S') */
    int m = (int)M->nrows, n = (int)M->ncols, k = (int)OMIN(m, n), lwork = -1, info = 0;
    int ldu = OMAX(1, m), ldvt = OMAX(1, k);
    float wq;
    p_sgesvd("S", "S", &m, &n, M->d, &m, svals, U->d, &ldu, Vt->d, &ldvt, &wq, &lwork, &info,
             1, 1);
    lwork = (int)wq + 1;
    float *work = (float *)malloc(sizeof(float) * (size_t)lwork);
    p_sgesvd("S", "S", &m, &n, M->d, &m, svals, U->d, &ldu, Vt->d, &ldvt, work, &lwork, &info,
             1, 1);
    free(work);
}
static void resize_cols32(mat32 **M, int64_t k) {   /* first k columns */
    mat32 *R = oref32_matrix_new((*M)->nrows, k);
    memcpy(R->d, (*M)->d, (size_t)((*M)->nrows * k) * sizeof(float));
    oref32_matrix_delete(*M);
    *M = R;
}
static void resize_rows32(mat32 **M, int64_t k) {   /* first k rows */
    mat32 *R = oref32_matrix_new(k, (*M)->ncols);
    for (int64_t j = 0; j < R->ncols; ++j)
        for (int64_t i = 0; i < k; ++i) MEL32(R, i, j) = MEL32(*M, i, j);
    oref32_matrix_delete(*M);
    *M = R;
}

/* ONEAPI mvf compute_evals_and_evecs_of_symm_matrix: ssyev('V','U') in place */
static void syev32(mat32 *S, float *evals) {
    int n = (int)S->nrows, lda = (int)OMAX(1, S->ncols), lwork = -1, info = 0;
    float wq;
    p_ssyev("V", "U", &n, S->d, &lda, evals, &wq, &lwork, &info, 1, 1);
    lwork = (int)wq + 1;
    float *work = (float *)malloc(sizeof(float) * (size_t)lwork);
    p_ssyev("V", "U", &n, S->d, &lda, evals, work, &lwork, &info, 1, 1);
    free(work);
}
/* keep columns [off, off + k) (and rows, for a square S) */
static void keep_cols32(mat32 **M, int64_t off, int64_t k) {
    mat32 *R = oref32_matrix_new((*M)->nrows, k);
    memcpy(R->d, (*M)->d + (size_t)(off * (*M)->nrows), (size_t)((*M)->nrows * k) * sizeof(float));
    oref32_matrix_delete(*M);
    *M = R;
}
static void keep_square32(mat32 **S, int64_t off, int64_t k) {
    mat32 *R = oref32_matrix_new(k, k);
    for (int64_t j = 0; j < k; ++j)
        for (int64_t i = 0; i < k; ++i) MEL32(R, i, j) = MEL32(*S, off + i, off + j);
    oref32_matrix_delete(*S);
    *S = R;
}

/* ONEAPI/rank_revealing_algorithms_one_api.c:72-240 low_rank_svd_rand_decomp_fixed_rank
 * in single precision: Y = M RN (:95), q - 1 power iterations with QR on
 * alternate half steps (:110-135), Q = qr(Y) (:143-144); vnum 1 (or > 2):
 * B^T = M^T Q, compact QR, sgesvd of R-hat, U = Q V-hat, V = Q-hat U-hat, first
 * k kept (:146-188); vnum 2: B = Q^T M, ssyev of the upper triangle of BB^T,
 * sigma = sqrt(lambda) (ascending, no clamp), U = Q U-hat, V = B^T (U-hat
 * Sigma^-1), last k kept (:190-234). */
int oref32_fixed_rank(mat32 *M, int64_t k, int64_t p, int64_t vnum, int64_t q, int64_t s,
                      mat32 **U, mat32 **S, mat32 **V) {
    if (!p_sgemm) return -2;
    int64_t m = M->nrows, n = M->ncols, l = k + p;
    if (k <= 0 || l > OMIN(m, n) || s <= 0) return -1;
    mat32 *RN = oref32_matrix_new(n, l);
    {
        int64_t N = n * l;
        if (g_inj) {
            if (g_inj_count < N) { oref32_matrix_delete(RN); return -3; }
            for (int64_t i = 0; i < N; ++i) RN->d[i] = (float)g_inj[i];
        } else {
            double *tmp = (double *)malloc(sizeof(double) * (size_t)N);
            oref_gaussian(g_seed, 0, N, tmp);
            for (int64_t i = 0; i < N; ++i) RN->d[i] = (float)tmp[i];
            free(tmp);
        }
    }
    mat32 *Y = oref32_matrix_new(m, l);
    mm32(M, RN, Y);
    oref32_matrix_delete(RN);
    mat32 *Z = oref32_matrix_new(n, l), *Yo = oref32_matrix_new(m, l), *Zo = oref32_matrix_new(n, l);
    for (int64_t j = 1; j < q; ++j) {
        if ((2 * j - 2) % s == 0) { qr_getQ32(Y, Yo); mtm32(M, Yo, Z); }
        else mtm32(M, Y, Z);
        if ((2 * j - 1) % s == 0) { qr_getQ32(Z, Zo); mm32(M, Zo, Y); }
        else mm32(M, Z, Y);
    }
    oref32_matrix_delete(Z); oref32_matrix_delete(Yo); oref32_matrix_delete(Zo);
    mat32 *Q = oref32_matrix_new(m, l);
    qr_getQ32(Y, Q);
    *U = oref32_matrix_new(m, l);
    *S = oref32_matrix_new(l, l);
    *V = oref32_matrix_new(n, l);
    if (vnum == 1 || vnum > 2) {
        mat32 *Bt = oref32_matrix_new(n, l);
        mtm32(M, Q, Bt);
        /* compact QR: R = triu(first l rows of the factored Bt), Q-hat = sorgqr */
        mat32 *Qh = oref32_matrix_new(n, l), *Rh = oref32_matrix_new(l, l);
        {
            int mm_ = (int)n, nn_ = (int)l, lwork = -1, info = 0;
            float wq;
            copy32(Qh, Bt);
            float *tau = (float *)calloc((size_t)l + 1, sizeof(float));
            p_sgeqrf(&mm_, &nn_, Qh->d, &mm_, tau, &wq, &lwork, &info);
            lwork = (int)wq + 1;
            float *work = (float *)malloc(sizeof(float) * (size_t)lwork);
            p_sgeqrf(&mm_, &nn_, Qh->d, &mm_, tau, work, &lwork, &info);
            free(work);
            for (int64_t i = 0; i < l; ++i)
                for (int64_t jj = i; jj < l; ++jj) MEL32(Rh, i, jj) = MEL32(Qh, i, jj);
            lwork = -1;
            p_sorgqr(&mm_, &nn_, &nn_, Qh->d, &mm_, tau, &wq, &lwork, &info);
            lwork = (int)wq + 1;
            work = (float *)malloc(sizeof(float) * (size_t)lwork);
            p_sorgqr(&mm_, &nn_, &nn_, Qh->d, &mm_, tau, work, &lwork, &info);
            free(work);
            free(tau);
        }
        mat32 *Uh = oref32_matrix_new(l, l), *Vht = oref32_matrix_new(l, l);
        float *sv = (float *)calloc((size_t)l + 1, sizeof(float));
        svd32(Rh, Uh, sv, Vht);
        for (int64_t i = 0; i < l; ++i) MEL32(*S, i, i) = sv[i];
        mmt32(Q, Vht, *U);
        mm32(Qh, Uh, *V);
        free(sv);
        oref32_matrix_delete(Bt); oref32_matrix_delete(Qh); oref32_matrix_delete(Rh);
        oref32_matrix_delete(Uh); oref32_matrix_delete(Vht);
        keep_cols32(U, 0, k); keep_cols32(V, 0, k); keep_square32(S, 0, k);
    } else {
        mat32 *B = oref32_matrix_new(l, n), *BBt = oref32_matrix_new(l, l);
        mtm32(Q, M, B);
        mmt32(B, B, BBt);
        mat32 *Uh = oref32_matrix_new(l, l);
        for (int64_t i = 0; i < l; ++i)      /* matrix_copy_symmetric: upper triangle */
            for (int64_t jj = i; jj < l; ++jj) MEL32(Uh, i, jj) = MEL32(BBt, i, jj);
        float *ev = (float *)calloc((size_t)l + 1, sizeof(float));
        syev32(Uh, ev);
        for (int64_t i = 0; i < l; ++i) MEL32(*S, i, i) = sqrtf(ev[i]);
        mm32(Q, Uh, *U);
        mat32 *Sinv = oref32_matrix_new(l, l), *UhSinv = oref32_matrix_new(l, l);
        for (int64_t i = 0; i < l; ++i) MEL32(Sinv, i, i) = 1.0f / MEL32(*S, i, i);
        mm32(Uh, Sinv, UhSinv);
        mtm32(B, UhSinv, *V);
        free(ev);
        oref32_matrix_delete(B); oref32_matrix_delete(BBt); oref32_matrix_delete(Uh);
        oref32_matrix_delete(Sinv); oref32_matrix_delete(UhSinv);
        keep_cols32(U, l - k, k); keep_cols32(V, l - k, k); keep_square32(S, l - k, k);
    }
    oref32_matrix_delete(Y);
    oref32_matrix_delete(Q);
    return 0;
}

/* ONEAPI/rank_revealing_algorithms_one_api.c:605-813 randQB_pb2: the blocked QB
 * of randQB_pb_new in single precision (kstep clamp :618-620, TOL mode when
 * nstep <= 0 :623-628, q power iterations with QR on alternate half steps
 * :678-709, re-orthogonalisation against all earlier blocks on even steps
 * :725-749, Bp = Qp^T A :757, A -= Qp Bp, frank = (step + 1) kstep :786 and
 * the strict norm < TOL stop :788-794, first-frank resize in TOL mode
 * :797-799). */
int oref32_randQB_pb2(mat32 *M, int64_t kstep, int64_t nstep, float TOL, int64_t q, int64_t s,
                      int64_t *frank, mat32 **Qo, mat32 **Bo, float *final_norm) {
    if (!p_sgemm) return -2;
    int64_t m = M->nrows, n = M->ncols, tolMode = 0;
    if (kstep > (int64_t)(OMIN(m, n) / 2)) kstep = (int64_t)OMIN(m, n) / 10;
    if (nstep <= 0) { tolMode = 1; if (kstep <= 0) return -1; nstep = (int64_t)(OMIN(m, n) / kstep); }
    if (kstep <= 0 || nstep <= 0 || s <= 0) return -1;
    int64_t l = kstep * nstep;
    mat32 *RN = oref32_matrix_new(n, l);
    {   /* initialize_random_matrix: float Gaussians, column-major */
        int64_t N = n * l;
        if (g_inj) {
            if (g_inj_count < N) { oref32_matrix_delete(RN); return -3; }
            for (int64_t i = 0; i < N; ++i) RN->d[i] = (float)g_inj[i];
        } else {
            double *tmp = (double *)malloc(sizeof(double) * (size_t)N);
            oref_gaussian(g_seed, 0, N, tmp);
            for (int64_t i = 0; i < N; ++i) RN->d[i] = (float)tmp[i];
            free(tmp);
        }
    }
    mat32 *A = oref32_matrix_new(m, n);
    *Qo = oref32_matrix_new(m, l);
    *Bo = oref32_matrix_new(l, n);
    mat32 *RNp = oref32_matrix_new(n, kstep), *Yp = oref32_matrix_new(m, kstep);
    mat32 *Qp = oref32_matrix_new(m, kstep), *Bp = oref32_matrix_new(kstep, n);
    mat32 *QpBp = oref32_matrix_new(m, n), *AtQp = oref32_matrix_new(n, kstep);
    mat32 *AtQp2 = oref32_matrix_new(n, kstep), *QptA = oref32_matrix_new(kstep, n);
    copy32(A, M);
    int64_t fr = 0;
    float normval = 0.0f;
    for (int64_t step = 0; step < nstep; ++step) {
        const int64_t c0 = step * kstep;
        memcpy(RNp->d, RN->d + (size_t)(c0 * n), (size_t)(n * kstep) * sizeof(float));
        mm32(A, RNp, Yp);
        for (int64_t j = 1; j <= q; ++j) {
            if ((2 * j - 2) % s == 0) {
                qr_getQ32(Yp, Qp);
                mtm32(Qp, A, QptA);
                build_transpose32(AtQp, QptA);
            } else {
                mtm32(A, Yp, AtQp);
            }
            if ((2 * j - 1) % s == 0) {
                qr_getQ32(AtQp, AtQp2);
                mm32(A, AtQp2, Yp);
            } else {
                mm32(A, AtQp, Yp);
            }
        }
        qr_getQ32(Yp, Qp);
        if (step > 0 && step % 2 == 0) {
            const int64_t nj = step * kstep;
            mat32 *Qj = oref32_matrix_new(m, nj), *QjtQp = oref32_matrix_new(nj, kstep);
            mat32 *QjQjtQp = oref32_matrix_new(m, kstep);
            memcpy(Qj->d, (*Qo)->d, (size_t)(m * nj) * sizeof(float));
            mtm32(Qj, Qp, QjtQp);
            mm32(Qj, QjtQp, QjQjtQp);
            copy32(Yp, Qp);
            sub32(Yp, QjQjtQp);
            qr_getQ32(Yp, Qp);
            oref32_matrix_delete(Qj); oref32_matrix_delete(QjtQp); oref32_matrix_delete(QjQjtQp);
        }
        mtm32(Qp, A, Bp);
        mm32(Qp, Bp, QpBp);
        sub32(A, QpBp);
        memcpy((*Qo)->d + (size_t)(c0 * m), Qp->d, (size_t)(m * kstep) * sizeof(float));
        for (int64_t jj = 0; jj < n; ++jj)
            for (int64_t i = 0; i < kstep; ++i) MEL32(*Bo, c0 + i, jj) = MEL32(Bp, i, jj);
        fr = (step + 1) * kstep;
        if (tolMode) {
            normval = oref32_frobenius(A->d, m * n);
            if (normval < TOL) break;
        }
    }
    if (tolMode) { resize_cols32(Qo, fr); resize_rows32(Bo, fr); }
    if (frank) *frank = fr;
    if (final_norm) *final_norm = tolMode ? normval : oref32_frobenius(A->d, m * n);
    oref32_matrix_delete(A); oref32_matrix_delete(RN); oref32_matrix_delete(RNp);
    oref32_matrix_delete(Yp); oref32_matrix_delete(Qp); oref32_matrix_delete(Bp);
    oref32_matrix_delete(QpBp); oref32_matrix_delete(AtQp); oref32_matrix_delete(AtQp2);
    oref32_matrix_delete(QptA);
    return 0;
}

/* ONEAPI/rank_revealing_algorithms_one_api.c:244-304 low_rank_svd_rand_decomp_fromQB:
 * BB^T (knew x knew, knew = Q->ncols), its SVD by sgesvd (:267), sigma_i =
 * sqrt(St_ii) (descending, :275-279), U = Q Uhat (:284), V = B^T (Uhat
 * Sigma^-1) with Uhat Sigma^-1 formed as a product (:291-293).  No
 * truncation: U m x knew, S knew x knew, V n x knew. */
int oref32_svd_fromQB(mat32 *Q, mat32 *B, mat32 **U, mat32 **S, mat32 **V) {
    if (!p_sgemm) return -2;
    int64_t m = Q->nrows, n = B->ncols, kn = Q->ncols;
    if (B->nrows != kn) return -1;
    *U = oref32_matrix_new(m, kn);
    *S = oref32_matrix_new(kn, kn);
    *V = oref32_matrix_new(n, kn);
    mat32 *Uhat = oref32_matrix_new(kn, kn), *Vt = oref32_matrix_new(kn, kn);
    mat32 *BBt = oref32_matrix_new(kn, kn);
    mmt32(B, B, BBt);
    float *st = (float *)calloc((size_t)kn + 1, sizeof(float));
    svd32(BBt, Uhat, st, Vt);
    for (int64_t i = 0; i < kn; ++i) MEL32(*S, i, i) = sqrtf(st[i]);
    mm32(Q, Uhat, *U);
    mat32 *Sinv = oref32_matrix_new(kn, kn), *UhatSinv = oref32_matrix_new(kn, kn);
    for (int64_t i = 0; i < kn; ++i) MEL32(Sinv, i, i) = 1.0f / MEL32(*S, i, i);
    mm32(Uhat, Sinv, UhatSinv);
    mtm32(B, UhatSinv, *V);
    free(st);
    oref32_matrix_delete(BBt); oref32_matrix_delete(Vt); oref32_matrix_delete(Uhat);
    oref32_matrix_delete(Sinv); oref32_matrix_delete(UhatSinv);
    return 0;
}
