// pipeline.h -- device-resident L2 algorithms shared by the C API layers.
#pragma once

#include "rsvd_internal.h"

namespace rsvd {

// Source of the Gaussian test matrix Omega (n x cols): either an injected
// device matrix (column j at dev + j*ld) or the seeded Philox stream of an
// n-row matrix.
struct OmegaSrc {
  uint64_t seed = 0;
  const double* dev = nullptr;
  int64_t ld = 0;
};

struct SvdInfo {
  int64_t l = 0;
  int qr_fallbacks = 0;
  int jacobi_sweeps = 0;
};

struct QbInfo {
  int steps = 0;
  int qr_fallbacks = 0;
  int kstep = 0;
  int nstep = 0;
  int tol_mode = 0;
  double last_norm = 0;
};

struct BlockrandOut {
  DBuf<double> U, S, V;  // m x frank, frank, n x frank
  int frank = 0;
  int l = 0;
  QbInfo qb;
  SvdInfo svd;
};

int omega_block(Context* c, const OmegaSrc& om, int64_t n, int64_t col0, int64_t cols,
                DBuf<double>& buf, const double** out, int64_t* ld);

// U (m x k), S (k), V (n x k) from an orthonormal Q (m x l), following the
// vnum branches of rra.c:133-225 (ordering and truncation included).
int svd_from_Q_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda,
                   const double* Q, int64_t ldq, int64_t l, int64_t k, int vnum, double* U,
                   int64_t ldu, double* S, double* V, int64_t ldv, SvdInfo* info);

int fixed_rank_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                   int vnum, int q, int s, const OmegaSrc& om, double* U, int64_t ldu, double* S,
                   double* V, int64_t ldv, SvdInfo* info);

// kstep clamp and TOL-mode nstep of randQB_pb_new (is_new) or plain checks.
int qb_adjust(int64_t m, int64_t n, int* kstep, int* nstep, int is_new, int* tolMode);

// Q (m x l, ld ldq) and B (l x n, ld ldb) must hold l = kstep*nstep columns /
// rows after qb_adjust.  *frank = columns produced, *l_alloc = kstep*nstep.
int qb_blocked_f64(Context* c, const double* M, int64_t m, int64_t n, int64_t ldm, int kstep,
                   int nstep, double TOL, int q, int s, int is_new, const OmegaSrc& om, double* Q,
                   int64_t ldq, double* B, int64_t ldb, int* frank, int* l_alloc, QbInfo* info);

// p / nstep / mode arithmetic of rra.c:251-263.
int blockrand_plan(int64_t m, int64_t n, int k, int p, int kstep, int* p_eff, int* nstep,
                   int* rankMode);

int blockrand_f64(Context* c, const double* M, int64_t m, int64_t n, int64_t ldm, int k, int p,
                  double TOL, int vnum, int kstep, int q, int s, const OmegaSrc& om,
                  BlockrandOut* out);

// A = U diag(sigma) V^T + noise_std * N(0,1), U, V Haar (m x r, n x r).  Writes
// the double matrix to A64 and / or its float rounding to A32 (when both are
// given, A64 holds exactly the float values).
int lowrank_test_matrix(Context* c, int64_t m, int64_t n, int64_t r, const double* sigma_host,
                        double noise_std, uint64_t seed, double* A64, int64_t lda64, float* A32,
                        int64_t lda32);

}  // namespace rsvd
