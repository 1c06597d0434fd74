// pipeline_mp.h -- mixed-precision fixed-rank rSVD on a float32 matrix.
#pragma once

#include "rsvd_internal.h"

namespace rsvd {

// Precision of the sketch and power-iteration passes over A (the projection
// B^T = A^T Q is always split-bf16): bf16, split-bf16, or fp8 e4m3 with
// per-128-column scales.
enum MpPrecision { MP_BF16 = 1, MP_BF16X3 = 2, MP_FP8 = 3 };

struct MpInfo {
  int qr_fallbacks = 0;
  int jacobi_sweeps = 0;
};

// U (m x k, fp32), S (k, double, reference order), V (n x k, fp32) of the
// fixed-rank rSVD of the fp32 matrix A.  dOmega (n x (k+p) fp32, ld n) or the
// seeded stream rounded to fp32.
int fixed_rank_mp(Context* c, const float* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                  int vnum, int q, int s, uint64_t seed, const float* dOmega, int prec, float* U,
                  int64_t ldu, double* S, float* V, int64_t ldv, MpInfo* info);

// ||A - U diag(S) V^T||_F with A double (m x n) and U, V double or float
// (uv_f32), accumulated in FP64 over column chunks of A.
int lowrank_residual(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                     const void* U, int64_t ldu, const double* S, const void* V, int64_t ldv,
                     int uv_f32, double* out);

}  // namespace rsvd
