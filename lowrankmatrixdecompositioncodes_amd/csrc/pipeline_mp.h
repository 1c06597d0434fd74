// pipeline_mp.h -- mixed-precision fixed-rank rSVD on a float32 matrix.
#pragma once

#include "pipeline.h"

namespace rsvd {

// Precision of the sketch and power-iteration passes over A (the projection
// B^T = A^T Q is split-bf16, except for MP_FP16): bf16, split-bf16, or fp8
// e4m3 with per-128-column scales, for every pass; or bf16 (MP_BF16_SPLIT) /
// fp8 (MP_FP8_SPLIT) for every pass but the last, which is split-bf16; or
// MP_FP16: every pass over A, the projection included, with A in fp16 (per
// 128-column scales) against the skinny operand in two fp16 planes.
enum MpPrecision {
  MP_BF16 = 1, MP_BF16X3 = 2, MP_FP8 = 3, MP_BF16_SPLIT = 4, MP_FP8_SPLIT = 5, MP_FP16 = 6
};

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

// U (m x k), S (k), V (n x k) from an orthonormal fp32 Q (m x l) of the fp32
// A, the vnum branches of rra.c:133-225: B^T = A^T Q in split-bf16 (or the
// caller's Bt_in, n x l, ld ldbt), the small problem in FP64, U and V
// products in split-bf16 (exact = 0) or with fp32-exact products (exact = 1).
// fp16_ascale (optional): the fp16 chunk scales of A, to form B^T on the fp16
// path instead.  *sw = Jacobi sweeps, *fb += QR fallbacks.
int svd_from_Q_mp(Context* c, const float* A, int64_t m, int64_t n, int64_t lda, const float* Q,
                  int64_t ldq, int64_t l, int64_t k, int vnum, const float* Bt_in, int64_t ldbt,
                  int exact, float* U, int64_t ldu, double* S, float* V, int64_t ldv, int* sw,
                  int* fb, const float* fp16_ascale = nullptr);

// randQB_pb_new (is_new = 1) / randQB_pb on a float32 A: a float32 working
// copy R of A is deflated block by block (R -= Qp Bp with fp32-exact products,
// Bp = Qp^T R in split-bf16 and its error cleared on the next block, FP64 sum
// of squares of R for the TOL test); the sketch and power passes over R run
// at precision prec; each Qp is orthonormalised in FP64 and, from the second
// block on, re-orthogonalised against all earlier blocks.  Q: m x l fp32
// (l = kstep*nstep after qb_adjust); Bt (optional): n x l fp32, B^T by
// column blocks.  dRN (optional): n x l fp32 Gaussian columns, else the
// seeded stream rounded to float.
int qb_blocked_mp(Context* c, const float* M, int64_t m, int64_t n, int64_t ldm, int kstep,
                  int nstep, double TOL, int q, int s, int is_new, uint64_t seed,
                  const float* dRN, int prec, float* Q, int64_t ldq, float* Bt, int64_t ldbt,
                  int* frank, int* l_alloc, QbInfo* info);

struct BlockrandMpOut {
  DBuf<float> U, V;   // m x frank, n x frank
  DBuf<double> S;     // frank
  int frank = 0;
  int l = 0;
  QbInfo qb;
  SvdInfo svd;
};

// low_rank_svd_blockrand_decomp_fixed_rank_or_prec (rra.c:239-381) on a
// float32 A: qb_blocked_mp, then svd_from_Q_mp on the QB's Q and B with
// fp32-exact U and V products.
int blockrand_mp(Context* c, const float* M, int64_t m, int64_t n, int64_t ldm, int k, int p,
                 double TOL, int vnum, int kstep, int q, int s, uint64_t seed, const float* dRN,
                 int prec, BlockrandMpOut* out);

// ||A - U diag(S) V^T||_F with A double (m x n) and U, V double or float
// (uv_f32), accumulated in FP64 over column chunks of A.
int lowrank_residual(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                     const void* U, int64_t ldu, const double* S, const void* V, int64_t ldv,
                     int uv_f32, double* out);

}  // namespace rsvd
