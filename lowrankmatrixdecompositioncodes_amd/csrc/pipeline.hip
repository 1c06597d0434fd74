// pipeline.hip -- the RSVDPACK L2 algorithms run resident in HBM (FP64).
//
//   fixed_rank_f64 : low_rank_svd_rand_decomp_fixed_rank  (MKL/rra.c:73-234)
//   qb_blocked_f64 : randQB_pb (rra.c:1425-1569) / randQB_pb_new (rra.c:1576-1801)
//   blockrand_f64  : low_rank_svd_blockrand_decomp_fixed_rank_or_prec (rra.c:239-381)
//
// The control flow (loop counts, re-orthogonalisation schedule, vnum
// branches, output ordering and truncation, kstep / nstep / frank integer
// arithmetic) follows the reference; the arithmetic runs on the GPU kernels of
// gemm.hip, qr.hip and jacobi.hip.  Gathers / scatters of the reference (RN
// column blocks, Q columns, B rows) are strided views here, not copies.
#include <cmath>
#include <cstring>

#include "pipeline.h"

namespace rsvd {

int omega_block(Context* c, const OmegaSrc& om, int64_t n, int64_t col0, int64_t cols,
                DBuf<double>& buf, const double** out, int64_t* ld) {
  if (om.dev) {
    *out = om.dev + col0 * om.ld;
    *ld = om.ld;
    return 0;
  }
  RSVD_TRY(buf.alloc((size_t)(n * cols)));
  RSVD_TRY(gaussian_f64(c, om.seed, n, cols, col0, n, buf.p, n));
  *out = buf.p;
  *ld = n;
  return 0;
}

// Symmetrise from the upper triangle (dsyev('U') reads only the upper part).
__global__ void symmetrize_upper_kernel(int64_t n, double* __restrict__ G) {
  int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % n, j = idx / n;
    if (i > j) G[idx] = G[j + i * n];
  }
}

int symmetrize_upper_f64(Context* c, int64_t n, double* G) {
  if (n == 0) return 0;
  int64_t b = (n * n + 255) / 256;
  hipLaunchKernelGGL(symmetrize_upper_kernel, dim3((unsigned)(b > 4096 ? 4096 : b)), dim3(256), 0,
                     c->stream, n, G);
  RSVD_LAUNCH_CHECK();
  return 0;
}

__global__ void sqrt_kernel(int64_t n, const double* __restrict__ in, double* __restrict__ out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    // a PSD eigenvalue may round to -eps; a NaN stays NaN (fmax would hide it)
    out[j] = in[j] < 0.0 ? 0.0 : sqrt(in[j]);
}

static int grid1(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

int sqrt_f64(Context* c, int64_t n, const double* in, double* out) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(sqrt_kernel, dim3(grid1(n)), dim3(256), 0, c->stream, n, in, out);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int svd_from_Q_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda,
                   const double* Q, int64_t ldq, int64_t l, int64_t k, int vnum, double* U,
                   int64_t ldu, double* S, double* V, int64_t ldv, SvdInfo* info) {
  if (k <= 0 || k > l) return set_error("svd_from_Q: k=%lld outside [1, l=%lld]", (long long)k,
                                        (long long)l);
  if (vnum == 1 || vnum > 2) {
    // Bt = A^T Q (n x l); [Qhat, Rhat] = qr(Bt); Rhat = Uhat Sigma Vhat^T;
    // U = Q Vhat, V = Qhat Uhat; descending, first k kept (rra.c:133-174).
    if (n < l) return set_error("vnum=1 needs n >= l (n=%lld, l=%lld)", (long long)n,
                                (long long)l);
    DBuf<double> Bt, R, Vh, sig;
    RSVD_TRY(Bt.alloc((size_t)(n * l)));
    RSVD_TRY(R.alloc((size_t)(l * l)));
    RSVD_TRY(Vh.alloc((size_t)(l * l)));
    RSVD_TRY(sig.alloc((size_t)l));
    RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'T', 'N', n, l, m, 1.0, A, lda, Q, ldq, 0.0, Bt.p, n));
    int fb = 0;
    RSVD_TRY(orth_f64(c, n, l, Bt.p, n, R.p, &fb));
    int sw = 0;
    RSVD_TRY(jacobi_f64(c, l, l, R.p, l, Vh.p, l, 60, &sw));
    RSVD_TRY(colnorms_f64(c, l, l, R.p, l, sig.p, 1));  // R -> Uhat, sig = Sigma
    RSVD_TRY(sort_pairs_f64(c, l, sig.p, 1, R.p, l, l, Vh.p, l, l));
    RSVD_TRY(gemm_f64(c, 'N', 'N', m, k, l, 1.0, Q, ldq, Vh.p, l, 0.0, U, ldu));
    RSVD_TRY(gemm_f64(c, 'N', 'N', n, k, l, 1.0, Bt.p, n, R.p, l, 0.0, V, ldv));
    RSVD_HIP(hipMemcpyAsync(S, sig.p, sizeof(double) * k, hipMemcpyDeviceToDevice, c->stream));
    if (info) { info->jacobi_sweeps = sw; info->qr_fallbacks += fb; }
    return 0;
  }
  // vnum == 2: B = Q^T A (l x n); BB^T = Uhat Lambda Uhat^T; sigma = sqrt(lambda);
  // U = Q Uhat; V = B^T Uhat Sigma^-1; ascending, last k kept (rra.c:175-225).
  DBuf<double> B, G, Uh, lam, sig;
  RSVD_TRY(B.alloc((size_t)(l * n)));
  RSVD_TRY(G.alloc((size_t)(l * l)));
  RSVD_TRY(Uh.alloc((size_t)(l * l)));
  RSVD_TRY(lam.alloc((size_t)l));
  RSVD_TRY(sig.alloc((size_t)l));
  RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'T', 'N', l, n, m, 1.0, Q, ldq, A, lda, 0.0, B.p, l));
  RSVD_TRY(gemm_f64(c, 'N', 'T', l, l, n, 1.0, B.p, l, B.p, l, 0.0, G.p, l));
  hipLaunchKernelGGL(symmetrize_upper_kernel, dim3(grid1(l * l)), dim3(256), 0, c->stream, l, G.p);
  RSVD_LAUNCH_CHECK();
  int sw = 0;
  RSVD_TRY(symeig_jacobi_f64(c, l, G.p, l, Uh.p, l, lam.p, 60, &sw));
  RSVD_TRY(sort_pairs_f64(c, l, lam.p, 0, Uh.p, l, l, nullptr, 0, 0));
  hipLaunchKernelGGL(sqrt_kernel, dim3(grid1(l)), dim3(256), 0, c->stream, l,
                     (const double*)lam.p, sig.p);
  RSVD_LAUNCH_CHECK();
  const int64_t off = l - k;
  RSVD_TRY(gemm_f64(c, 'N', 'N', m, k, l, 1.0, Q, ldq, Uh.p + off * l, l, 0.0, U, ldu));
  RSVD_TRY(scale_cols_f64(c, l, l, Uh.p, l, sig.p, 1));  // Uhat Sigma^-1
  RSVD_TRY(gemm_f64(c, 'T', 'N', n, k, l, 1.0, B.p, l, Uh.p + off * l, l, 0.0, V, ldv));
  RSVD_HIP(hipMemcpyAsync(S, sig.p + off, sizeof(double) * k, hipMemcpyDeviceToDevice,
                          c->stream));
  if (info) info->jacobi_sweeps = sw;
  return 0;
}

int fixed_rank_f64(Context* c, const double* A, int64_t m, int64_t n, int64_t lda, int k, int p,
                   int vnum, int q, int s, const OmegaSrc& om, double* U, int64_t ldu, double* S,
                   double* V, int64_t ldv, SvdInfo* info) {
  const int64_t l = (int64_t)k + p;
  if (k <= 0 || p < 0) return set_error("fixed_rank: need k > 0 and p >= 0 (k=%d, p=%d)", k, p);
  if (l > (m < n ? m : n))
    return set_error("fixed_rank: k+p=%lld exceeds min(m,n)=%lld", (long long)l,
                     (long long)(m < n ? m : n));
  if (s <= 0) return set_error("fixed_rank: s must be positive (s=%d)", s);
  DBuf<double> Ob, Y, Z;
  const double* Om;
  int64_t ldo;
  RSVD_TRY(omega_block(c, om, n, 0, l, Ob, &Om, &ldo));               // rra.c:90-91
  RSVD_TRY(Y.alloc((size_t)(m * l)));
  RSVD_TRY(Z.alloc((size_t)(n * l)));
  RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'N', 'N', m, l, n, 1.0, A, lda, Om, ldo, 0.0, Y.p, m));  // rra.c:95
  int fb = 0, f1 = 0;
  for (int j = 1; j < q; ++j) {                                        // rra.c:101-126
    if ((2 * j - 2) % s == 0) { RSVD_TRY(orth_f64(c, m, l, Y.p, m, nullptr, &f1)); fb += f1; }
    RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'T', 'N', n, l, m, 1.0, A, lda, Y.p, m, 0.0, Z.p, n));
    if ((2 * j - 1) % s == 0) { RSVD_TRY(orth_f64(c, n, l, Z.p, n, nullptr, &f1)); fb += f1; }
    RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'N', 'N', m, l, n, 1.0, A, lda, Z.p, n, 0.0, Y.p, m));
  }
  RSVD_TRY(orth_f64(c, m, l, Y.p, m, nullptr, &f1));                   // rra.c:129-130
  fb += f1;
  if (info) { info->qr_fallbacks = fb; info->l = l; }
  return svd_from_Q_f64(c, A, m, n, lda, Y.p, m, l, k, vnum, U, ldu, S, V, ldv, info);
}

int qb_adjust(int64_t m, int64_t n, int* kstep, int* nstep, int is_new, int* tolMode) {
  *tolMode = 0;
  const int64_t mn = m < n ? m : n;
  if (is_new) {
    if (*kstep > (int)(mn / 2)) *kstep = (int)(mn / 10);              // rra.c:1589-1592
    if (*nstep <= 0) {                                                  // rra.c:1595-1602
      *tolMode = 1;
      if (*kstep <= 0) return set_error("randQB: kstep=%d after clamping", *kstep);
      *nstep = (int)(mn / *kstep);
    }
  }
  if (*kstep <= 0 || *nstep <= 0)
    return set_error("randQB: need kstep > 0 and nstep > 0 (kstep=%d, nstep=%d)", *kstep, *nstep);
  return 0;
}

int qb_blocked_f64(Context* c, const double* M, int64_t m, int64_t n, int64_t ldm, int kstep,
                   int nstep, double TOL, int q, int s, int is_new, const OmegaSrc& om, double* Q,
                   int64_t ldq, double* B, int64_t ldb, int* frank, int* l_alloc, QbInfo* info) {
  int tolMode = 0;
  RSVD_TRY(qb_adjust(m, n, &kstep, &nstep, is_new, &tolMode));
  if (s <= 0) return set_error("randQB: s must be positive (s=%d)", s);
  const int64_t l = (int64_t)kstep * nstep;
  if (l_alloc) *l_alloc = (int)l;
  if (kstep > m) return set_error("randQB: kstep=%d exceeds m=%lld", kstep, (long long)m);
  DBuf<double> A, Yp, AtQp, QtQ, sums, Ob;
  RSVD_TRY(A.alloc((size_t)(m * n)));
  RSVD_TRY(Yp.alloc((size_t)(m * kstep)));
  RSVD_TRY(AtQp.alloc((size_t)(n * kstep)));
  RSVD_TRY(sums.alloc(1));
  RSVD_TRY(copy2d_f64(c, m, n, M, ldm, A.p, m));                       // rra.c:1630
  int fr = 0, fb = 0, f1 = 0, steps = 0;
  double normval = 0;
  for (int step = 0; step < nstep; ++step) {
    const int64_t c0 = (int64_t)step * kstep;
    const double* RNp;
    int64_t ldr;
    RSVD_TRY(omega_block(c, om, n, c0, kstep, Ob, &RNp, &ldr));       // rra.c:1639-1642
    RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'N', 'N', m, kstep, n, 1.0, A.p, m, RNp, ldr, 0.0, Yp.p, m));
    for (int j = 1; j <= q; ++j) {                                      // rra.c:1652-1686
      if ((2 * j - 2) % s == 0) { RSVD_TRY(orth_f64(c, m, kstep, Yp.p, m, nullptr, &f1)); fb += f1; }
      // rra.c:1657-1658 forms Qp^T A and transposes it; A^T Qp directly here
      RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'T', 'N', n, kstep, m, 1.0, A.p, m, Yp.p, m, 0.0, AtQp.p, n));
      if ((2 * j - 1) % s == 0) { RSVD_TRY(orth_f64(c, n, kstep, AtQp.p, n, nullptr, &f1)); fb += f1; }
      RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'N', 'N', m, kstep, n, 1.0, A.p, m, AtQp.p, n, 0.0, Yp.p, m));
    }
    RSVD_TRY(orth_f64(c, m, kstep, Yp.p, m, nullptr, &f1));             // rra.c:1690
    fb += f1;
    const bool reortho = is_new ? (step > 0 && step % 2 == 0) : (step > 0);  // :1703 / :1500
    if (reortho) {
      // Yp = Qp - Qj (Qj^T Qp); Qp = qr(Yp)   (rra.c:1716-1722)
      const int64_t nj = c0;
      RSVD_TRY(QtQ.alloc((size_t)(nj * kstep)));
      RSVD_TRY(gemm_f64(c, 'T', 'N', nj, kstep, m, 1.0, Q, ldq, Yp.p, m, 0.0, QtQ.p, nj));
      RSVD_TRY(gemm_f64(c, 'N', 'N', m, kstep, nj, -1.0, Q, ldq, QtQ.p, nj, 1.0, Yp.p, m));
      RSVD_TRY(orth_f64(c, m, kstep, Yp.p, m, nullptr, &f1));
      fb += f1;
    }
    // Bp = Qp^T A straight into rows c0.. of B; A -= Qp Bp with fused ||A||^2
    double* Bp = B + c0;
    RSVD_TRY(gemm_f64_tag(c, "apass_f64", 'T', 'N', kstep, n, m, 1.0, Yp.p, m, A.p, m, 0.0, Bp, ldb));  // :1741
    RSVD_TRY(gemm_f64_deflate(c, m, n, kstep, Yp.p, m, Bp, ldb, A.p, m, sums.p));        // :1750
    RSVD_TRY(copy2d_f64(c, m, kstep, Yp.p, m, Q + c0 * ldq, ldq));                       // :1760
    fr = (step + 1) * kstep;                                            // rra.c:1770
    steps = step + 1;
    if (tolMode) {                                                      // rra.c:1771-1778
      double ss = 0;
      RSVD_TRY(read_small(c, &ss, sums.p, sizeof ss));
      normval = sqrt(ss);
      if (normval < TOL) break;
    }
  }
  if (frank) *frank = fr;
  if (info) {
    info->steps = steps;
    info->qr_fallbacks = fb;
    info->last_norm = normval;
    info->kstep = kstep;
    info->nstep = nstep;
    info->tol_mode = tolMode;
  }
  return 0;
}

int blockrand_plan(int64_t m, int64_t n, int k, int p, int kstep, int* p_eff, int* nstep,
                   int* rankMode) {
  if (kstep <= 0) return set_error("blockrand: kstep must be positive (kstep=%d)", kstep);
  *rankMode = (k <= 0) ? 0 : 1;                                         // rra.c:251-256
  const int64_t mn = m < n ? m : n;
  if (p < kstep && (p + kstep) < mn) p = kstep;                        // rra.c:260-262
  *p_eff = p;
  *nstep = (k + p) / kstep;                                             // rra.c:263 (integer)
  return 0;
}

int blockrand_f64(Context* c, const double* M, int64_t m, int64_t n, int64_t ldm, int k, int p,
                  double TOL, int vnum, int kstep, int q, int s, const OmegaSrc& om,
                  BlockrandOut* out) {
  int p_eff = 0, nstep = 0, rankMode = 0;
  RSVD_TRY(blockrand_plan(m, n, k, p, kstep, &p_eff, &nstep, &rankMode));
  p = p_eff;
  // randQB_pb_new sizes (after its own kstep clamp / TOL-mode nstep)
  int ks = kstep, ns = nstep, tm = 0;
  RSVD_TRY(qb_adjust(m, n, &ks, &ns, 1, &tm));
  const int64_t lmax = (int64_t)ks * ns;
  DBuf<double> Q, B;
  RSVD_TRY(Q.alloc((size_t)(m * lmax)));
  RSVD_TRY(B.alloc((size_t)(lmax * n)));
  int frank = 0, lalloc = 0;
  QbInfo qi{};
  RSVD_TRY(qb_blocked_f64(c, M, m, n, ldm, kstep, nstep, TOL, q, s, 1, om, Q.p, m, B.p, lmax,
                          &frank, &lalloc, &qi));
  const int64_t l = qi.tol_mode ? frank : lalloc;                       // B->nrows
  if (rankMode == 1) frank = k;                                         // rra.c:273-279
  else frank = (int)llround((double)(frank / (frank + p + 1e-6)) * frank);
  out->frank = frank;
  out->l = (int)l;
  out->qb = qi;
  if (frank <= 0 || frank > l)
    return set_error("blockrand: rank %d outside [1, l=%lld] (k=%d p=%d kstep=%d)", frank,
                     (long long)l, k, p, kstep);
  RSVD_TRY(out->U.alloc((size_t)(m * frank)));
  RSVD_TRY(out->S.alloc((size_t)frank));
  RSVD_TRY(out->V.alloc((size_t)(n * frank)));
  SvdInfo si{};
  RSVD_TRY(svd_from_Q_f64(c, M, m, n, ldm, Q.p, m, l, frank, vnum, out->U.p, m, out->S.p,
                          out->V.p, n, &si));                           // rra.c:288-380
  out->svd = si;
  return 0;
}

}  // namespace rsvd
