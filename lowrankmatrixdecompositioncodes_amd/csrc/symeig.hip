// symeig.hip -- symmetric eigensolver of the rSVD pipeline on the GPU, the
// device replacement of LAPACKE_dsyev on BB^T (MKL/mvf.c:1206-1209, vnum = 2,
// called from rra.c:191) and of compute_evals_and_evecs_of_symm_matrix.
//
// Two-sided cyclic block Jacobi.  The n columns are cut into blocks of B = 16;
// a round pairs the blocks with the circle method (every block meets every
// other once per sweep).  A round has two roles, which one launch
// (symj_step_kernel) runs side by side -- the apply of round r next to the
// solve of round r + 1:
//   solve : one workgroup per block pair forms the pair's S x S (S = 2B)
//           diagonal tile H of G and runs an inner sweep of two-sided Jacobi
//           on it (B disjoint rotations per step; all pairs in the first round
//           of a sweep, only the cross-block pairs afterwards), accumulating
//           the S x S orthogonal J with H <- J^T H J;
//   apply : every S x S tile of G in pair order becomes J_a^T T J_b (the
//           upper tiles are computed and mirrored, so G stays exactly
//           symmetric) and every S-row strip of V's pair columns becomes
//           V J -- hundreds of independent workgroups.
// Unlike one-sided Jacobi no Gram matrix is formed: the tile is read from G
// itself, and the O(n^2) work of a round is spread over the whole GPU.
// A sweep in which no pair met had |g_pq| > 1e-10 sqrt(|g_pp g_qq|) ends the
// loop: Jacobi converges quadratically, so that sweep's rotations leave the
// off-diagonals near the rounding floor.
// Off-diagonals below 4 eps ||G||_F are not rotated (they are below the
// absolute accuracy of any backward-stable eigensolver, LAPACK's included).
#include <cmath>
#include <cstring>

#include "jacobi_rot.h"
#include "rsvd_internal.h"

namespace rsvd {

namespace {


__device__ __forceinline__ int player(int pos, int r, int N) {
  return pos == 0 ? 0 : 1 + ((pos - 1 + r) % (N - 1));
}

// The columns of block pair w in round r: slot s (0..2B-1) is column
// first[s / B] + s % B, or -1 when the slot is padding (dummy block or past n).
template <int B>
struct PairCols {
  int64_t first[2];
  int64_t n;
  __device__ __forceinline__ PairCols(int w, int r, int N, int nblocks, int64_t n_) : n(n_) {
    const int ba = player(w, r, N), bb = player(N - 1 - w, r, N);
    first[0] = ba < nblocks ? (int64_t)ba * B : -1;
    first[1] = bb < nblocks ? (int64_t)bb * B : -1;
  }
  __device__ __forceinline__ int64_t col(int s) const {
    const int64_t f = first[s >= B];
    const int64_t c = f + (s & (B - 1));
    return (f >= 0 && c < n) ? c : -1;
  }
};

template <int B>
__device__ __forceinline__ int64_t pair_col(int w, int s, int r, int N, int nblocks, int64_t n) {
  return PairCols<B>(w, r, N, nblocks, n).col(s);
}

// Local pairing of step t of the inner sweep.  Round 0 of a sweep runs the
// full cyclic sweep of the 2B columns (S - 1 steps, every pair once, within
// and across the two blocks); the other rounds only meet the B x B cross pairs
// (B steps: local column k of block a with B + (k + t) mod B of block b) --
// the pairs inside one block were already met in round 0.  Every pair of
// columns still meets once per sweep, and the inner steps per sweep fall from
// about 2n to about n.
template <int B>
__device__ __forceinline__ void local_pair(int k, int t, bool cross, int* p, int* q) {
  constexpr int S = 2 * B;
  if (cross) {
    *p = k;
    *q = B + (k + t) % B;
  } else {
    *p = player(k, t, S);
    *q = player(S - 1 - k, t, S);
  }
}

// The inner sweep of one block pair on its S x S tile (S = 2B), in LDS: H in
// Ha (double-buffered with Hb), J accumulated in Jl (row-major, padded HP).
// A step: thread t = (bi, bj) (bi = t / B, bj = t % B) computes the rotation
// of local pair bj (the 16 lanes of each row of a wave compute the 16
// rotations), takes that of pair bi from the lane of its row that computed it
// (a cross-lane read, no LDS round trip and no extra barrier), applies
// R_bi^T . R_bj to the 2 x 2 block of H at (pair bi rows, pair bj columns) --
// the blocks partition H -- and R_bi to rows 2 bj, 2 bj + 1 of J's columns of
// pair bi.  One barrier per step.  Returns the buffer holding the final H.
template <int B, bool CROSS>
__device__ double* jac_inner_t(double* Ha, double* Hb, double* Jl, double flo, int* any_sig) {
  constexpr int S = 2 * B, HP = S + 1;
  static_assert(B == 16, "a row of 16 lanes holds the B rotations of a step");
  const int tid = threadIdx.x;
  const int bi = tid / B, bj = tid % B;
  const int src = (tid & 48) | bi;  // lane of this row that owns pair bi
  double* Hc = Ha;
  double* Hn = Hb;
  int sig = 0;
  constexpr int steps = CROSS ? B : S - 1;
  for (int t = 0; t < steps; ++t) {
    int pi, qi, pj, qj;
    local_pair<B>(bi, t, CROSS, &pi, &qi);
    local_pair<B>(bj, t, CROSS, &pj, &qj);
    const int x0 = 2 * bj;
    // every LDS read of the step first, so that one latency covers them
    const double aj = Hc[pj * HP + pj], bbj = Hc[qj * HP + qj], gj = Hc[pj * HP + qj];
    const double h00 = Hc[pi * HP + pj], h01 = Hc[pi * HP + qj];
    const double h10 = Hc[qi * HP + pj], h11 = Hc[qi * HP + qj];
    const double j00 = Jl[x0 * HP + pi], j01 = Jl[x0 * HP + qi];
    const double j10 = Jl[(x0 + 1) * HP + pi], j11 = Jl[(x0 + 1) * HP + qi];
    double cj, sj;
    int sgj;
    rotation(aj, bbj, gj, flo, &cj, &sj, &sgj);
    sig |= sgj;
    const double ci = __shfl(cj, src, 64), si = __shfl(sj, src, 64);
    // X = H_blk R_j, then R_i^T X
    const double x00 = cj * h00 - sj * h01, x01 = sj * h00 + cj * h01;
    const double x10 = cj * h10 - sj * h11, x11 = sj * h10 + cj * h11;
    double y00 = ci * x00 - si * x10, y01 = ci * x01 - si * x11;
    double y10 = si * x00 + ci * x10, y11 = si * x01 + ci * x11;
    if (bi == bj && si != 0.0) y01 = y10 = 0.0;  // the annihilated pair
    Hn[pi * HP + pj] = y00;
    Hn[pi * HP + qj] = y01;
    Hn[qi * HP + pj] = y10;
    Hn[qi * HP + qj] = y11;
    Jl[x0 * HP + pi] = ci * j00 - si * j01;
    Jl[x0 * HP + qi] = si * j00 + ci * j01;
    Jl[(x0 + 1) * HP + pi] = ci * j10 - si * j11;
    Jl[(x0 + 1) * HP + qi] = si * j10 + ci * j11;
    __syncthreads();
    double* tmp = Hc;
    Hc = Hn;
    Hn = tmp;
  }
  *any_sig = sig;
  return Hc;
}

template <int B>
__device__ double* jac_inner(double* Ha, double* Hb, double* Jl, bool cross, double flo,
                             int* any_sig) {
  return cross ? jac_inner_t<B, true>(Ha, Hb, Jl, flo, any_sig)
               : jac_inner_t<B, false>(Ha, Hb, Jl, flo, any_sig);
}

// J and the final tile H of pair w out (column-major S x S each); the flag
// collects the rotation-size bits of rotation() (bit 0: above kStopTol, bit
// 1: a rotation the loose stop counts).
template <int B>
__device__ void jac_out(const double* Hf, const double* Jl, int w, double* __restrict__ Jout,
                        double* __restrict__ Hout, int any_sig,
                        unsigned long long* __restrict__ flag) {
  constexpr int S = 2 * B, HP = S + 1;
  const int tid = threadIdx.x;
  double* __restrict__ Jw = Jout + (size_t)w * S * S;
  double* __restrict__ Hw = Hout + (size_t)w * S * S;
  for (int e = tid; e < S * S; e += 256) {
    const int i = e % S, j = e / S;
    Jw[i + j * S] = Jl[i * HP + j];
    Hw[i + j * S] = Hf[i * HP + j];
  }
  const unsigned long long m =
      (__any(any_sig & 1) ? 1ULL : 0ULL) | (__any(any_sig & 2) ? 2ULL : 0ULL);
  if (m && (tid & 63) == 0) atomicOr(flag, m);
}

// First round of the iteration: the pair tiles are read from G.
template <int B>
__global__ void __launch_bounds__(256)
symj_solve_kernel(int64_t n, const double* __restrict__ G, int64_t ldg, int nblocks, int N,
                  int round, const double* __restrict__ fro2, double* __restrict__ Jout,
                  double* __restrict__ Hout, unsigned long long* __restrict__ flag) {
  constexpr int S = 2 * B, HP = S + 1;
  __shared__ double Ha[S * HP], Hb[S * HP], Jl[S * HP];
  __shared__ int64_t cols[S];
  const int w = blockIdx.x;
  const int tid = threadIdx.x;
  if (tid < S) cols[tid] = pair_col<B>(w, tid, round, N, nblocks, n);
  __syncthreads();
  for (int e = tid; e < S * S; e += 256) {
    const int i = e % S, j = e / S;
    const int64_t ci = cols[i], cj = cols[j];
    Ha[i * HP + j] = (ci >= 0 && cj >= 0) ? G[ci + cj * ldg] : 0.0;
    Jl[i * HP + j] = (i == j) ? 1.0 : 0.0;
  }
  const double flo = 4.0 * 2.220446049250313e-16 * sqrt(*fro2);
  __syncthreads();
  int sig = 0;
  const double* Hf = jac_inner<B>(Ha, Hb, Jl, round > 0, flo, &sig);
  jac_out<B>(Hf, Jl, w, Jout, Hout, sig, flag);
}

typedef double d4_t __attribute__((ext_vector_type(4)));

// The 16 x 16 block (bm, bn) of the product op(P) Q with inner dimension 32,
// P and Q in LDS (row-major, leading dimensions lp and lq; op(P) = P^T reads
// P[k][m]), on the FP64 matrix cores: eight v_mfma_f64_16x16x4f64 steps along
// k.  srcA carries Q[k][n] and srcB op(P)[m][k] (lane l: n or m = l % 16,
// k = 4 s + l / 16), so lane l's accumulator register r holds
// C[16 bm + l % 16][16 bn + 4 r + l / 16].
template <bool TA>
__device__ __forceinline__ d4_t blk_mfma(const double* __restrict__ P, int lp,
                                         const double* __restrict__ Q, int lq, int bm, int bn) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int k = 4 * s + lk;
    const double q = Q[k * lq + 16 * bn + li];
    const double p = TA ? P[k * lp + 16 * bm + li] : P[(16 * bm + li) * lp + k];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(q, p, acc, 0, 0, 0);
  }
  return acc;
}

// Stores block (bm, bn) (the layout of blk_mfma) into C (leading dimension ldc).
__device__ __forceinline__ void blk_store(const d4_t& acc, double* __restrict__ C, int ldc,
                                          int bm, int bn) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) C[(16 * bm + li) * ldc + 16 * bn + 4 * r + lk] = acc[r];
}

// The wave's 16 x 16 quadrant (wave / 2, wave % 2) of the 32 x 32 product
// op(P) Q, P and Q padded to HP = 33.  Two LDS reads per 16 multiply-adds a
// lane, where a register-blocked scalar product reads one per two.
template <bool TA>
__device__ __forceinline__ d4_t tile_mfma(const double* __restrict__ P,
                                          const double* __restrict__ Q) {
  const int wave = threadIdx.x >> 6;
  return blk_mfma<TA>(P, 33, Q, 33, wave >> 1, wave & 1);
}

__device__ __forceinline__ void tile_store(const d4_t& acc, double* __restrict__ C) {
  const int wave = threadIdx.x >> 6;
  blk_store(acc, C, 33, wave >> 1, wave & 1);
}

// Apply round r's rotations (the apply role of symj_step_kernel): blocks
// [0, ntg) take the upper pair tiles (w1 <= w2) of Gin and write J_w1^T T J_w2
// to both triangles of Gout; the rest take 2S-row strips of V (in place):
// V(rows, cols(w)) <- V(rows, cols(w)) J_w.  Every element of Gout and V is
// written by exactly one workgroup.  Every global read of a workgroup is
// issued before its first wait (one memory latency), ahead of the skip test;
// padding slots read an in-bounds element and drop it.
template <int B>
__device__ void jac_apply(int bid, int64_t n, const double* __restrict__ Gin,
                          double* __restrict__ Gout, int64_t ldg, double* __restrict__ V,
                          int64_t ldv, int nblocks, int N, int round,
                          const double* __restrict__ Jall, int ntg, double* T, double* Ja,
                          double* Jb, double* X, bool skip) {
  static_assert(B == 16, "the tile products are 32 x 32, one 16 x 16 quadrant a wave");
  constexpr int S = 2 * B, HP = S + 1;
  const int pairs = N / 2;
  const int tid = threadIdx.x;
  if (bid < ntg) {
    int w1 = 0;
    while (bid >= pairs - w1) {
      bid -= pairs - w1;
      ++w1;
    }
    const int w2 = w1 + bid;
    const PairCols<B> pa(w1, round, N, nblocks, n), pb(w2, round, N, nblocks, n);
    const double* J1 = Jall + (size_t)w1 * S * S;
    const double* J2 = Jall + (size_t)w2 * S * S;
    // thread: row r of the tile, columns c0 + 8 it
    const int r = tid % S, c0 = tid / S;
    const int64_t gr = pa.col(r);
    double tv[4], av[4], bv[4];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int c = c0 + 8 * it;
      const int64_t gc = pb.col(c);
      const double g = Gin[(gr < 0 ? 0 : gr) + (gc < 0 ? 0 : gc) * ldg];
      tv[it] = (gr >= 0 && gc >= 0) ? g : 0.0;
      av[it] = J1[r + c * S];
      bv[it] = J2[r + c * S];
    }
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int c = c0 + 8 * it;
      T[r * HP + c] = tv[it];
      Ja[r * HP + c] = av[it];
      Jb[r * HP + c] = bv[it];
    }
    if (skip) return;
    __syncthreads();
    tile_store(tile_mfma<false>(T, Jb), X);  // X = T J_w2
    __syncthreads();
    // Y = J_w1^T X into T (T is not read again: its last reads were before
    // the barrier above)
    tile_store(tile_mfma<true>(Ja, X), T);
    __syncthreads();
    if (w1 == w2) {
      // diagonal pair tile: the upper triangle, mirrored
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int c = c0 + 8 * it;
        const int64_t gc = pb.col(c);
        if (gr >= 0 && gc >= 0) Gout[gr + gc * ldg] = (r <= c) ? T[r * HP + c] : T[c * HP + r];
      }
    } else {
      const int64_t gr2 = pb.col(r);
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int c = c0 + 8 * it;
        const int64_t gc = pb.col(c), gc2 = pa.col(c);
        if (gr >= 0 && gc >= 0) Gout[gr + gc * ldg] = T[r * HP + c];
        // mirror: row cols(w2)[r], column cols(w1)[c]
        if (gr2 >= 0 && gc2 >= 0) Gout[gr2 + gc2 * ldg] = T[c * HP + r];
      }
    }
    return;
  }
  if (V == nullptr) return;
  // V strip: rows [r0, r0 + 2S) of the pair's columns, in two S-row halves
  // (T and X).  Thread: row r of the strip, columns cq + 4 it; J_w element
  // (tid % S, tid / S + 8 it).
  bid -= ntg;
  const int w = bid % pairs;
  const int64_t r0 = (int64_t)(bid / pairs) * (2 * S);
  const PairCols<B> pc(w, round, N, nblocks, n);
  const double* Jw = Jall + (size_t)w * S * S;
  const int r = tid % (2 * S), cq = tid / (2 * S);
  const bool row_ok = r0 + r < n;
  double vv[8], jv[4];
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    const int64_t gc = pc.col(cq + 4 * it);
    const double v = V[(row_ok ? r0 + r : 0) + (gc < 0 ? 0 : gc) * ldv];
    vv[it] = (gc >= 0 && row_ok) ? v : 0.0;
  }
  const int jr = tid % S, jc0 = tid / S;
#pragma unroll
  for (int it = 0; it < 4; ++it) jv[it] = Jw[jr + (jc0 + 8 * it) * S];
  double* Vh = r < S ? T + r * HP : X + (r - S) * HP;  // this thread's row of the strip
#pragma unroll
  for (int it = 0; it < 8; ++it) Vh[cq + 4 * it] = vv[it];
#pragma unroll
  for (int it = 0; it < 4; ++it) Jb[jr * HP + jc0 + 8 * it] = jv[it];
  if (skip) return;
  __syncthreads();
  const d4_t o1 = tile_mfma<false>(T, Jb), o2 = tile_mfma<false>(X, Jb);
  __syncthreads();
  tile_store(o1, T);
  tile_store(o2, X);
  __syncthreads();
  if (!row_ok) return;
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    const int64_t gc = pc.col(cq + 4 * it);
    if (gc >= 0) V[(r0 + r) + gc * ldv] = Vh[cq + 4 * it];
  }
}

// One launch per round: the apply of round r (blocks nsolve..) and, in the
// same launch, the inner sweeps of round r1 = the next round (blocks
// 0..nsolve-1), so a round costs max(apply, solve) instead of their sum.  A
// solve workgroup builds its pair's tile of G_{r+1} = J_r^T G_r J_r itself:
// its two blocks a and b sat in pairs wa and wb of round r, so the diagonal
// blocks are sub-blocks of those pairs' final tiles (Hin, kept by round r's
// solves) and the off-diagonal block is Ja^T G_r(cols(wa), cols(wb)) Jb with
// Ja, Jb the columns of J_wa, J_wb that belong to a and b.  G is read from
// Gin and written to Gout (double-buffered), J and the tiles alternate too.
template <int B>
__global__ void __launch_bounds__(256)
symj_step_kernel(int64_t n, const double* __restrict__ Gin, double* __restrict__ Gout,
                 int64_t ldg, double* __restrict__ V, int64_t ldv, int nblocks, int N, int round,
                 int round1, int nsolve, const double* __restrict__ fro2,
                 const double* __restrict__ Jin, const double* __restrict__ Hin,
                 double* __restrict__ Jout, double* __restrict__ Hout, int ntg,
                 unsigned long long* __restrict__ flag,
                 const unsigned long long* __restrict__ prev, unsigned long long mask) {
  constexpr int S = 2 * B, HP = S + 1, BP = B + 1;
  // Device-side stop: every launch of a sweep reads the previous sweep's
  // rotation flags (complete: that sweep's launches have finished) and does
  // nothing when they say it converged, so the host can queue a sweep before
  // it has read the last one's flags.  Read here and tested after the other
  // global reads are issued (before any global write), so one memory latency
  // covers them all.
  const unsigned long long pv = prev ? *prev : mask;
  const bool skip = (pv & mask) == 0;
  // Apply role: four S x HP regions T, Ja, Jb, X.  Solve role: Ha, Jl and
  // Tab (later Hb) S x HP each, then JBq, X3 and JAq S x BP each; no region
  // changes tenant before the inner sweep.  Both footprints let four
  // workgroups share a CU.
  constexpr int kSm = (4 * S * HP > 3 * S * HP + 3 * S * BP) ? 4 * S * HP : 3 * S * HP + 3 * S * BP;
  __shared__ double sm[kSm];
  const int tid = threadIdx.x;
  if ((int)blockIdx.x >= nsolve) {
    jac_apply<B>((int)blockIdx.x - nsolve, n, Gin, Gout, ldg, V, ldv, nblocks, N, round, Jin, ntg,
                 sm, sm + S * HP, sm + 2 * S * HP, sm + 3 * S * HP, skip);
    return;
  }
  double* Ha = sm;
  double* Jl = sm + S * HP;
  double* Tab = sm + 2 * S * HP;
  double* Hb = Tab;  // from the first inner step on
  double* JBq = sm + 3 * S * HP;  // JBq[i * BP + c] = J_wb(i, hb B + c)
  double* X3 = JBq + S * BP;      // Tab JBq
  double* JAq = X3 + S * BP;      // JAq[i * BP + c] = J_wa(i, ha B + c)
  const int w = blockIdx.x;
  // the blocks of pair w in round1 and where they sat in round r
  const int xa = player(w, round1, N), xb = player(N - 1 - w, round1, N);
  auto locate = [&](int x, int* wp, int* h) {
    const int pos = (x == 0) ? 0 : 1 + (((x - 1 - round) % (N - 1)) + (N - 1)) % (N - 1);
    if (pos < N / 2) {
      *wp = pos;
      *h = 0;
    } else {
      *wp = N - 1 - pos;
      *h = 1;
    }
  };
  int wa, ha, wb, hb;
  locate(xa, &wa, &ha);
  locate(xb, &wb, &hb);
  const PairCols<B> pa(wa, round, N, nblocks, n), pb(wb, round, N, nblocks, n);
  const double* JA = Jin + (size_t)wa * S * S;
  const double* JB = Jin + (size_t)wb * S * S;
  const double* HA = Hin + (size_t)wa * S * S;
  const double* HB = Hin + (size_t)wb * S * S;
  // every global read first: thread row i of Tab (columns j0 + 8 it), of JAq
  // and JBq (columns j0 + 8 it), and entry (dc, dd) of the two diagonal blocks
  const int i = tid % S, j0 = tid / S;
  const int dc = tid % B, dd = tid / B;
  const int64_t gi = pa.col(i);
  double tv[4], av[2], bv[2];
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    // padding slots read column 0 / row 0 (in bounds) and drop the value
    const int64_t gj = pb.col(j0 + 8 * it);
    const double g = Gin[(gi < 0 ? 0 : gi) + (gj < 0 ? 0 : gj) * ldg];
    tv[it] = (gi >= 0 && gj >= 0) ? g : 0.0;
  }
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    av[it] = JA[i + (ha * B + j0 + 8 * it) * S];
    bv[it] = JB[i + (hb * B + j0 + 8 * it) * S];
  }
  const double hav = HA[(ha * B + dc) + (ha * B + dd) * S];
  const double hbv = HB[(hb * B + dc) + (hb * B + dd) * S];
  // the rotation floor goes through LDS: its load then completes with the
  // others instead of being deferred to its first use in the inner sweep
  __shared__ double flo_s;
  if (tid == 0) flo_s = 4.0 * 2.220446049250313e-16 * sqrt(*fro2);
  // J starts as the identity (its region has no other tenant)
#pragma unroll
  for (int it = 0; it < 4; ++it) Jl[i * HP + j0 + 8 * it] = (i == j0 + 8 * it) ? 1.0 : 0.0;
#pragma unroll
  for (int it = 0; it < 4; ++it) Tab[i * HP + j0 + 8 * it] = tv[it];
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    JAq[i * BP + j0 + 8 * it] = av[it];
    JBq[i * BP + j0 + 8 * it] = bv[it];
  }
  // diagonal blocks of the new tile from round r's final tiles
  Ha[dc * HP + dd] = hav;
  Ha[(B + dc) * HP + B + dd] = hbv;
  // the skip test after the LDS stores, so that no global read waits for it
  if (skip) return;
  __syncthreads();
  // off-diagonal block H_ab = J_a^T G_r(cols(wa), cols(wb)) J_b on the FP64
  // matrix cores: X3 = Tab JBq (S x B, waves 0 and 1), then H_ab = JAq^T X3
  // (B x B, wave 0), mirrored into H_ba
  const int wave = tid >> 6;
  if (wave < 2) blk_store(blk_mfma<false>(Tab, HP, JBq, BP, wave, 0), X3, BP, wave, 0);
  __syncthreads();
  if (wave == 0) {
    const d4_t h = blk_mfma<true>(JAq, BP, X3, BP, 0, 0);
    const int lane = tid & 63, c = lane & 15, d0 = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Ha[c * HP + B + 4 * r + d0] = h[r];
      Ha[(B + 4 * r + d0) * HP + c] = h[r];
    }
  }
  __syncthreads();
  const double flo = flo_s;
  int sig = 0;
  const double* Hf = jac_inner<B>(Ha, Hb, Jl, round1 > 0, flo, &sig);
  jac_out<B>(Hf, Jl, w, Jout, Hout, sig, flag);
}

__global__ void symj_identity_kernel(int64_t n, double* __restrict__ V, int64_t ldv) {
  const int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % n, j = idx / n;
    V[i + j * ldv] = (i == j) ? 1.0 : 0.0;
  }
}

__global__ void symj_diag_kernel(int64_t n, const double* __restrict__ G, int64_t ldg,
                                 double* __restrict__ d) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j < n) d[j] = G[j + j * ldg];
}

template <int B>
int run_sweeps(Context* c, int64_t n, double* G, int64_t ldg, double* V, int64_t ldv,
               int max_sweeps, int loose, int* sweep_out) {
  constexpr int S = 2 * B;
  const int nblocks = (int)((n + B - 1) / B);
  const int N = nblocks + (nblocks & 1);
  const int pairs = N / 2;
  const int ntg = pairs * (pairs + 1) / 2;
  const int ntv = V ? pairs * (int)((n + 2 * S - 1) / (2 * S)) : 0;
  DBuf<double> G2, Jb, Hb, fro2;
  DBuf<unsigned long long> flags;
  const size_t js = (size_t)pairs * S * S;
  RSVD_TRY(G2.alloc((size_t)(ldg * n)));
  RSVD_TRY(Jb.alloc(2 * js));
  RSVD_TRY(Hb.alloc(2 * js));
  RSVD_TRY(fro2.alloc(1));
  RSVD_TRY(flags.alloc((size_t)max_sweeps + 2));
  RSVD_TRY(sumsq_f64(c, n, n, G, ldg, fro2.p));
  RSVD_HIP(hipMemsetAsync(flags.p, 0, sizeof(unsigned long long) * (max_sweeps + 2), c->stream));
  double* cur = G;
  double* nxt = G2.p;
  int jc = 0;
  int sweep = 0;
  KTimer kt(c, "symeig");
  if (n > 1) {
    // round 0 of sweep 0 from G itself; afterwards every launch applies
    // round r and solves the next round
    hipLaunchKernelGGL(symj_solve_kernel<B>, dim3(pairs), dim3(256), 0, c->stream, n,
                       (const double*)G, ldg, nblocks, N, 0, (const double*)fro2.p, Jb.p, Hb.p,
                       flags.p);
    RSVD_LAUNCH_CHECK();
    // The first launches of sweep s are queued before the host looks at
    // sweep s - 1's flags (copied to pinned memory behind sweep s - 1 and
    // waited for while those launches run); when s - 1 converged they return
    // at once (the device-side stop above), the rest of the sweep is never
    // queued, and the state is the one sweep s started from.
    const unsigned long long mask = loose ? 2ULL : 1ULL;
    constexpr int kAhead = 4;  // launches of sweep s queued before the wait
    volatile unsigned long long* slot = static_cast<unsigned long long*>(c->pinned);
    hipEvent_t ev[2] = {nullptr, nullptr};
    RSVD_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    RSVD_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int rc = 0;
    bool done = false;
    for (; sweep < max_sweeps && rc == 0 && !done; ++sweep) {
      double* cur0 = cur;
      const int jc0 = jc;
      const unsigned long long* prev = sweep > 0 ? flags.p + sweep - 1 : nullptr;
      for (int r = 0; r < N - 1; ++r) {
        const bool wrap = r + 1 == N - 1;
        const int r1 = wrap ? 0 : r + 1;
        unsigned long long* fl = flags.p + (wrap ? sweep + 1 : sweep);
        hipLaunchKernelGGL(symj_step_kernel<B>, dim3(pairs + ntg + ntv), dim3(256), 0, c->stream,
                           n, (const double*)cur, nxt, ldg, V, ldv, nblocks, N, r, r1, pairs,
                           (const double*)fro2.p, (const double*)(Jb.p + jc * js),
                           (const double*)(Hb.p + jc * js), Jb.p + (1 - jc) * js,
                           Hb.p + (1 - jc) * js, ntg, fl, prev, mask);
        if (hipGetLastError() != hipSuccess) {
          rc = set_error("symeig: step launch failed");
          break;
        }
        double* tmp = cur;
        cur = nxt;
        nxt = tmp;
        jc = 1 - jc;
        // after the first kAhead launches (or the whole sweep, if shorter):
        // sweep - 1's verdict
        if (sweep > 0 && (r + 1 == kAhead || (r + 1 == N - 1 && N - 1 < kAhead))) {
          if (hipEventSynchronize(ev[(sweep - 1) & 1]) != hipSuccess) {
            rc = set_error("symeig: flag wait failed");
            break;
          }
          // no pair met in sweep - 1 was above kStopTol (or counted by the
          // loose stop): this sweep's launches did nothing
          if ((slot[(sweep - 1) & 1] & mask) == 0) {
            cur = cur0;
            jc = jc0;
            done = true;
            break;
          }
        }
      }
      if (rc || done) break;
      if (hipMemcpyAsync((void*)(slot + (sweep & 1)), flags.p + sweep, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
          hipEventRecord(ev[sweep & 1], c->stream) != hipSuccess) {
        rc = set_error("symeig: flag copy failed");
        break;
      }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) return rc;
  }
  if (cur != G)
    RSVD_HIP(hipMemcpy2DAsync(G, ldg * sizeof(double), cur, ldg * sizeof(double),
                              n * sizeof(double), n, hipMemcpyDeviceToDevice, c->stream));
  *sweep_out = sweep;
  return 0;
}

}  // namespace

int symeig_jacobi_f64(Context* c, int64_t n, double* G, int64_t ldg, double* V, int64_t ldv,
                      double* evals, int max_sweeps, int* sweeps_out, int loose) {
  if (sweeps_out) *sweeps_out = 0;
  if (n <= 0) return 0;
  if (ldg < n || (V && ldv < n)) return set_error("symeig: leading dimension < n");
  if (V) {
    int blocks = (int)((n * n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(symj_identity_kernel, dim3(blocks), dim3(256), 0, c->stream, n, V, ldv);
    RSVD_LAUNCH_CHECK();
  }
  int sweep = 0;
  // B = 16: each thread owns one 2x2 block and computes two rotations per
  // step; a larger B trades launches for a slower, LDS-bound inner sweep
  const int rc = run_sweeps<16>(c, n, G, ldg, V, ldv, max_sweeps, loose, &sweep);
  if (rc) return rc;
  if (evals) {
    hipLaunchKernelGGL(symj_diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, n, (const double*)G, ldg, evals);
    RSVD_LAUNCH_CHECK();
  }
  if (sweeps_out) *sweeps_out = sweep;
  return 0;
}

}  // namespace rsvd
