// gemm_lp.hip -- low-precision matrix-core GEMM for the sketch and power-
// iteration passes over A (Y = A*Omega, Z = A^T*Y, Y = A*Z) and the
// split-precision projection B^T = A^T*Q of the mixed-precision rSVD.
//
// C (fp32) = alpha * op(A) * op(B) + beta * C.  Operands are read from HBM in
// their storage precision (fp32 here, so A is never copied) and rounded to
// bf16 while being staged through LDS, in a layout with the contraction index
// contiguous: [row][k] for op(A) and [col][k] for op(B), so each lane's MFMA
// fragment (8 consecutive k) is one 16-byte LDS read.
//
// MFMA: v_mfma_f32_16x16x32_bf16.  Lane l supplies srcA[i = l%16][k = 8(l/16)..+7]
// and srcB[k = 8(l/16)..+7][j = l%16]; the fp32 accumulator holds
// D[i = 4(l/16) + r][j = l%16] in register r.  As in gemm.hip, srcA is fed from
// op(B) and srcB from op(A), so accumulator lanes run along M and column-major
// stores of C are coalesced.
//
// SPLIT = 3 evaluates each product as hi*hi + hi*lo + lo*hi with x = hi + lo,
// hi = bf16(x), lo = bf16(x - hi): 16 significant bits per operand (a
// relative error up to ~2^-16 per product, which averages out in long sums),
// accumulated in fp32, at 3 bf16 MFMAs.  SPLIT = 6 splits each operand into
// three bf16 planes x = hi + mid + lo, which hold all 24 bits of an fp32
// value, and evaluates hi*hi + hi*mid + mid*hi + mid*mid + hi*lo + lo*hi (the
// dropped terms are below 2^-24 of the product): fp32-exact products,
// accumulated in fp32, at 6 bf16 MFMAs, with a K step of 32.
//
// Workgroup: 256 threads, 2x2 waves, tile 128 x 128, K step 64.  The grid is
// ordered so that the tiles sharing a panel of the large operand are launched
// together (their reads of that panel then hit the shared last-level cache).
#include <hip/hip_bf16.h>

#include "rsvd_internal.h"

namespace rsvd {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f4v_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int LBM = 128, LBN = 128, LPAD = 8, LTHREADS = 256;
// K step and bf16 planes per operand of a SPLIT mode
__host__ __device__ constexpr int lp_bk(int split) { return split == 6 ? 32 : 64; }
__host__ __device__ constexpr int lp_planes(int split) {
  return split == 6 ? 3 : (split == 3 ? 2 : 1);
}

// Staging of a 128 x LBK operand tile into LDS rows [row][k] (bf16, k
// contiguous), optionally as a (hi, lo) split.  Global view: element
// (row0 + rr, k0 + kk) at base[(k0+kk) + (row0+rr)*ld] when the operand is
// contiguous along k (KCONT), else at base[(row0+rr) + (k0+kk)*ld].  Loads are
// float4; `full` (uniform per workgroup) selects the unchecked path for tiles
// wholly inside the matrix with 16-byte aligned rows / columns.
typedef float f4s_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));

// Four consecutive k of one row into the NP planes at offset `off` of each
// (plane stride `ps` elements): plane 0 = bf16(x), plane 1 = bf16(x - p0),
// plane 2 = bf16(x - p0 - p1).  The differences are exact in fp32.  With
// hi_only (uniform) only plane 0 is formed and written.
template <int NP>
__device__ __forceinline__ void put4(__bf16* base, int ps, int off, f4s_t x, bool hi_only) {
  bf16x4_t h, m, l;
  if (NP > 1 && hi_only) {
#pragma unroll
    for (int e = 0; e < 4; ++e) h[e] = (__bf16)x[e];
    *reinterpret_cast<bf16x4_t*>(base + off) = h;
    return;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const __bf16 a = (__bf16)x[e];
    h[e] = a;
    if (NP > 1) {
      const float r1 = x[e] - (float)a;
      const __bf16 b = (__bf16)r1;
      m[e] = b;
      if (NP > 2) l[e] = (__bf16)(r1 - (float)b);
    }
  }
  *reinterpret_cast<bf16x4_t*>(base + off) = h;
  if (NP > 1) *reinterpret_cast<bf16x4_t*>(base + ps + off) = m;
  if (NP > 2) *reinterpret_cast<bf16x4_t*>(base + 2 * ps + off) = l;
}

// Row-contiguous operands are staged with the 16-byte k chunk c of LDS row r
// stored at chunk c ^ lds_swz(r).  The stager writes 4-row x 4-k micro-tiles
// (rows four apart across lanes), which at the padded row pitch alone put
// four lanes on one bank; the XOR spreads them over all banks and keeps the
// 16-byte fragment reads conflict-free.
__device__ __forceinline__ int lds_swz(int r) { return (r >> 3) & 3; }

template <bool KCONT, int SPLIT>
struct Stager;

// k-contiguous operand: an item is 4 consecutive k of one row; 16 lanes
// cover the 64 k of a row (256 contiguous bytes).  8 items per thread.
template <int SPLIT>
struct Stager<true, SPLIT> {
  static constexpr int LBK = lp_bk(SPLIT), LROW = LBK + LPAD, NP = lp_planes(SPLIT);
  static constexpr int ITEMS = (LBM * LBK / 4) / LTHREADS;
  f4s_t v[ITEMS];
  __device__ __forceinline__ void load(const float* __restrict__ base, int64_t ld, int64_t row0,
                                       int64_t nrows, int64_t k0, int64_t kend, int tid,
                                       bool full) {
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const int f = tid + it * LTHREADS;
      const int rr = f / (LBK / 4), kq = f % (LBK / 4);
      const int64_t gr = row0 + rr, gk = k0 + kq * 4;
      const float* p = base + gk + gr * ld;
      if (full) {
        v[it] = *reinterpret_cast<const f4s_t*>(p);
      } else {
        const bool rok = gr < nrows;
#pragma unroll
        for (int e = 0; e < 4; ++e) v[it][e] = (rok && gk + e < kend) ? p[e] : 0.0f;
      }
    }
  }
  // planes of LBM * LROW elements each, contiguous from `base`
  __device__ __forceinline__ void store(__bf16* __restrict__ base, int tid, bool hi_only) {
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const int f = tid + it * LTHREADS;
      const int rr = f / (LBK / 4), kq = f % (LBK / 4);
      put4<NP>(base, LBM * LROW, rr * LROW + kq * 4, v[it], hi_only);
    }
  }
};

// row-contiguous operand: an item is a 4-row x 4-k micro-tile (four float4
// loads down four columns), transposed in registers into four 4-k rows.
// Lane j of wave w takes row quad (j % 8) + 8 w and k quad j / 8 (+ 8 for
// the second item): 8 lanes read 128 contiguous bytes of a column, and the
// LDS writes spread over 32 bank pairs.
template <int SPLIT>
struct Stager<false, SPLIT> {
  static constexpr int LBK = lp_bk(SPLIT), LROW = LBK + LPAD, NP = lp_planes(SPLIT);
  static constexpr int ITEMS = (LBM * LBK / 16) / LTHREADS;
  f4s_t v[ITEMS][4];  // [item][column c] = rows 4rq..4rq+3 at k = 4kq + c
  __device__ __forceinline__ void coords(int tid, int it, int* rq, int* kq) const {
    const int lane = tid & 63, w = tid >> 6;
    *rq = (lane & 7) + 8 * w;
    *kq = (lane >> 3) + 8 * it;
  }
  __device__ __forceinline__ void load(const float* __restrict__ base, int64_t ld, int64_t row0,
                                       int64_t nrows, int64_t k0, int64_t kend, int tid,
                                       bool full) {
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      int rq, kq;
      coords(tid, it, &rq, &kq);
      const int64_t gr = row0 + 4 * rq, gk = k0 + 4 * kq;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float* p = base + gr + (gk + c) * ld;
        if (full) {
          v[it][c] = *reinterpret_cast<const f4s_t*>(p);
        } else {
          const bool kok = gk + c < kend;
#pragma unroll
          for (int e = 0; e < 4; ++e) v[it][c][e] = (kok && gr + e < nrows) ? p[e] : 0.0f;
        }
      }
    }
  }
  __device__ __forceinline__ void store(__bf16* __restrict__ base, int tid, bool hi_only) {
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      int rq, kq;
      coords(tid, it, &rq, &kq);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const f4s_t row = {v[it][0][e], v[it][1][e], v[it][2][e], v[it][3][e]};
        const int r = 4 * rq + e;
        const int off = r * LROW + ((((kq >> 1) ^ lds_swz(r)) << 3) | ((kq & 1) << 2));
        put4<NP>(base, LBM * LROW, off, row, hi_only);
      }
    }
  }
};

}  // namespace

// C[z] = alpha * op(A) op(B) (+ beta C), K range of slice z.  Two LDS stages:
// while the MFMAs read stage t, the registers loaded during step t-1 are
// written to stage t+1 and the loads for step t+2 are issued; one barrier
// per K step.
template <bool TA, bool TB, int SPLIT>
__global__ void __launch_bounds__(LTHREADS, 2)
gemm_bf16_kernel(int64_t M, int64_t N, int64_t K, int64_t kchunk, float alpha,
                 const float* __restrict__ A, int64_t lda, const float* __restrict__ B, int64_t ldb,
                 float beta, float* __restrict__ C, int64_t ldc, int64_t zstride, int n_fast,
                 int aligned, double* __restrict__ sumsq, int64_t klo) {
  constexpr int NBUF = lp_planes(SPLIT);  // bf16 planes per operand
  constexpr int LBK = lp_bk(SPLIT), LROW = LBK + LPAD;
  // LDS stages: one for the split modes (74 KB with two planes, 61 KB with
  // three), so that two workgroups share a CU and hide each other's staging
  // and epilogue (measured faster than two stages at one workgroup per CU);
  // two for plain bf16, which fits twice either way
  constexpr int NST = (SPLIT == 1) ? 2 : 1;
  __shared__ __attribute__((aligned(16))) __bf16 As[NST][NBUF][LBM * LROW];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[NST][NBUF][LBN * LROW];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;
  // XCD-aware tile order: workgroups are dispatched round-robin over the 8
  // XCDs (block b -> XCD b % 8).  Remap so that consecutive logical tiles --
  // the tiles sharing one panel of the large operand -- land on one XCD and
  // share its L2.
  int64_t tm, tn;
  {
    const int64_t nx = (int64_t)gridDim.x * gridDim.y;
    const int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    const int64_t per = nx / 8;
    const int64_t logical = (b < 8 * per) ? (b % 8) * per + b / 8 : b;
    const int64_t fx = logical % gridDim.x, fy = logical / gridDim.x;
    if (n_fast) { tn = fx; tm = fy; }
    else { tm = fx; tn = fy; }
  }
  const int64_t m0 = tm * LBM, n0 = tn * LBN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = (kbeg + kchunk < K) ? kbeg + kchunk : K;
  float* Cz = C + (int64_t)blockIdx.z * zstride;
  // unchecked loads when the whole tile is inside and rows are 16-B aligned
  const bool mn_full = aligned && (m0 + LBM <= M) && (n0 + LBN <= N);

  // op(A)(m, k): TA -> A[k + m*lda] (k-contiguous) ; !TA -> A[m + k*lda]
  // op(B)(k, n): TB -> B[n + k*ldb] (n-contiguous) ; !TB -> B[k + n*ldb] (k-contiguous)
  Stager<TA, SPLIT> sa;
  Stager<!TB, SPLIT> sb;
  // LDS offset of fragment (row, k chunk at kof) of op(A) / op(B): the
  // operands staged row-contiguously (!TA, TB) are chunk-swizzled
  auto offA = [&](int row, int kof) {
    return row * LROW + (TA ? kof : (((kof >> 3) ^ lds_swz(row)) << 3));
  };
  auto offB = [&](int row, int kof) {
    return row * LROW + (!TB ? kof : (((kof >> 3) ^ lds_swz(row)) << 3));
  };

  f4v_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4v_t{0.f, 0.f, 0.f, 0.f};

  const int64_t nsteps = (kend - kbeg + LBK - 1) / LBK;
  auto load_step = [&](int64_t t) {
    const int64_t k0 = kbeg + t * LBK;
    const bool full = mn_full && (k0 + LBK <= kend);
    sa.load(A, lda, m0, M, k0, kend, tid, full);
    sb.load(B, ldb, n0, N, k0, kend, tid, full);
  };
  // SPLIT = 6: a K step below klo takes the hi * hi product only, so only
  // plane 0 is staged and read for it
  auto hi_step = [&](int64_t t) { return SPLIT == 6 && kbeg + t * LBK < klo; };
  if (nsteps > 0) {
    load_step(0);
    sa.store(As[0][0], tid, hi_step(0));
    sb.store(Bs[0][0], tid, hi_step(0));
    if (nsteps > 1) load_step(1);
    __syncthreads();
  }
  for (int64_t t = 0; t < nsteps; ++t) {
    const int cur = (NST == 2) ? (int)(t & 1) : 0;
    if (hi_step(t)) {
#pragma unroll
      for (int ks = 0; ks < LBK / 32; ++ks) {
        const int kof = ks * 32 + lk * 8;
        bf16x8_t fa[4], fb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
          fb[a] = *reinterpret_cast<const bf16x8_t*>(&Bs[cur][0][offB(wn0 + a * 16 + li, kof)]);
#pragma unroll
        for (int b = 0; b < 4; ++b)
          fa[b] = *reinterpret_cast<const bf16x8_t*>(&As[cur][0][offA(wm0 + b * 16 + li, kof)]);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[a], fa[b], acc[a][b], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < LBK / 32; ++ks) {
        const int kof = ks * 32 + lk * 8;
        // fragments of plane P: fb[P][a] from op(B), fa[P][b] from op(A)
        bf16x8_t fa[NBUF][4], fb[NBUF][4];
#pragma unroll
        for (int P = 0; P < NBUF; ++P) {
#pragma unroll
          for (int a = 0; a < 4; ++a)
            fb[P][a] =
                *reinterpret_cast<const bf16x8_t*>(&Bs[cur][P][offB(wn0 + a * 16 + li, kof)]);
#pragma unroll
          for (int b = 0; b < 4; ++b)
            fa[P][b] =
                *reinterpret_cast<const bf16x8_t*>(&As[cur][P][offA(wm0 + b * 16 + li, kof)]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            f4v_t d = acc[a][b];
            if (SPLIT == 6) {  // smallest terms first
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0][a], fa[NBUF - 1][b], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[NBUF - 1][a], fa[0][b], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[1][a], fa[1][b], d, 0, 0, 0);
            }
            if (SPLIT == 3 || SPLIT == 6) {
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0][a], fa[1][b], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[1][a], fa[0][b], d, 0, 0, 0);
            }
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0][a], fa[0][b], d, 0, 0, 0);
          }
      }
    }
    if constexpr (NST == 2) {
      if (t + 1 < nsteps) {
        const int nxt = cur ^ 1;
        sa.store(As[nxt][0], tid, hi_step(t + 1));
        sb.store(Bs[nxt][0], tid, hi_step(t + 1));
        if (t + 2 < nsteps) load_step(t + 2);
      }
      __syncthreads();
    } else {
      __syncthreads();  // every wave is done with the stage
      if (t + 1 < nsteps) {
        sa.store(As[0][0], tid, hi_step(t + 1));
        sb.store(Bs[0][0], tid, hi_step(t + 1));
        if (t + 2 < nsteps) load_step(t + 2);
      }
      __syncthreads();
    }
  }

  double ss = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t gm = m0 + wm0 + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gn = n0 + wn0 + a * 16 + 4 * lk + r;
        if (gm < M && gn < N) {
          float* cp = Cz + gm + gn * ldc;
          float v = alpha * acc[a][b][r];
          if (beta != 0.0f) v += beta * (*cp);
          *cp = v;
          ss += (double)v * (double)v;
        }
      }
    }
  if (sumsq) {
    // FP64 sum of squares of the stored tile, one partial per workgroup
    __shared__ double red[LTHREADS / 64];
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
    if (lane == 0) red[wave] = ss;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < LTHREADS / 64; ++w) t += red[w];
      sumsq[blockIdx.x + (int64_t)gridDim.x * blockIdx.y] = t;
    }
  }
}

__global__ void splitk_reduce_f32_kernel(int64_t M, int64_t N, int splits,
                                         const float* __restrict__ W, float beta,
                                         float* __restrict__ C, int64_t ldc) {
  int64_t total = M * N;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % M, j = idx / M;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += W[(int64_t)z * total + idx];
    float* cp = C + i + j * ldc;
    if (beta != 0.0f) s += beta * (*cp);
    *cp = s;
  }
}

int gemm_bf16(Context* c, const char* tag, char ta, char tb, int64_t M, int64_t N, int64_t K,
              float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
              float* C, int64_t ldc, int split) {
  if (M < 0 || N < 0 || K < 0) return set_error("gemm_bf16: negative dimension");
  if (split != 1 && split != 3) return set_error("gemm_bf16: split must be 1 or 3");
  const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
  // skinny right operand and enough row blocks to fill the GPU: the
  // A-streaming kernel, which reads op(A) once
  if (M > 0 && N > 0 && K >= 32 && !tB && apass_suitable(c, M, N))
    return apass_bf16(c, tag, tA, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, split);
  return gemm_bf16_general(c, tag, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, split);
}

// The tiled kernel; sumsq_part (one double per tile, optional) receives the
// FP64 sums of squares of the stored tiles and disables the k split.
static int gemm_bf16_tiled(Context* c, const char* tag, char ta, char tb, int64_t M, int64_t N,
                           int64_t K, float alpha, const float* A, int64_t lda, const float* B,
                           int64_t ldb, float beta, float* C, int64_t ldc, int split,
                           double* sumsq_part, int64_t klo = 0) {
  const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
  if (split != 1 && split != 3 && split != 6)
    return set_error("gemm_bf16: split must be 1, 3 or 6");
  const int64_t gm = (M + LBM - 1) / LBM, gn = (N + LBN - 1) / LBN;
  const int n_fast = (gn <= gm) ? 1 : 0;
  const int64_t gx = n_fast ? gn : gm, gy = n_fast ? gm : gn;
  if (gy > 65535) return set_error("gemm_bf16: grid too large");
  int splits = 1;
  const int64_t tiles = gm * gn;
  const int64_t target = 2LL * (c->num_cus > 0 ? c->num_cus : 256);
  if (!sumsq_part && tiles < target && K >= 8 * 256) {
    int64_t maxs = K / 512;
    if (maxs > 32) maxs = 32;
    splits = pick_k_slices(tiles, target, maxs);  // two workgroups per CU
  }
  int64_t kchunk = K;
  if (splits > 1) {
    kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + lp_bk(split) - 1) / lp_bk(split) * lp_bk(split);
    splits = (int)((K + kchunk - 1) / kchunk);
  }
  DBuf<float> work;
  float* Cout = C;
  int64_t ldo = ldc, zstride = 0;
  float b_eff = beta;
  if (splits > 1) {
    RSVD_TRY(work.alloc((size_t)(M * N * splits)));
    Cout = work.p;
    ldo = M;
    zstride = M * N;
    b_eff = 0.0f;
  }
  dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)splits);
  // float4 loads need 16-byte aligned base pointers and leading dimensions
  const int aligned = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) % 16 == 0) &&
                      (lda % 4 == 0) && (ldb % 4 == 0);
  {
    KTimer kt(c, tag);
#define RSVD_LP_LAUNCH(TA_, TB_, S_)                                                         \
  hipLaunchKernelGGL((gemm_bf16_kernel<TA_, TB_, S_>), grid, dim3(LTHREADS), 0,        \
                     c->stream, M, N, K, kchunk, alpha, A, lda, B, ldb, b_eff, Cout, ldo,    \
                     zstride, n_fast, aligned, sumsq_part, klo)
    if (split == 1) {
      if (!tA && !tB) RSVD_LP_LAUNCH(false, false, 1);
      else if (tA && !tB) RSVD_LP_LAUNCH(true, false, 1);
      else if (!tA && tB) RSVD_LP_LAUNCH(false, true, 1);
      else RSVD_LP_LAUNCH(true, true, 1);
    } else if (split == 3) {
      if (!tA && !tB) RSVD_LP_LAUNCH(false, false, 3);
      else if (tA && !tB) RSVD_LP_LAUNCH(true, false, 3);
      else if (!tA && tB) RSVD_LP_LAUNCH(false, true, 3);
      else RSVD_LP_LAUNCH(true, true, 3);
    } else {
      if (!tA && !tB) RSVD_LP_LAUNCH(false, false, 6);
      else if (tA && !tB) RSVD_LP_LAUNCH(true, false, 6);
      else if (!tA && tB) RSVD_LP_LAUNCH(false, true, 6);
      else RSVD_LP_LAUNCH(true, true, 6);
    }
#undef RSVD_LP_LAUNCH
    RSVD_LAUNCH_CHECK();
  }
  if (splits > 1) {
    int64_t total = M * N;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(splitk_reduce_f32_kernel, dim3(blocks), dim3(256), 0, c->stream, M, N,
                       splits, (const float*)work.p, beta, C, ldc);
    RSVD_LAUNCH_CHECK();
  }
  return 0;
}

int gemm_bf16_general(Context* c, const char* tag, char ta, char tb, int64_t M, int64_t N,
                      int64_t K, float alpha, const float* A, int64_t lda, const float* B,
                      int64_t ldb, float beta, float* C, int64_t ldc, int split) {
  if (M < 0 || N < 0 || K < 0) return set_error("gemm_bf16: negative dimension");
  if (M == 0 || N == 0) return 0;
  if (K == 0) return gemm_f32(c, ta, tb, M, N, 0, alpha, A, lda, B, ldb, beta, C, ldc);
  return gemm_bf16_tiled(c, tag, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, split,
                         nullptr);
}

__global__ void sum_partials_lp_kernel(int64_t n, const double* __restrict__ part,
                                       double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

int gemm_bf16_deflate(Context* c, int64_t M, int64_t N, int64_t K, const float* Q, int64_t ldq,
                      const float* Bt, int64_t ldbt, float* C, int64_t ldc, double* sumsq,
                      int64_t klo) {
  if (M < 0 || N < 0 || K < 0) return set_error("deflate: negative dimension");
  if (M == 0 || N == 0) {
    RSVD_HIP(hipMemsetAsync(sumsq, 0, sizeof(double), c->stream));
    return 0;
  }
  if (K == 0) return set_error("deflate: empty block");
  const int64_t tiles = ((M + LBM - 1) / LBM) * ((N + LBN - 1) / LBN);
  DBuf<double> part;
  RSVD_TRY(part.alloc((size_t)tiles));
  if (klo % 32 != 0) return set_error("deflate: klo must be a multiple of 32");
  RSVD_TRY(gemm_bf16_tiled(c, "deflate_bf16x6", 'N', 'T', M, N, K, -1.0f, Q, ldq, Bt, ldbt, 1.0f,
                           C, ldc, 6, part.p, klo));
  hipLaunchKernelGGL(sum_partials_lp_kernel, dim3(1), dim3(256), 0, c->stream, tiles,
                     (const double*)part.p, sumsq);
  RSVD_LAUNCH_CHECK();
  return 0;
}

}  // namespace rsvd
