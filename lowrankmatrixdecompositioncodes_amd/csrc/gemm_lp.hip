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
#include <type_traits>

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

// XCD-aware tile order: workgroups are dispatched round-robin over the 8 XCDs
// (block b -> XCD b % 8).  Remap so that consecutive logical tiles -- the
// tiles sharing one panel of the large operand -- land on one XCD and share
// its L2.  n_fast: the column-tile index runs fastest (over gridDim.x).
__device__ __forceinline__ void xcd_tile(int n_fast, int64_t* tm, int64_t* tn) {
  const int64_t nx = (int64_t)gridDim.x * gridDim.y;
  const int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  const int64_t per = nx / 8;
  const int64_t logical = (b < 8 * per) ? (b % 8) * per + b / 8 : b;
  const int64_t fx = logical % gridDim.x, fy = logical / gridDim.x;
  if (n_fast) { *tn = fx; *tm = fy; }
  else { *tm = fx; *tn = fy; }
}

// C[z] = alpha * op(A) op(B) (+ beta C), K range of slice z.  Two LDS stages:
// while the MFMAs read stage t, the registers loaded during step t-1 are
// written to stage t+1 and the loads for step t+2 are issued; one barrier
// per K step.
template <bool TA, bool TB, int SPLIT>
__global__ void __launch_bounds__(LTHREADS, 2)
gemm_bf16_kernel(int64_t M, int64_t N, int64_t K, int64_t kchunk, float alpha,
                 const float* __restrict__ A, int64_t lda, const float* __restrict__ B, int64_t ldb,
                 float beta, float* __restrict__ C, int64_t ldc, int64_t zstride, int n_fast,
                 int aligned, double* __restrict__ sumsq, int64_t klo, int sym) {
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
  int64_t tm, tn;
  xcd_tile(n_fast, &tm, &tn);
  // a Gram (sym): the tiles below the diagonal are the mirror of those above
  if (sym && tm > tn) return;
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

// C(i, j) <- C(j, i) for i > j: the strictly lower triangle from the upper one
__global__ void mirror_lower_f32_kernel(int64_t n, float* __restrict__ C, int64_t ldc) {
  const int64_t total = n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx % n, j = idx / n;
    if (i > j) C[i + j * ldc] = C[j + i * ldc];
  }
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
  // A^T A with beta = 0: only the tiles on and above the diagonal are formed,
  // the lower triangle is copied from the upper one afterwards
  const bool sym = A == B && lda == ldb && tA && !tB && M == N && beta == 0.0f && !sumsq_part;
  int splits = 1;
  const int64_t tiles = sym ? gm * (gm + 1) / 2 : gm * gn;
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
                     zstride, n_fast, aligned, sumsq_part, klo, sym ? 1 : 0)
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
  if (sym) {
    const int64_t total = M * M;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(mirror_lower_f32_kernel, dim3(blocks), dim3(256), 0, c->stream, M, C,
                       ldc);
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

// out[b] = sum of part[b * chunk .. (b + 1) * chunk) (clipped to n), one
// workgroup per b: a fixed order, so the FP64 total is the same on every run.
__global__ void sum_partials_lp_kernel(int64_t n, int64_t chunk, const double* __restrict__ part,
                                       double* __restrict__ out) {
  __shared__ double red[256];
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = (lo + chunk < n) ? lo + chunk : n;
  double s = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// ---- Deflation C -= Q B^T on an LDS ring -------------------------------------
// The deflation's operands are small (Q: M x K, B^T: N x K, K = one or two QB
// blocks) and its output is the whole residual.  The operands are split into
// the three bf16 planes once per call, in MFMA fragment order
// (split3_frag_kernel), and each 128 x 64 tile streams them into two LDS
// stages by direct loads (global_load_lds_dwordx4, no registers) one k block
// ahead; 73 KiB of LDS, so that two workgroups share a CU and overlap each
// other's prologue and stores.  Every accumulator takes the products of the
// tiled kernel in the same order, so C is the same to the bit.
typedef __attribute__((address_space(3))) void* lds_ptr_t;
typedef __attribute__((address_space(1))) void* gbl_ptr_t;

// X (rows x K, column-major, ld) -> out[((P * KB + kb) * RT + t) * 64 + lane]
// = plane P of rows 16 t + lane % 16, k = 32 kb + 8 (lane / 16) .. + 7
// (planes as put4; rows past `rows` are zero).  Planes 1 and 2 are written
// for k blocks at or above kblo only.
__global__ void split3_frag_kernel(int64_t rows, int64_t KB, int64_t RT, int64_t kblo,
                                   const float* __restrict__ X, int64_t ldx,
                                   bf16x8_t* __restrict__ out) {
  const int64_t total = KB * RT * 64;
  const int64_t plane = total;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * 256) {
    const int lane = (int)(idx & 63);
    const int64_t t = (idx >> 6) % RT, kb = (idx >> 6) / RT;
    const int64_t r = 16 * t + (lane & 15);
    const int64_t k0 = 32 * kb + 8 * (lane >> 4);
    bf16x8_t h, m, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float x = (r < rows) ? X[r + (k0 + e) * ldx] : 0.0f;
      const __bf16 a = (__bf16)x;
      const float r1 = x - (float)a;
      const __bf16 b = (__bf16)r1;
      h[e] = a;
      m[e] = b;
      l[e] = (__bf16)(r1 - (float)b);
    }
    out[idx] = h;
    if (kb >= kblo) {
      out[plane + idx] = m;
      out[2 * plane + idx] = l;
    }
  }
}

constexpr int kDBN = 64;  // ring tile columns (4 row tiles of B)

// C -= Q B^T over k blocks [0, KB) (hi * hi only below kblo), C's FP64 sum of
// squares per tile into sumsq[tile].  Qs / Bs: split3_frag_kernel layouts
// with RTQ / RTB row tiles (multiples of 8 and 4).
__global__ void __launch_bounds__(LTHREADS, 2)
deflate_ring_kernel(int64_t M, int64_t N, int KB, int kblo, const bf16x8_t* __restrict__ Qs,
                    int64_t RTQ, const bf16x8_t* __restrict__ Bs, int64_t RTB,
                    float* __restrict__ C, int64_t ldc, int n_fast,
                    double* __restrict__ sumsq) {
  // [stage][chunk][lane], chunk = plane * 12 + row tile (Q: 0..7, B: 8..11)
  __shared__ bf16x8_t Ls[2][36][64];
  // A direct load's LDS target is picked between its ring slot and Ldummy on
  // a condition that always holds but that the compiler cannot prove; the
  // target then carries no LDS memory operand and the compiler adds no wait
  // for the loads in flight before each LDS read (it cannot tell the stages
  // apart).  Ordering comes from the explicit waits and barrier.
  __shared__ bf16x8_t Ldummy[64];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lk = lane >> 4;
  int64_t tm, tn;
  xcd_tile(n_fast, &tm, &tn);
  const int64_t m0 = tm * LBM, n0 = tn * kDBN;
  const int wa = (wave & 1) * 4, wb = (wave >> 1) * 2;  // the wave's first Q / B row tile
  const bf16x8_t* Qb = Qs + tm * 8 * 64 + lane;
  const bf16x8_t* Bb = Bs + tn * 4 * 64 + lane;
  const int64_t qkb = RTQ * 64, qpl = (int64_t)KB * qkb;
  const int64_t bkb = RTB * 64, bpl = (int64_t)KB * bkb;
  // direct loads of k block kb into stage st: chunks taken by the four waves
  // in turn -- 3 per wave (plane 0 only) below kblo, 9 above
  auto issue = [&](int kb, int st) {
    const int np = (kb < kblo) ? 1 : 3;
    if (np == 1) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int ch = wave + 4 * q, t = ch;  // 0..11
        const bf16x8_t* src = (t < 8) ? Qb + kb * qkb + t * 64 : Bb + kb * bkb + (t - 8) * 64;
        lds_ptr_t d = (ch < 36) ? (lds_ptr_t)&Ls[st][ch][0] : (lds_ptr_t)&Ldummy[0];
        __builtin_amdgcn_global_load_lds((gbl_ptr_t)src, d, 16, 0, 0);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const int ch = wave + 4 * q, P = ch / 12, t = ch % 12;  // 0..35
        const bf16x8_t* src = (t < 8) ? Qb + P * qpl + kb * qkb + t * 64
                                      : Bb + P * bpl + kb * bkb + (t - 8) * 64;
        lds_ptr_t d = (ch < 36) ? (lds_ptr_t)&Ls[st][ch][0] : (lds_ptr_t)&Ldummy[0];
        __builtin_amdgcn_global_load_lds((gbl_ptr_t)src, d, 16, 0, 0);
      }
    }
  };

  f4v_t acc[2][4];  // [B row tile a][Q row tile b]
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4v_t{0.f, 0.f, 0.f, 0.f};

  // one k block on stage st (a compile-time index, so that the fragment
  // reads have constant offsets)
  auto step = [&](int kb, auto stc) {
    constexpr int st = decltype(stc)::value;
    // this wave's loads of block kb have landed and its reads of the other
    // stage (block kb - 1) have returned; after the barrier block kb is
    // complete in LDS and the other stage is free for block kb + 1
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4));
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (kb + 1 < KB) issue(kb + 1, st ^ 1);
    if (kb < kblo) {
      bf16x8_t fa[4], fb[2];
#pragma unroll
      for (int b = 0; b < 4; ++b) fa[b] = Ls[st][wa + b][lane];
#pragma unroll
      for (int a = 0; a < 2; ++a) fb[a] = Ls[st][8 + wb + a][lane];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[a], fa[b], acc[a][b], 0, 0, 0);
    } else {
      bf16x8_t fa[3][4], fb[3][2];
#pragma unroll
      for (int P = 0; P < 3; ++P) {
#pragma unroll
        for (int b = 0; b < 4; ++b) fa[P][b] = Ls[st][12 * P + wa + b][lane];
#pragma unroll
        for (int a = 0; a < 2; ++a) fb[P][a] = Ls[st][12 * P + 8 + wb + a][lane];
      }
      // per accumulator the tiled kernel's order (smallest terms first);
      // consecutive MFMAs go to different accumulators
      constexpr int PB[6] = {0, 2, 1, 0, 1, 0}, PA[6] = {2, 0, 1, 1, 0, 0};
#pragma unroll
      for (int term = 0; term < 6; ++term)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[PB[term]][a], fa[PA[term]][b],
                                                                acc[a][b], 0, 0, 0);
    }
  };
  // this lane's 32 C values, loaded before the first direct loads so that
  // their latency overlaps stage 0's instead of following the k loop (no
  // other workgroup touches this tile, so they are the values read later)
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 32;
  float cpre[2][4][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t gm = m0 + wm0 + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gn = n0 + wn0 + a * 16 + 4 * lk + r;
        cpre[a][b][r] = (gm < M && gn < N) ? C[gm + gn * ldc] : 0.0f;
      }
    }
  issue(0, 0);
  for (int kb = 0; kb < KB; kb += 2) {
    step(kb, std::integral_constant<int, 0>{});
    if (kb + 1 < KB) step(kb + 1, std::integral_constant<int, 1>{});
  }

  // C <- C - acc and the tile's FP64 sum of squares
  double ss = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t gm = m0 + wm0 + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gn = n0 + wn0 + a * 16 + 4 * lk + r;
        if (gm < M && gn < N) {
          float v = -1.0f * acc[a][b][r];
          v += 1.0f * cpre[a][b][r];
          C[gm + gn * ldc] = v;
          ss += (double)v * (double)v;
        }
      }
    }
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

int gemm_bf16_deflate(Context* c, int64_t M, int64_t N, int64_t K, const float* Q, int64_t ldq,
                      const float* Bt, int64_t ldbt, float* C, int64_t ldc, double* sumsq,
                      int64_t klo) {
  if (M < 0 || N < 0 || K < 0) return set_error("deflate: negative dimension");
  if (M == 0 || N == 0) {
    RSVD_HIP(hipMemsetAsync(sumsq, 0, sizeof(double), c->stream));
    return 0;
  }
  if (K == 0) return set_error("deflate: empty block");
  if (klo % 32 != 0) return set_error("deflate: klo must be a multiple of 32");
  const int64_t gm = (M + LBM - 1) / LBM, gnr = (N + kDBN - 1) / kDBN;
  const int n_fast = (gnr <= gm) ? 1 : 0;
  const int64_t gx = n_fast ? gnr : gm, gy = n_fast ? gm : gnr;
  DBuf<double> part;
  int64_t tiles;
  if (K % 32 == 0 && gy <= 65535) {
    // the LDS ring on operands split once
    const int64_t KB = K / 32, kblo = klo / 32;
    const int64_t RTQ = gm * 8, RTB = gnr * 4;
    tiles = gm * gnr;
    RSVD_TRY(part.alloc((size_t)tiles));
    DBuf<bf16x8_t> qs, bs;
    RSVD_TRY(qs.alloc((size_t)(3 * KB * RTQ * 64)));
    RSVD_TRY(bs.alloc((size_t)(3 * KB * RTB * 64)));
    KTimer kt(c, "deflate_bf16x6");
    auto blocks_for = [](int64_t items) {
      const int64_t b = (items + 255) / 256;
      return (unsigned)(b < 8192 ? b : 8192);
    };
    hipLaunchKernelGGL(split3_frag_kernel, dim3(blocks_for(KB * RTQ * 64)), dim3(256), 0,
                       c->stream, M, KB, RTQ, kblo, Q, ldq, qs.p);
    RSVD_LAUNCH_CHECK();
    hipLaunchKernelGGL(split3_frag_kernel, dim3(blocks_for(KB * RTB * 64)), dim3(256), 0,
                       c->stream, N, KB, RTB, kblo, Bt, ldbt, bs.p);
    RSVD_LAUNCH_CHECK();
    hipLaunchKernelGGL(deflate_ring_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(LTHREADS), 0,
                       c->stream, M, N, (int)KB, (int)kblo, (const bf16x8_t*)qs.p, RTQ,
                       (const bf16x8_t*)bs.p, RTB, C, ldc, n_fast, part.p);
    RSVD_LAUNCH_CHECK();
  } else {
    tiles = gm * ((N + LBN - 1) / LBN);
    RSVD_TRY(part.alloc((size_t)tiles));
    RSVD_TRY(gemm_bf16_tiled(c, "deflate_bf16x6", 'N', 'T', M, N, K, -1.0f, Q, ldq, Bt, ldbt,
                             1.0f, C, ldc, 6, part.p, klo));
  }
  // the per-tile partials: 256-way in slices of `chunk`, then the slice sums
  const int64_t chunk = (tiles + 255) / 256;
  const int64_t nsl = (tiles + chunk - 1) / chunk;
  DBuf<double> slice;
  RSVD_TRY(slice.alloc((size_t)nsl));
  hipLaunchKernelGGL(sum_partials_lp_kernel, dim3((unsigned)nsl), dim3(256), 0, c->stream, tiles,
                     chunk, (const double*)part.p, slice.p);
  RSVD_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_partials_lp_kernel, dim3(1), dim3(256), 0, c->stream, nsl, nsl,
                     (const double*)slice.p, sumsq);
  RSVD_LAUNCH_CHECK();
  return 0;
}

}  // namespace rsvd
