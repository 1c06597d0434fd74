// apass.hip -- the A-streaming passes of the mixed-precision rSVD on the bf16
// matrix cores: C = alpha * op(A) * S + beta * C with A the large fp32 matrix
// (m x n, column-major, never copied) and S a skinny fp32 panel (K x N, N up to
// a few hundred): Y = A Omega and Y = A Z (op = N), Z = A^T Q and B^T = A^T Q
// (op = T) -- the dominant cost of rra.c:96-131 in the reference.
//
// Design for a skinny N: one workgroup owns 64 rows of op(A) and ALL the
// columns of C (up to 576 = 4 waves x 9 tiles of 16), so A is read from HBM
// exactly once per pass; only the skinny operand is re-read, and it is shared
// by every workgroup at the same k, so it comes from L2.
//   * S is first rounded to bf16 (and, for split-bf16, its residual lo =
//     bf16(s - hi)) and re-laid out in MFMA fragment order: for every k block
//     of 32 and n tile of 16 a contiguous 1 KiB block whose 64 lanes each hold
//     8 consecutive k of one column.  A wave then loads a B fragment with one
//     coalesced 16-byte load per lane, prefetched one k block ahead.
//   * A's 64 x 32 tile of a k block is loaded into registers three blocks
//     ahead (16-byte loads when op(A) is k-contiguous, 4-byte loads along the
//     contiguous m otherwise -- both coalesced), rounded to bf16 (hi, lo) and
//     written to LDS in fragment order, three stages, one barrier per block;
//     each wave reads its A fragments one block ahead.
//   * v_mfma_f32_16x16x32_bf16 with srcA = the S fragment and srcB = the A
//     fragment, so the accumulator lanes run along m and C's column-major
//     stores are coalesced.  Split-bf16 (SPLIT = 3) adds hi*lo and lo*hi.
// One workgroup of 4 waves per CU (launch bounds 256, 1) leaves each wave
// the full register file: 144 accumulators, double-buffered S and A
// fragments (the k loop is unrolled by two so the buffers swap by name).
#include <hip/hip_bf16.h>

#include <cmath>

#include "rsvd_internal.h"

namespace rsvd {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f4v_t __attribute__((ext_vector_type(4)));
typedef float f4s_t __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kBM = 64;               // rows of op(A) per workgroup
constexpr int kNTW = 9;               // n tiles (16 columns) per wave
constexpr int kBNT = kWaves * kNTW;   // n tiles per workgroup (576 columns)

// St[((p * KB + kb) * NTP + nt) * 64 + lane][e] = plane p of
// S(32 kb + 8 (lane / 16) + e, 16 nt + lane % 16), zero outside K x N.
__global__ void tile_skinny_kernel(int64_t K, int64_t N, const float* __restrict__ S, int64_t lds,
                                   int64_t KB, int64_t NTP, int planes,
                                   bf16x8_t* __restrict__ St) {
  const int64_t total = KB * NTP * 64;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(idx & 63);
    const int64_t t = idx >> 6;
    const int64_t nt = t % NTP, kb = t / NTP;
    const int64_t n = nt * 16 + (lane & 15);
    const int64_t k0 = kb * 32 + 8 * (lane >> 4);
    bf16x8_t h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t k = k0 + e;
      const float v = (n < N && k < K) ? S[k + n * lds] : 0.0f;
      const __bf16 hb = (__bf16)v;
      h[e] = hb;
      l[e] = (__bf16)(v - (float)hb);
    }
    St[t * 64 + lane] = h;
    if (planes == 2) St[(KB * NTP + t) * 64 + lane] = l;
  }
}

// K must be a multiple of 32 here (the caller adds the remainder with a
// second product).  Rows past M are clamped to row M - 1 for the loads (their
// results are never stored) and the S tiles past N are zero, so the k loop
// has no per-lane predicates and no branches: the memory-counter waits the
// compiler inserts can then leave the prefetches of the next k block in
// flight across the MFMAs of the current one.
template <bool TA, int SPLIT, bool VEC>
__global__ void __launch_bounds__(kThreads, 1)
apass_kernel(int64_t M, int64_t N, int64_t K, float alpha, const float* __restrict__ A,
             int64_t lda, const bf16x8_t* __restrict__ St, int64_t KB, int64_t NTP, float beta,
             float* __restrict__ C, int64_t ldc) {
  constexpr int NP = (SPLIT == 3) ? 2 : 1;
  constexpr int MT = kBM / 16;
  // three A stages: block kb + 2 is written while block kb + 1 is read into
  // registers and block kb is consumed
  __shared__ bf16x8_t As[3][NP][MT][64];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t ntb = (int64_t)blockIdx.y * kBNT + wave * kNTW;  // this wave's first n tile

  // A staging: thread -> row r of the tile, k group g (8 consecutive k)
  const int r = TA ? (tid >> 2) : (tid & 63);
  const int g = TA ? (tid & 3) : (tid >> 6);
  const int64_t gm = (m0 + r < M) ? m0 + r : M - 1;
  const float* Arow = TA ? A + gm * lda + 8 * g : A + gm + (int64_t)(8 * g) * lda;
  const int64_t last = KB - 1;
  float av[8];
  auto loadA = [&](int64_t kb) {
    kb = kb < last ? kb : last;
    if (TA) {
      const float* p = Arow + kb * 32;
      if (VEC) {
        const f4s_t x = *reinterpret_cast<const f4s_t*>(p);
        const f4s_t y = *reinterpret_cast<const f4s_t*>(p + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          av[e] = x[e];
          av[4 + e] = y[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) av[e] = p[e];
      }
    } else {
      const float* p = Arow + kb * 32 * lda;
#pragma unroll
      for (int e = 0; e < 8; ++e) av[e] = p[e * lda];
    }
  };
  auto storeA = [&](int64_t kb) {
    const int st = (int)(kb % 3);
    bf16x8_t h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const __bf16 hb = (__bf16)av[e];
      h[e] = hb;
      if (SPLIT == 3) l[e] = (__bf16)(av[e] - (float)hb);
    }
    const int mt = r >> 4, ln = (r & 15) + 16 * g;
    As[st][0][mt][ln] = h;
    if (SPLIT == 3) As[st][NP - 1][mt][ln] = l;
  };
  auto readA = [&](int64_t kb, bf16x8_t* h, bf16x8_t* l) {
    const int st = (int)(kb % 3);
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      h[i] = As[st][0][i][lane];
      if (SPLIT == 3) l[i] = As[st][NP - 1][i][lane];
    }
  };
  const bf16x8_t* Sh = St + ntb * 64 + lane;
  const bf16x8_t* Sl = St + (KB * NTP + ntb) * 64 + lane;
  auto loadS = [&](int64_t kb, bf16x8_t* h, bf16x8_t* l) {
    kb = kb < last ? kb : last;
#pragma unroll
    for (int j = 0; j < kNTW; ++j) {
      h[j] = Sh[(kb * NTP + j) * 64];
      if (SPLIT == 3) l[j] = Sl[(kb * NTP + j) * 64];
    }
  };

  f4v_t acc[kNTW][MT];
#pragma unroll
  for (int j = 0; j < kNTW; ++j)
#pragma unroll
    for (int i = 0; i < MT; ++i) acc[j][i] = f4v_t{0.f, 0.f, 0.f, 0.f};

  // one k block: prefetch S and the A fragments of block kb + 1 (issued
  // first and kept ahead of the MFMAs by a scheduling barrier), MFMAs on
  // block kb, stage A block kb + 2, fetch A block kb + 3
  auto step = [&](int64_t kb, bf16x8_t* sch, bf16x8_t* scl, bf16x8_t* ach, bf16x8_t* acl,
                  bf16x8_t* snh, bf16x8_t* snl, bf16x8_t* anh, bf16x8_t* anl) {
    loadS(kb + 1, snh, snl);
    readA(kb + 1, anh, anl);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < kNTW; ++j)
#pragma unroll
      for (int i = 0; i < MT; ++i)
        acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sch[j], ach[i], acc[j][i], 0, 0, 0);
    if (SPLIT == 3) {
#pragma unroll
      for (int j = 0; j < kNTW; ++j)
#pragma unroll
        for (int i = 0; i < MT; ++i)
          acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sch[j], acl[i], acc[j][i], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < kNTW; ++j)
#pragma unroll
        for (int i = 0; i < MT; ++i)
          acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(scl[j], ach[i], acc[j][i], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    storeA(kb + 2);
    loadA(kb + 3);
    __syncthreads();
  };

  bf16x8_t s0h[kNTW], s0l[kNTW], s1h[kNTW], s1l[kNTW];
  bf16x8_t a0h[MT], a0l[MT], a1h[MT], a1l[MT];
  loadA(0);
  storeA(0);
  loadA(1);
  storeA(1);
  loadA(2);
  __syncthreads();
  readA(0, a0h, a0l);
  loadS(0, s0h, s0l);
  int64_t kb = 0;
  for (; kb + 1 < KB; kb += 2) {
    step(kb, s0h, s0l, a0h, a0l, s1h, s1l, a1h, a1l);
    step(kb + 1, s1h, s1l, a1h, a1l, s0h, s0l, a0h, a0l);
  }
  if (kb < KB) step(kb, s0h, s0l, a0h, a0l, s1h, s1l, a1h, a1l);

  // D[n][m] of tile (j, i): lane holds n = 4 (lane / 16) + rr, m = lane % 16
#pragma unroll
  for (int j = 0; j < kNTW; ++j) {
#pragma unroll
    for (int i = 0; i < kBM / 16; ++i) {
      const int64_t m = m0 + i * 16 + (lane & 15);
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int64_t n = (ntb + j) * 16 + 4 * (lane >> 4) + rr;
        if (m < M && n < N) {
          float* cp = C + m + n * ldc;
          float v = alpha * acc[j][i][rr];
          if (beta != 0.0f) v += beta * (*cp);
          *cp = v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- apass2
// Eight waves (two per SIMD) and a 256 x 288 tile per workgroup: waves
// (wm, wn) = (wave % 4, wave / 4) own 64 rows x 144 columns (4 x 9 tiles of
// 16 x 16, 144 accumulators; two waves per SIMD leave 256 registers each).
// Per k block of 32 the workgroup needs 288 columns of the skinny operand (hi
// and lo planes, 36 KiB) and 256 rows of A (32 KiB): 2.5x fewer bytes per
// output from L2 than the 64 x 576 tile.  At m = 65536, l <= 576 the grid is
// 512 workgroups, two full waves of one workgroup per CU (a 192-row tile gave
// 683: 2.67 waves, the last one a third idle, and measured 8% slower).  For
// op(A) = A^T the tile is 128 x 288 (two row tiles per wave).
//   * The skinny operand goes HBM/L2 -> LDS with direct loads
//     (global_load_lds_dwordx4, no registers) into a ring of three stages,
//     issued two k blocks ahead; each wave reads its fragments from LDS one
//     column tile ahead of the MFMAs that use them.
//   * A goes straight into registers in fragment order (lane l: row l % 16,
//     k 8 (l / 16) .. + 7), one k block ahead, and is split into bf16 hi / lo
//     at the end of the block.  The two waves that share rows (wn = 0, 1)
//     load the same lines at the same time, so the second read hits L1.
//   * One barrier per k block: after it, the next stage is complete for every
//     wave and the stage read in this block may be overwritten.
// The k range can be split over gridDim.z-like slices (ks): slice partials go
// to a workspace and a second kernel adds them.  Block ids are mapped so that
// the column blocks of one row block and k slice run on the same XCD next to
// each other (they read the same A lines).
typedef __attribute__((address_space(3))) void* lds_ptr_t;
typedef __attribute__((address_space(1))) void* gbl_ptr_t;

constexpr int kW2 = 8;
constexpr int kT2 = 64 * kW2;
constexpr int kMT2 = 4;               // row tiles per wave (op(A) = A)
constexpr int kBM2 = 4 * 16 * kMT2;   // rows per workgroup
constexpr int kMT2T = 2;              // row tiles per wave (op(A) = A^T)
constexpr int kNTW2 = 9;
constexpr int kBNT2 = 2 * kNTW2;
constexpr int kST2 = 3;
// op(A) = A^T with 288 < l <= 544: one 128 x 544 tile per workgroup (17
// column tiles a wave, two LDS stages, 140 KiB), so that every A tile is
// loaded once instead of by two column blocks; the k range is split to fill
// the CUs.  At m = 65536, n = 16384, l = 544: 4.02 -> 3.19 ms a pass with the
// slice reduction (the same tile with three stages does not fit the LDS).
constexpr int kNTWW = 17;
constexpr int kBNTW = 2 * kNTWW;

// v_mfma_f32_16x16x32 on 8 16-bit values per lane: bf16 (SPLIT 1, 3) or fp16
// (SPLIT 2), the fp16 fragments carried in the bf16 vector type bit for bit.
template <int SPLIT>
__device__ __forceinline__ f4v_t mfma16(bf16x8_t a, bf16x8_t b, f4v_t c) {
  if constexpr (SPLIT == 2)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a),
                                                  __builtin_bit_cast(f16x8_t, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// SPLIT 1: bf16 operands, one product.  SPLIT 3: split-bf16, hi hi + hi lo +
// lo hi.  SPLIT 2: fp16, A scaled by 1 / a_c per 128-column chunk of A
// (ascale) and rounded to fp16, the skinny operand in two fp16 planes (hi +
// lo, scaled per column by 1 / b_n, with a_c folded in for op(A) = A), two
// products hi hi + lo hi; the epilogue multiplies by b_n (bscale) and, for
// op(A) = A^T, by the output row's a_c.
template <bool TA, int SPLIT, bool VEC, int MT, int NTW = kNTW2, int ST = kST2>
__global__ void __launch_bounds__(kT2)
apass2_kernel(int64_t M, int64_t N, int64_t KBtot, int64_t KBper, float alpha,
              const float* __restrict__ A, int64_t lda, const bf16x8_t* __restrict__ St,
              int64_t NTP, float beta, float* __restrict__ C, int64_t ldc,
              float* __restrict__ Wk, int gx, int gy, int ks, const float* __restrict__ ascale,
              const float* __restrict__ bscale) {
  static_assert(ST == 2 || ST == 3, "a ring of two or three stages");
  constexpr int BNT = 2 * NTW;  // column tiles per workgroup
  constexpr int NP = (SPLIT >= 2) ? 2 : 1;
  constexpr int NCH = NP * BNT;             // 1 KiB chunks per stage
  constexpr int NLD = (NCH + kW2 - 1) / kW2;  // chunk loads per wave per stage
  __shared__ bf16x8_t Ss[ST][NP][BNT][64];
  __shared__ bf16x8_t Sdummy[64];

  const int tid = threadIdx.x, lane = tid & 63;
  // wave-uniform values in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 3, wn = wave >> 2;
  const int id = blockIdx.x;
  int bx, by, kz;
  if ((gx & 7) == 0) {
    const int xcd = id & 7, slot = id >> 3;
    by = slot % gy;
    const int t = slot / gy;
    kz = t % ks;
    bx = (t / ks) * 8 + xcd;
  } else {
    bx = id % gx;
    const int t = id / gx;
    by = t % gy;
    kz = t / gy;
  }
  bx = __builtin_amdgcn_readfirstlane(bx);
  by = __builtin_amdgcn_readfirstlane(by);
  kz = __builtin_amdgcn_readfirstlane(kz);
  if ((int64_t)bx * (64 * MT) >= M) return;  // padding of the row-block count to 8
  const int64_t kb0 = (int64_t)kz * KBper;
  const int64_t kbe = (kb0 + KBper < KBtot) ? kb0 + KBper : KBtot;
  const int64_t nk = kbe - kb0;
  const int64_t klast = kbe - 1;
  const int64_t mw = (int64_t)bx * (64 * MT) + wm * 16 * MT;  // this wave's first row
  const int64_t ntw = (int64_t)by * BNT;          // workgroup's first n tile

  // ---- S stage loads: chunk c = (plane c / 18, tile c % 18), wave-strided;
  // the padding slots reload chunk 0 into Sdummy so every wave issues NLD
  // per-wave chunk offsets, fixed over the k loop (wave-uniform): source
  // element offset at k block 0 and the chunk's slot within an LDS stage
  // (padding slots go to Sdummy).  Used for op(A) = A^T, where it removes
  // most of the scalar work per k block; the op(A) = A variant, at the
  // register limit, spills with it and recomputes them instead.
  int64_t sbase[NLD];
  int soff[NLD];
  bool sreal[NLD];
#pragma unroll
  for (int q = 0; q < NLD; ++q) {
    const int c = wave + kW2 * q;
    sreal[q] = c < NCH;
    const int pl = sreal[q] ? c / BNT : 0, t = sreal[q] ? c % BNT : 0;
    sbase[q] = ((int64_t)pl * KBtot * NTP + ntw + t) * 64;
    soff[q] = pl * BNT + t;
  }
  const int64_t kstride = NTP * 64;  // St elements per k block
  auto loadS = [&](int64_t kb, bf16x8_t (*dst)[BNT][64]) {
    kb = kb < klast ? kb : klast;
    if constexpr (TA) {
      const int64_t koff = kb * kstride;
#pragma unroll
      for (int q = 0; q < NLD; ++q) {
        lds_ptr_t d = sreal[q] ? (lds_ptr_t)&dst[0][soff[q]][0] : (lds_ptr_t)&Sdummy[0];
        __builtin_amdgcn_global_load_lds((gbl_ptr_t)(St + (sbase[q] + koff) + lane), d, 16, 0,
                                         0);
      }
    } else {
#pragma unroll
      for (int q = 0; q < NLD; ++q) {
        const int c = wave + kW2 * q;
        const bool real = c < NCH;
        const int pl = real ? c / BNT : 0, t = real ? c % BNT : 0;
        const bf16x8_t* g = St + ((pl * KBtot + kb) * NTP + ntw + t) * 64 + lane;
        lds_ptr_t d = real ? (lds_ptr_t)&dst[pl][t][0] : (lds_ptr_t)&Sdummy[0];
        __builtin_amdgcn_global_load_lds((gbl_ptr_t)g, d, 16, 0, 0);
      }
    }
  };

  // ---- A: lane -> row mw + 16 i + lane % 16, k 8 (lane / 16) .. + 7
  const float* Ar[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) {
    int64_t r = mw + 16 * i + (lane & 15);
    r = r < M ? r : M - 1;
    Ar[i] = TA ? A + r * lda + 8 * (lane >> 4) : A + r + (int64_t)(8 * (lane >> 4)) * lda;
  }
  float raw[MT][8];
  int64_t kb_raw = kb0;  // k block held in raw
  auto loadA = [&](int64_t kb) {
    kb = kb < klast ? kb : klast;
    kb_raw = kb;
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      if (TA) {
        const float* p = Ar[i] + kb * 32;
        if (VEC) {
          const f4s_t x = *reinterpret_cast<const f4s_t*>(p);
          const f4s_t y = *reinterpret_cast<const f4s_t*>(p + 4);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            raw[i][e] = x[e];
            raw[i][4 + e] = y[e];
          }
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) raw[i][e] = p[e];
        }
      } else {
        const float* p = Ar[i] + kb * 32 * lda;
#pragma unroll
        for (int e = 0; e < 8; ++e) raw[i][e] = p[e * lda];
      }
    }
  };
  // one row tile of A for k block kb (op(A) = A: eight 4-byte loads along k)
  auto loadA_row = [&](int64_t kb, int i) {
    kb = kb < klast ? kb : klast;
    kb_raw = kb;
    const float* p = Ar[i] + kb * 32 * lda;
#pragma unroll
    for (int e = 0; e < 8; ++e) raw[i][e] = p[e * lda];
  };
  bf16x8_t ah[MT], al[MT];
  // fp16: 1 / a_c of the chunk of A's columns the values in raw come from --
  // the k block's chunk (op(A) = A) or the workgroup's 128 rows (op(A) = A^T)
  static_assert(SPLIT != 2 || !TA || 64 * MT == 128, "fp16: one chunk of A per row block");
  const float inv_row = (SPLIT == 2 && TA) ? 1.0f / ascale[bx] : 1.0f;
  auto split = [&]() {
    if constexpr (SPLIT == 2) {
      const float sc = TA ? inv_row : 1.0f / ascale[kb_raw >> 2];
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int e = 0; e < 8; ++e)
          ah[i][e] = __builtin_bit_cast(__bf16, (_Float16)(raw[i][e] * sc));
    } else {
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const __bf16 hb = (__bf16)raw[i][e];
          ah[i][e] = hb;
          if (SPLIT == 3) al[i][e] = (__bf16)(raw[i][e] - (float)hb);
        }
    }
  };

  f4v_t acc[NTW][MT];
#pragma unroll
  for (int j = 0; j < NTW; ++j)
#pragma unroll
    for (int i = 0; i < MT; ++i) acc[j][i] = f4v_t{0.f, 0.f, 0.f, 0.f};

  // one k block on stage `cur`: the MFMAs with the S fragments read one tile
  // ahead, the A loads of kb + 1 and the A split for kb + 1, and the S(kb + 2)
  // loads.  op(A) = A issues A(kb + 1) one row tile per column pair, spread
  // over the MFMAs (its 32 four-byte loads per block, issued together, kept
  // the load path busy in bursts), and S(kb + 2) after the split: the split's
  // wait for A then covers no S load that is not due anyway, so S keeps a full
  // block of lead (measured standalone at m = 65536, n = 16384, l = 544:
  // 3.55 -> 3.22 ms, bitwise the same output).  op(A) = A^T, whose A loads
  // are 16-byte vectors, measured 3% slower that way and keeps both at the
  // start of the block.
  auto step = [&](int64_t kb, bf16x8_t (*cur)[BNT][64], bf16x8_t (*nx2)[BNT][64]) {
    if constexpr (ST == 2) {
      // two stages: S(kb + 1) into the stage read in the previous block,
      // issued before the A loads so that the split's wait for A(kb + 1)
      // (vector memory counters retire in order) also covers it
      loadS(kb + 1, nx2);
      if constexpr (TA) loadA(kb + 1);
    } else if constexpr (TA) {
      loadA(kb + 1);
      loadS(kb + 2, nx2);
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (!TA) {
      // column tiles in pairs, so that consecutive MFMAs on one accumulator are
      // 2 MT apart (measured 8% faster for op(A) = A; slower for op(A) = A^T)
      constexpr int NPR = (NTW + 1) / 2;
      bf16x8_t sh[2][2], sl[2][2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        sh[0][q] = cur[0][wn * NTW + q][lane];
        if (SPLIT >= 2) sl[0][q] = cur[NP - 1][wn * NTW + q][lane];
      }
      static_assert(NPR >= MT, "one row tile of A per column pair");
#pragma unroll
      for (int jp = 0; jp < NPR; ++jp) {
        const int b = jp & 1;
        if (jp < MT) loadA_row(kb + 1, jp);
        if (jp + 1 < NPR) {
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int jn = 2 * (jp + 1) + q;
            if (jn < NTW) {
              sh[b ^ 1][q] = cur[0][wn * NTW + jn][lane];
              if (SPLIT >= 2) sl[b ^ 1][q] = cur[NP - 1][wn * NTW + jn][lane];
            }
          }
        }
        const int j0 = 2 * jp, j1 = 2 * jp + 1;
#pragma unroll
        for (int i = 0; i < MT; ++i)
          acc[j0][i] = mfma16<SPLIT>(sh[b][0], ah[i], acc[j0][i]);
        if (j1 < NTW) {
#pragma unroll
          for (int i = 0; i < MT; ++i)
            acc[j1][i] = mfma16<SPLIT>(sh[b][1], ah[i], acc[j1][i]);
        }
        if (SPLIT == 3) {
#pragma unroll
          for (int i = 0; i < MT; ++i)
            acc[j0][i] = mfma16<SPLIT>(sh[b][0], al[i], acc[j0][i]);
          if (j1 < NTW) {
#pragma unroll
            for (int i = 0; i < MT; ++i)
              acc[j1][i] = mfma16<SPLIT>(sh[b][1], al[i], acc[j1][i]);
          }
        }
        if (SPLIT >= 2) {
#pragma unroll
          for (int i = 0; i < MT; ++i)
            acc[j0][i] = mfma16<SPLIT>(sl[b][0], ah[i], acc[j0][i]);
          if (j1 < NTW) {
#pragma unroll
            for (int i = 0; i < MT; ++i)
              acc[j1][i] = mfma16<SPLIT>(sl[b][1], ah[i], acc[j1][i]);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      bf16x8_t sh[2], sl[2];
      sh[0] = cur[0][wn * NTW][lane];
      if (SPLIT >= 2) sl[0] = cur[NP - 1][wn * NTW][lane];
#pragma unroll
      for (int j = 0; j < NTW; ++j) {
        const int b = j & 1;
        if (j + 1 < NTW) {
          sh[b ^ 1] = cur[0][wn * NTW + j + 1][lane];
          if (SPLIT >= 2) sl[b ^ 1] = cur[NP - 1][wn * NTW + j + 1][lane];
        }
#pragma unroll
        for (int i = 0; i < MT; ++i)
          acc[j][i] = mfma16<SPLIT>(sh[b], ah[i], acc[j][i]);
        if (SPLIT == 3) {
#pragma unroll
          for (int i = 0; i < MT; ++i)
            acc[j][i] = mfma16<SPLIT>(sh[b], al[i], acc[j][i]);
        }
        if (SPLIT >= 2) {
#pragma unroll
          for (int i = 0; i < MT; ++i)
            acc[j][i] = mfma16<SPLIT>(sl[b], ah[i], acc[j][i]);
        }
        // keep the fragment reads one tile ahead (not all hoisted: registers)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    split();
    if constexpr (!TA && ST == 3) loadS(kb + 2, nx2);
    __syncthreads();
  };

  loadS(kb0, Ss[0]);
  if constexpr (ST == 3) loadS(kb0 + 1, Ss[1]);
  loadA(kb0);
  split();
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int st = 0;
  for (int64_t s = 0; s < nk; ++s) {
    if constexpr (ST == 2) {
      step(kb0 + s, Ss[st], Ss[st ^ 1]);
      st ^= 1;
    } else {
      const int st2 = (st >= 1) ? st - 1 : st + 2;  // (s + 2) % 3
      step(kb0 + s, Ss[st], Ss[st2]);
      st = (st == 2) ? 0 : st + 1;
    }
  }
  // drain the trailing direct loads before the workgroup exits
  __builtin_amdgcn_s_waitcnt(0);

  // D[n][m] of tile (j, i): lane holds n = 4 (lane / 16) + rr, m = lane % 16
  float* out = (ks > 1) ? Wk + (int64_t)kz * M * (NTP * 16) : C;
  const int64_t ldo = (ks > 1) ? M : ldc;
  const float rowsc = (SPLIT == 2 && TA) ? ascale[bx] : 1.0f;
#pragma unroll
  for (int j = 0; j < NTW; ++j) {
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const int64_t m = mw + i * 16 + (lane & 15);
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int64_t n = (ntw + wn * NTW + j) * 16 + 4 * (lane >> 4) + rr;
        if (m < M && n < N) {
          float* cp = out + m + n * ldo;
          const float a = (SPLIT == 2) ? acc[j][i][rr] * (bscale[n] * rowsc) : acc[j][i][rr];
          if (ks > 1) {
            *cp = a;
          } else {
            float v = alpha * a;
            if (beta != 0.0f) v += beta * (*cp);
            *cp = v;
          }
        }
      }
    }
  }
}

// C = alpha * sum_z Wk[z] + beta * C over the M x N output.
__global__ void apass2_reduce_kernel(int64_t M, int64_t N, int ks, const float* __restrict__ Wk,
                                     int64_t ldw, float alpha, float beta, float* __restrict__ C,
                                     int64_t ldc) {
  const int64_t total = M * N;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = idx % M, n = idx / M;
    float v = 0.0f;
    for (int z = 0; z < ks; ++z) v += Wk[(int64_t)z * M * ldw + m + n * M];
    v *= alpha;
    float* cp = C + m + n * ldc;
    if (beta != 0.0f) v += beta * (*cp);
    *cp = v;
  }
}

// ---------------------------------------------------------------- fp8
// FP8 (e4m3) A passes with per-128-column scales.  a_c = max |A| over the
// 128-column chunk c of A, / 448 (the largest e4m3 value):
//   * op(A) = A (Y = A S, the chunk runs along K): A is stored as A / a_c and
//     a_c is folded into the skinny operand, S'[k, :] = a_{k/128} S[k, :];
//   * op(A) = A^T (Z = A^T S, the chunk runs along the output rows): A / a_c,
//     and output row j is multiplied by a_{j/128} in the epilogue.
// The skinny operand is quantised with one scale per output column, b_n =
// max_k |S'[k, n]| / 448, applied in the epilogue.  Products on
// v_mfma_f32_16x16x32_fp8_fp8 (fragments: 8 e4m3 values per lane along k,
// the layout of the bf16 16x16x32 instruction), fp32 accumulation.
typedef long fp8x8_t;

__device__ __forceinline__ fp8x8_t pack_fp8x8(const float (&v)[8], float scale) {
  int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * scale, v[1] * scale, 0, false);
  lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * scale, v[3] * scale, lo, true);
  int hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4] * scale, v[5] * scale, 0, false);
  hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6] * scale, v[7] * scale, hi, true);
  return (fp8x8_t)(((unsigned long)(unsigned)hi << 32) | (unsigned)lo);
}

// scale_bits[c] = bits of max |A| over columns 128 c .. 128 c + 127 of A
// (M x N, ld lda); non-negative floats order like their bit patterns.
// max |A| over the 128-column chunk blockIdx.x and the row slice blockIdx.y
// into scale_bits (float bits, atomicMax: the max of non-negative floats is
// the max of their bit patterns, and max is order-independent).  16-byte
// loads along the columns when A allows them, four columns in flight.
__global__ void chunk_absmax_kernel(int64_t M, int64_t N, const float* __restrict__ A,
                                    int64_t lda, unsigned* __restrict__ scale_bits) {
  const int64_t c = blockIdx.x;
  const int64_t j0 = c * 128, j1 = (j0 + 128 < N) ? j0 + 128 : N;
  const int64_t rs = (((M + gridDim.y - 1) / gridDim.y) + 3) & ~int64_t(3);
  const int64_t r0 = (int64_t)blockIdx.y * rs, r1 = (r0 + rs < M) ? r0 + rs : M;
  float mx = 0.0f;
  if (r0 < r1) {
    const bool vec = (lda & 3) == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    const int64_t r4 = vec ? (r1 - r0) >> 2 : 0;  // float4 per column
    for (int64_t j = j0; j < j1; j += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (j + u >= j1) break;
        const float* col = A + (j + u) * lda + r0;
        for (int64_t q = threadIdx.x; q < r4; q += blockDim.x) {
          const float4 v = *reinterpret_cast<const float4*>(col + 4 * q);
          mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        }
        for (int64_t r = r0 + 4 * r4 + threadIdx.x; r < r1; r += blockDim.x)
          mx = fmaxf(mx, fabsf(A[r + (j + u) * lda]));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(scale_bits + c, __float_as_uint(mx));
}

// scale_bits -> a_c = max / top (1 for an all-zero chunk); top = the largest
// value the scaled chunk may take (448 for e4m3, 32768 for fp16)
__global__ void chunk_scale_finish_kernel(int64_t nc, const unsigned* __restrict__ bits,
                                          float top, float* __restrict__ scale) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < nc) {
    const float mx = __uint_as_float(bits[c]);
    scale[c] = (mx > 0.0f) ? mx / top : 1.0f;
  }
}

// b[n] = max_k |S[k, n] a_{k/128}| / top (a optional), one workgroup per column
__global__ void skinny_colscale_kernel(int64_t K, const float* __restrict__ S, int64_t lds,
                                       const float* __restrict__ ascale, float top,
                                       float* __restrict__ b) {
  __shared__ float red[4];
  const int64_t n = blockIdx.x;
  float mx = 0.0f;
  for (int64_t k = threadIdx.x; k < K; k += 256) {
    const float v = S[k + n * lds] * (ascale ? ascale[k >> 7] : 1.0f);
    mx = fmaxf(mx, fabsf(v));
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    b[n] = (mx > 0.0f) ? mx / top : 1.0f;
  }
}

// St8[(kb * NTP + nt) * 64 + lane] = e4m3 of S(32 kb + 8 (lane / 16) + e,
// 16 nt + lane % 16) * a_{k/128} / b_n (a optional), zero outside K x N.
__global__ void tile_skinny8_kernel(int64_t K, int64_t N, const float* __restrict__ S,
                                    int64_t lds, int64_t KB, int64_t NTP,
                                    const float* __restrict__ ascale,
                                    const float* __restrict__ bscale, fp8x8_t* __restrict__ St) {
  const int64_t total = KB * NTP * 64;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(idx & 63);
    const int64_t t = idx >> 6;
    const int64_t nt = t % NTP, kb = t / NTP;
    const int64_t n = nt * 16 + (lane & 15);
    const int64_t k0 = kb * 32 + 8 * (lane >> 4);
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t k = k0 + e;
      v[e] = (n < N && k < K) ? S[k + n * lds] * (ascale ? ascale[k >> 7] : 1.0f) : 0.0f;
    }
    St[t * 64 + lane] = pack_fp8x8(v, (n < N) ? 1.0f / bscale[n] : 0.0f);
  }
}

// Two fp16 planes of S' = S a_{k/128} / b_n (a optional) in the fragment
// order of tile_skinny_kernel: hi = fp16(S'), lo = fp16(S' - hi).
__global__ void tile_skinny16_kernel(int64_t K, int64_t N, const float* __restrict__ S,
                                     int64_t lds, int64_t KB, int64_t NTP,
                                     const float* __restrict__ ascale,
                                     const float* __restrict__ bscale,
                                     bf16x8_t* __restrict__ St) {
  const int64_t total = KB * NTP * 64;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(idx & 63);
    const int64_t t = idx >> 6;
    const int64_t nt = t % NTP, kb = t / NTP;
    const int64_t n = nt * 16 + (lane & 15);
    const int64_t k0 = kb * 32 + 8 * (lane >> 4);
    const float inv_b = (n < N) ? 1.0f / bscale[n] : 0.0f;
    bf16x8_t h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t k = k0 + e;
      const float v =
          (n < N && k < K) ? S[k + n * lds] * (ascale ? ascale[k >> 7] : 1.0f) * inv_b : 0.0f;
      const _Float16 hv = (_Float16)v;
      h[e] = __builtin_bit_cast(__bf16, hv);
      l[e] = __builtin_bit_cast(__bf16, (_Float16)(v - (float)hv));
    }
    St[t * 64 + lane] = h;
    St[(KB * NTP + t) * 64 + lane] = l;
  }
}

// The 64 x 576 A-streaming tile of apass_kernel with e4m3 fragments: A is
// quantised (times 1 / a_c) while staged into LDS, three stages; the skinny
// operand's fragments come from L2, double-buffered in registers.
template <bool TA, bool VEC>
__global__ void __launch_bounds__(kThreads, 1)
apass8_kernel(int64_t M, int64_t N, int64_t K, float alpha, const float* __restrict__ A,
              int64_t lda, const fp8x8_t* __restrict__ St, int64_t KB, int64_t NTP,
              const float* __restrict__ ascale, const float* __restrict__ bscale, float beta,
              float* __restrict__ C, int64_t ldc) {
  constexpr int MT = kBM / 16;
  __shared__ fp8x8_t As[3][MT][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t ntb = (int64_t)blockIdx.y * kBNT + wave * kNTW;
  const int r = TA ? (tid >> 2) : (tid & 63);
  const int g = TA ? (tid & 3) : (tid >> 6);
  const int64_t gm = (m0 + r < M) ? m0 + r : M - 1;
  const float* Arow = TA ? A + gm * lda + 8 * g : A + gm + (int64_t)(8 * g) * lda;
  // op(A) = A^T: this thread's row of op(A) is column gm of A, one scale
  const float rinv_row = TA ? 1.0f / ascale[gm >> 7] : 1.0f;
  const int64_t last = KB - 1;
  float av[8];
  auto loadA = [&](int64_t kb) {
    kb = kb < last ? kb : last;
    if (TA) {
      const float* p = Arow + kb * 32;
      if (VEC) {
        const f4s_t x = *reinterpret_cast<const f4s_t*>(p);
        const f4s_t y = *reinterpret_cast<const f4s_t*>(p + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          av[e] = x[e];
          av[4 + e] = y[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) av[e] = p[e];
      }
    } else {
      const float* p = Arow + kb * 32 * lda;
#pragma unroll
      for (int e = 0; e < 8; ++e) av[e] = p[e * lda];
    }
  };
  auto storeA = [&](int64_t kb) {
    // op(A) = A: the k block lies in one 128-column chunk of A
    const float sc = TA ? rinv_row : 1.0f / ascale[(kb < last ? kb : last) >> 2];
    As[kb % 3][r >> 4][(r & 15) + 16 * g] = pack_fp8x8(av, sc);
  };
  const fp8x8_t* Sp = St + ntb * 64 + lane;
  auto loadS = [&](int64_t kb, fp8x8_t* h) {
    kb = kb < last ? kb : last;
#pragma unroll
    for (int j = 0; j < kNTW; ++j) h[j] = Sp[(kb * NTP + j) * 64];
  };
  f4v_t acc[kNTW][MT];
#pragma unroll
  for (int j = 0; j < kNTW; ++j)
#pragma unroll
    for (int i = 0; i < MT; ++i) acc[j][i] = f4v_t{0.f, 0.f, 0.f, 0.f};
  auto step = [&](int64_t kb, fp8x8_t* sc, fp8x8_t* ac, fp8x8_t* sn, fp8x8_t* an) {
    loadS(kb + 1, sn);
#pragma unroll
    for (int i = 0; i < MT; ++i) an[i] = As[(kb + 1) % 3][i][lane];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < kNTW; ++j)
#pragma unroll
      for (int i = 0; i < MT; ++i)
        acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(sc[j], ac[i], acc[j][i], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    storeA(kb + 2);
    loadA(kb + 3);
    __syncthreads();
  };
  fp8x8_t s0[kNTW], s1[kNTW], a0[MT], a1[MT];
  loadA(0);
  storeA(0);
  loadA(1);
  storeA(1);
  loadA(2);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MT; ++i) a0[i] = As[0][i][lane];
  loadS(0, s0);
  int64_t kb = 0;
  for (; kb + 1 < KB; kb += 2) {
    step(kb, s0, a0, s1, a1);
    step(kb + 1, s1, a1, s0, a0);
  }
  if (kb < KB) step(kb, s0, a0, s1, a1);
#pragma unroll
  for (int j = 0; j < kNTW; ++j) {
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const int64_t m = m0 + i * 16 + (lane & 15);
      const float rs = (TA && m < M) ? ascale[m >> 7] : 1.0f;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int64_t n = (ntb + j) * 16 + 4 * (lane >> 4) + rr;
        if (m < M && n < N) {
          float* cp = C + m + n * ldc;
          float v = alpha * acc[j][i][rr] * bscale[n] * rs;
          if (beta != 0.0f) v += beta * (*cp);
          *cp = v;
        }
      }
    }
  }
}

}  // namespace

bool apass_suitable(Context* c, int64_t M, int64_t N) {
  const int64_t gx = (M + kBM - 1) / kBM;
  const int64_t gy = ((N + 15) / 16 + kBNT - 1) / kBNT;
  const int64_t cus = c->num_cus > 0 ? c->num_cus : 256;
  return gx * gy >= cus && gy <= 65535;
}

int apass_bf16(Context* c, const char* tag, bool ta, int64_t M, int64_t N, int64_t K, float alpha,
               const float* A, int64_t lda, const float* S, int64_t lds, float beta, float* C,
               int64_t ldc, int split) {
  if (M < 0 || N < 0 || K < 0) return set_error("apass: negative dimension");
  if (split != 1 && split != 3) return set_error("apass: split must be 1 or 3");
  if (M == 0 || N == 0) return 0;
  const int64_t K32 = K - K % 32;  // streamed here; the remainder below
  if (K32 == 0)
    return gemm_bf16_general(c, tag, ta ? 'T' : 'N', 'N', M, N, K, alpha, A, lda, S, lds, beta, C,
                             ldc, split);
  const int64_t KB = K32 / 32;
  const int64_t cus = c->num_cus > 0 ? c->num_cus : 256;
  // the LDS-ring kernel, 256 x 288 tiles for op(A) = A and 128 x 288 for
  // op(A) = A^T, whose rows are separate 128-byte streams per k block (columns
  // of A): 128 of them measured 6% faster than the 64 x 576 tile at m = 65536,
  // n = 16384, l = 544, and 192 or 256 of them slower (open DRAM pages and
  // TLB entries); for op(A) = A^T with 288 < l <= 544 the 128 x 544 tile
  // (kNTWW).  The k range is split while the grid is short of the CU count
  // (keeping at least 8 k blocks per slice).
  const int64_t bm2 = ta ? 64 * kMT2T : kBM2;
  const int64_t gx2 = (M + bm2 - 1) / bm2;
  const int64_t ntiles = (N + 15) / 16;
  const bool wide = ta && ntiles > kBNT2 && ntiles <= kBNTW;
  const int64_t bnt2 = wide ? kBNTW : kBNT2;
  const int64_t gy2 = (ntiles + bnt2 - 1) / bnt2;
  // k slices: one workgroup per CU runs at a time, so the grid runs in
  // ceil(gx gy ks / cus) waves; take the smallest ks (at least 8 k blocks a
  // slice) whose waves are within 3% of the best fill, e.g. 314 workgroups
  // (61% fill of the last of two waves) -> ks = 3 (942: 92%)
  int ks = 1;
  {
    double best = 0.0;
    double fill[9] = {0.0};
    for (int t = 1; t <= 8 && KB / t >= 8; ++t) {
      const double w = (double)(gx2 * gy2 * t) / (double)cus;
      fill[t] = w / std::ceil(w);
      if (fill[t] > best) best = fill[t];
    }
    for (int t = 1; t <= 8 && KB / t >= 8; ++t)
      if (fill[t] >= best - 0.03) { ks = t; break; }
  }
  const bool use2 = gx2 * gy2 * ks >= cus && gx2 * gy2 * ks <= 0x7fffffff;
  const int64_t gx = use2 ? gx2 : (M + kBM - 1) / kBM;
  const int64_t gy = use2 ? gy2 : ((N + 15) / 16 + kBNT - 1) / kBNT;
  if (gy > 65535 || gx > 0x7fffffff) return set_error("apass: grid too large");
  const int64_t NTP = gy * (use2 ? bnt2 : kBNT);
  const int planes = (split == 3) ? 2 : 1;
  DBuf<bf16x8_t> St;
  RSVD_TRY(St.alloc((size_t)(planes * KB * NTP * 64)));
  {
    const int64_t total = KB * NTP * 64;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(tile_skinny_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, K32,
                       N, S, lds, KB, NTP, planes, St.p);
    RSVD_LAUNCH_CHECK();
  }
  const bool vec = ta && (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (lda % 4 == 0);
  if (use2) {
    DBuf<float> Wk;
    if (ks > 1) RSVD_TRY(Wk.alloc((size_t)(ks * M * NTP * 16)));
    const int64_t KBper = (KB + ks - 1) / ks;
    // row blocks padded to a multiple of 8 so that the XCD-grouped block
    // mapping applies (the padding blocks exit at once)
    const int64_t gxp = (gx + 7) / 8 * 8;
    const unsigned blocks = (unsigned)(gxp * gy * ks);
    {
      KTimer kt(c, tag);
#define RSVD_APASS2(TA_, S_, V_, ...)                                                     \
  hipLaunchKernelGGL((apass2_kernel<TA_, S_, V_, TA_ ? kMT2T : kMT2, ##__VA_ARGS__>), dim3(blocks), \
                     dim3(kT2), 0, c->stream, M, N, KB,                                           \
                     KBper, alpha, A, lda, (const bf16x8_t*)St.p, NTP, beta, C, ldc, Wk.p,        \
                     (int)gxp, (int)gy, ks, (const float*)nullptr, (const float*)nullptr)
      if (ta && wide) {
        if (split == 3) {
          if (vec) RSVD_APASS2(true, 3, true, kNTWW, 2);
          else RSVD_APASS2(true, 3, false, kNTWW, 2);
        } else {
          if (vec) RSVD_APASS2(true, 1, true, kNTWW, 2);
          else RSVD_APASS2(true, 1, false, kNTWW, 2);
        }
      } else if (ta) {
        if (split == 3) {
          if (vec) RSVD_APASS2(true, 3, true);
          else RSVD_APASS2(true, 3, false);
        } else {
          if (vec) RSVD_APASS2(true, 1, true);
          else RSVD_APASS2(true, 1, false);
        }
      } else {
        if (split == 3) RSVD_APASS2(false, 3, false);
        else RSVD_APASS2(false, 1, false);
      }
#undef RSVD_APASS2
      RSVD_LAUNCH_CHECK();
      if (ks > 1) {
        int64_t rb = (M * N + 255) / 256;
        if (rb > 8192) rb = 8192;
        hipLaunchKernelGGL(apass2_reduce_kernel, dim3((unsigned)rb), dim3(256), 0, c->stream, M, N,
                           ks, (const float*)Wk.p, NTP * 16, alpha, beta, C, ldc);
        RSVD_LAUNCH_CHECK();
      }
    }
  } else {
  dim3 grid((unsigned)gx, (unsigned)gy);
  {
    KTimer kt(c, tag);
#define RSVD_APASS(TA_, S_, V_)                                                               \
  hipLaunchKernelGGL((apass_kernel<TA_, S_, V_>), grid, dim3(kThreads), 0, c->stream, M, N, K32, \
                     alpha, A, lda, (const bf16x8_t*)St.p, KB, NTP, beta, C, ldc)
    if (ta) {
      if (split == 3) {
        if (vec) RSVD_APASS(true, 3, true);
        else RSVD_APASS(true, 3, false);
      } else {
        if (vec) RSVD_APASS(true, 1, true);
        else RSVD_APASS(true, 1, false);
      }
    } else {
      if (split == 3) RSVD_APASS(false, 3, false);
      else RSVD_APASS(false, 1, false);
    }
#undef RSVD_APASS
    RSVD_LAUNCH_CHECK();
  }
  }
  if (K32 < K) {
    // the last K % 32 columns of op(A): C += alpha op(A)(:, K32:) S(K32:, :)
    const float* Ar = ta ? A + K32 : A + K32 * lda;
    RSVD_TRY(gemm_bf16_general(c, tag, ta ? 'T' : 'N', 'N', M, N, K - K32, alpha, Ar, lda,
                               S + K32, lds, 1.0f, C, ldc, split));
  }
  return 0;
}

static int column_scales(Context* c, int64_t M, int64_t N, const float* A, int64_t lda, float top,
                         float* scale);

int fp8_column_scales(Context* c, int64_t M, int64_t N, const float* A, int64_t lda, float* scale) {
  return column_scales(c, M, N, A, lda, 448.0f, scale);
}

int fp16_column_scales(Context* c, int64_t M, int64_t N, const float* A, int64_t lda,
                       float* scale) {
  return column_scales(c, M, N, A, lda, 32768.0f, scale);
}

static int column_scales(Context* c, int64_t M, int64_t N, const float* A, int64_t lda, float top,
                         float* scale) {
  const int64_t nc = (N + 127) / 128;
  if (nc == 0) return 0;
  DBuf<unsigned> bits;
  RSVD_TRY(bits.alloc((size_t)nc));
  RSVD_HIP(hipMemsetAsync(bits.p, 0, sizeof(unsigned) * nc, c->stream));
  // slices of ~2048 rows: enough workgroups (and loads in flight) to stream A
  // at HBM rate
  int64_t gy = (M + 2047) / 2048;
  if (gy > 256) gy = 256;
  if (gy < 1) gy = 1;
  hipLaunchKernelGGL(chunk_absmax_kernel, dim3((unsigned)nc, (unsigned)gy), dim3(256), 0,
                     c->stream, M, N, A, lda, bits.p);
  RSVD_LAUNCH_CHECK();
  hipLaunchKernelGGL(chunk_scale_finish_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0,
                     c->stream, nc, (const unsigned*)bits.p, top, scale);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int apass_fp8(Context* c, const char* tag, bool ta, int64_t M, int64_t N, int64_t K, float alpha,
              const float* A, int64_t lda, const float* ascale, const float* S, int64_t lds,
              float beta, float* C, int64_t ldc) {
  if (M < 0 || N < 0 || K < 0) return set_error("apass_fp8: negative dimension");
  if (!ascale) return set_error("apass_fp8: column scales of A required");
  if (M == 0 || N == 0) return 0;
  const int64_t K32 = K - K % 32;
  if (K32 == 0)
    return gemm_bf16_general(c, tag, ta ? 'T' : 'N', 'N', M, N, K, alpha, A, lda, S, lds, beta, C,
                             ldc, 3);
  const int64_t KB = K32 / 32;
  const int64_t gx = (M + kBM - 1) / kBM;
  const int64_t gy = ((N + 15) / 16 + kBNT - 1) / kBNT;
  if (gy > 65535 || gx > 0x7fffffff) return set_error("apass_fp8: grid too large");
  const int64_t NTP = gy * kBNT;
  DBuf<float> b;
  DBuf<fp8x8_t> St;
  RSVD_TRY(b.alloc((size_t)N));
  RSVD_TRY(St.alloc((size_t)(KB * NTP * 64)));
  const float* afold = ta ? nullptr : ascale;  // op(A) = A: a_c folded into S
  hipLaunchKernelGGL(skinny_colscale_kernel, dim3((unsigned)N), dim3(256), 0, c->stream, K32, S,
                     lds, afold, 448.0f, b.p);
  RSVD_LAUNCH_CHECK();
  {
    const int64_t total = KB * NTP * 64;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(tile_skinny8_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, K32,
                       N, S, lds, KB, NTP, afold, (const float*)b.p, St.p);
    RSVD_LAUNCH_CHECK();
  }
  const bool vec = ta && (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (lda % 4 == 0);
  {
    KTimer kt(c, tag);
    dim3 grid((unsigned)gx, (unsigned)gy);
    if (ta) {
      if (vec)
        hipLaunchKernelGGL((apass8_kernel<true, true>), grid, dim3(kThreads), 0, c->stream, M, N,
                           K32, alpha, A, lda, (const fp8x8_t*)St.p, KB, NTP, ascale,
                           (const float*)b.p, beta, C, ldc);
      else
        hipLaunchKernelGGL((apass8_kernel<true, false>), grid, dim3(kThreads), 0, c->stream, M, N,
                           K32, alpha, A, lda, (const fp8x8_t*)St.p, KB, NTP, ascale,
                           (const float*)b.p, beta, C, ldc);
    } else {
      hipLaunchKernelGGL((apass8_kernel<false, false>), grid, dim3(kThreads), 0, c->stream, M, N,
                         K32, alpha, A, lda, (const fp8x8_t*)St.p, KB, NTP, ascale,
                         (const float*)b.p, beta, C, ldc);
    }
    RSVD_LAUNCH_CHECK();
  }
  if (K32 < K) {
    // the last K % 32 columns of op(A) in split-bf16: C += alpha op(A)(:, K32:) S(K32:, :)
    const float* Ar = ta ? A + K32 : A + K32 * lda;
    RSVD_TRY(gemm_bf16_general(c, tag, ta ? 'T' : 'N', 'N', M, N, K - K32, alpha, Ar, lda,
                               S + K32, lds, 1.0f, C, ldc, 3));
  }
  return 0;
}

int apass_fp16(Context* c, const char* tag, bool ta, int64_t M, int64_t N, int64_t K, float alpha,
               const float* A, int64_t lda, const float* ascale, const float* S, int64_t lds,
               float beta, float* C, int64_t ldc) {
  if (M < 0 || N < 0 || K < 0) return set_error("apass_fp16: negative dimension");
  if (!ascale) return set_error("apass_fp16: column scales of A required");
  if (M == 0 || N == 0) return 0;
  const int64_t K32 = K - K % 32;
  const int64_t KB = K32 / 32;
  const int64_t cus = c->num_cus > 0 ? c->num_cus : 256;
  const int64_t bm2 = ta ? 64 * kMT2T : kBM2;
  const int64_t gx2 = (M + bm2 - 1) / bm2;
  const int64_t gy2 = ((N + 15) / 16 + kBNT2 - 1) / kBNT2;
  int ks = 1;
  {
    double best = 0.0;
    double fill[9] = {0.0};
    for (int t = 1; t <= 8 && KB / t >= 8; ++t) {
      const double w = (double)(gx2 * gy2 * t) / (double)cus;
      fill[t] = w / std::ceil(w);
      if (fill[t] > best) best = fill[t];
    }
    for (int t = 1; t <= 8 && KB / t >= 8; ++t)
      if (fill[t] >= best - 0.03) { ks = t; break; }
  }
  // the LDS-ring kernel only; shapes it does not fill run in split-bf16
  if (K32 == 0 || KB < 8 || gx2 * gy2 * ks < cus || gy2 > 65535 || gx2 * gy2 * ks > 0x7fffffff)
    return gemm_bf16(c, tag, ta ? 'T' : 'N', 'N', M, N, K, alpha, A, lda, S, lds, beta, C, ldc,
                     3);
  const int64_t NTP = gy2 * kBNT2;
  DBuf<float> b;
  DBuf<bf16x8_t> St;
  RSVD_TRY(b.alloc((size_t)N));
  RSVD_TRY(St.alloc((size_t)(2 * KB * NTP * 64)));
  const float* afold = ta ? nullptr : ascale;  // op(A) = A: a_c folded into S
  hipLaunchKernelGGL(skinny_colscale_kernel, dim3((unsigned)N), dim3(256), 0, c->stream, K32, S,
                     lds, afold, 32768.0f, b.p);
  RSVD_LAUNCH_CHECK();
  {
    const int64_t total = KB * NTP * 64;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(tile_skinny16_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, K32,
                       N, S, lds, KB, NTP, afold, (const float*)b.p, St.p);
    RSVD_LAUNCH_CHECK();
  }
  const bool vec = ta && (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (lda % 4 == 0);
  DBuf<float> Wk;
  if (ks > 1) RSVD_TRY(Wk.alloc((size_t)(ks * M * NTP * 16)));
  const int64_t KBper = (KB + ks - 1) / ks;
  const int64_t gxp = (gx2 + 7) / 8 * 8;
  const unsigned blocks = (unsigned)(gxp * gy2 * ks);
  {
    KTimer kt(c, tag);
#define RSVD_APASS16(TA_, V_)                                                                  \
  hipLaunchKernelGGL((apass2_kernel<TA_, 2, V_, TA_ ? kMT2T : kMT2>), dim3(blocks), dim3(kT2), 0, \
                     c->stream, M, N, KB, KBper, alpha, A, lda, (const bf16x8_t*)St.p, NTP, beta,  \
                     C, ldc, Wk.p, (int)gxp, (int)gy2, ks, ascale, (const float*)b.p)
    if (ta) {
      if (vec) RSVD_APASS16(true, true);
      else RSVD_APASS16(true, false);
    } else {
      RSVD_APASS16(false, false);
    }
#undef RSVD_APASS16
    RSVD_LAUNCH_CHECK();
    if (ks > 1) {
      int64_t rb = (M * N + 255) / 256;
      if (rb > 8192) rb = 8192;
      hipLaunchKernelGGL(apass2_reduce_kernel, dim3((unsigned)rb), dim3(256), 0, c->stream, M, N,
                         ks, (const float*)Wk.p, NTP * 16, alpha, beta, C, ldc);
      RSVD_LAUNCH_CHECK();
    }
  }
  if (K32 < K) {
    // the last K % 32 columns of op(A) in split-bf16
    const float* Ar = ta ? A + K32 : A + K32 * lda;
    RSVD_TRY(gemm_bf16_general(c, tag, ta ? 'T' : 'N', 'N', M, N, K - K32, alpha, Ar, lda,
                               S + K32, lds, 1.0f, C, ldc, 3));
  }
  return 0;
}

}  // namespace rsvd
