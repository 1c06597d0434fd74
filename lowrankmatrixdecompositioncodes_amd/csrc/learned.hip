// learned.hip -- the learned (q, rank) front end of the TOL-driven low-rank
// approximation: a small multilayer perceptron, trained and served by the
// kernels below, predicts the power-iteration count q and the rank from the
// energy profile of a cheap 64-column bf16 sketch, so that one fixed-rank rSVD
// (rra.c:73-234) can replace the adaptive blocked loop
// (low_rank_svd_blockrand_decomp_fixed_rank_or_prec, rra.c:239-381).  The
// adaptive loop stays as the fallback and the judge: a prediction whose FP64
// residual misses TOL is redone by it.
//
// Network (fp32; weights [out][in] row-major, packed W1 b1 W2 b2 W3 b3 W4 b4):
//   64 -> 128 -> 128 -> 128 (GELU, erf form) -> 6
// outputs 0..4: logits of q in {0..4}; output 5: log(rank) / log(min(m, n)).
// Loss per sample: logsumexp(o[0:5]) - o[q] + lam (o[5] - r)^2, batch mean.
//
// Training: one launch computes forward and backward for chunks of 16
// samples per workgroup (activations in LDS) and writes per-workgroup
// gradient partials; a second launch sums them in a fixed order (so a step is
// deterministic); a third applies Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "rng.h"
#include "rsvd_internal.h"

namespace rsvd {

namespace {

constexpr int kIn = 64, kH = 128, kOut = 6, kQ = 5;
constexpr int kSPB = 16;  // samples per workgroup
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kIn, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kH,
              kOffW4 = kOffB3 + kH, kOffB4 = kOffW4 + kOut * kH, kNParams = kOffB4 + kOut;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_prime(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// z = W in + b (and h = gelu(z)) for the SPB samples of the workgroup
template <int IN, int OUT>
__device__ void layer_fwd(const float* __restrict__ W, const float* __restrict__ b,
                          const float* in, float* z, float* h) {
  for (int e = threadIdx.x; e < kSPB * OUT; e += blockDim.x) {
    const int s = e / OUT, o = e % OUT;
    const float* w = W + o * IN;
    const float* x = in + s * IN;
    float acc = b[o];
    for (int i = 0; i < IN; ++i) acc += w[i] * x[i];
    z[e] = acc;
    if (h) h[e] = gelu(acc);
  }
  __syncthreads();
}

// weight / bias gradients of one layer from its output deltas (SPB x OUT)
// and its input activations (SPB x IN); delta_in = (delta W) * gelu'(zin)
template <int IN, int OUT>
__device__ void layer_bwd(const float* __restrict__ W, const float* delta, const float* hin,
                          const float* zin, float* __restrict__ gW, float* __restrict__ gb,
                          float* delta_in) {
  for (int e = threadIdx.x; e < OUT * IN; e += blockDim.x) {
    const int o = e / IN, i = e % IN;
    float acc = 0.0f;
    for (int s = 0; s < kSPB; ++s) acc += delta[s * OUT + o] * hin[s * IN + i];
    gW[e] = acc;
  }
  for (int o = threadIdx.x; o < OUT; o += blockDim.x) {
    float acc = 0.0f;
    for (int s = 0; s < kSPB; ++s) acc += delta[s * OUT + o];
    gb[o] = acc;
  }
  if (delta_in) {
    for (int e = threadIdx.x; e < kSPB * IN; e += blockDim.x) {
      const int s = e / IN, i = e % IN;
      float acc = 0.0f;
      for (int o = 0; o < OUT; ++o) acc += delta[s * OUT + o] * W[o * IN + i];
      delta_in[e] = acc * gelu_prime(zin[e]);
    }
  }
  __syncthreads();
}

// Forward (and with BWD the backward) pass for samples [16 wg, 16 wg + 16)
// of the batch; sample s of the batch is row idx[s] (or s) of X.
template <bool BWD>
__global__ void __launch_bounds__(256)
mlp_kernel(const float* __restrict__ theta, const float* __restrict__ X,
           const int* __restrict__ idx, const int* __restrict__ qlab,
           const float* __restrict__ rlab, float lam, int B, float* __restrict__ out,
           float* __restrict__ gpart, float* __restrict__ loss) {
  __shared__ float xs[kSPB * kIn];
  __shared__ float z1[kSPB * kH], h1[kSPB * kH], z2[kSPB * kH], h2[kSPB * kH];
  __shared__ float z3[kSPB * kH], h3[kSPB * kH], z4[kSPB * kOut];
  __shared__ float dA[kSPB * kH], dB[kSPB * kH], d4[kSPB * kOut];
  const int s0 = blockIdx.x * kSPB;
  const int ns = min(kSPB, B - s0);
  for (int e = threadIdx.x; e < kSPB * kIn; e += blockDim.x) {
    const int s = e / kIn, i = e % kIn;
    float v = 0.0f;
    if (s < ns) {
      const int row = idx ? idx[s0 + s] : s0 + s;
      v = X[(int64_t)row * kIn + i];
    }
    xs[e] = v;
  }
  __syncthreads();
  layer_fwd<kIn, kH>(theta + kOffW1, theta + kOffB1, xs, z1, h1);
  layer_fwd<kH, kH>(theta + kOffW2, theta + kOffB2, h1, z2, h2);
  layer_fwd<kH, kH>(theta + kOffW3, theta + kOffB3, h2, z3, h3);
  layer_fwd<kH, kOut>(theta + kOffW4, theta + kOffB4, h3, z4, nullptr);
  if (!BWD) {
    for (int e = threadIdx.x; e < ns * kOut; e += blockDim.x) out[(int64_t)s0 * kOut + e] = z4[e];
    return;
  }
  // output deltas of the mean loss
  if (threadIdx.x < kSPB) {
    const int s = threadIdx.x;
    float* d = d4 + s * kOut;
    if (s < ns) {
      const float* o = z4 + s * kOut;
      const int q = qlab[idx ? idx[s0 + s] : s0 + s];
      const float r = rlab[idx ? idx[s0 + s] : s0 + s];
      float mx = o[0];
      for (int c = 1; c < kQ; ++c) mx = fmaxf(mx, o[c]);
      float se = 0.0f;
      for (int c = 0; c < kQ; ++c) se += expf(o[c] - mx);
      const float lse = mx + logf(se);
      const float invB = 1.0f / (float)B;
      for (int c = 0; c < kQ; ++c) d[c] = (expf(o[c] - lse) - (c == q ? 1.0f : 0.0f)) * invB;
      const float dr = o[5] - r;
      d[5] = 2.0f * lam * dr * invB;
      loss[s0 + s] = lse - o[q] + lam * dr * dr;
    } else {
      for (int c = 0; c < kOut; ++c) d[c] = 0.0f;
    }
  }
  __syncthreads();
  float* g = gpart + (int64_t)blockIdx.x * kNParams;
  layer_bwd<kH, kOut>(theta + kOffW4, d4, h3, z3, g + kOffW4, g + kOffB4, dA);
  layer_bwd<kH, kH>(theta + kOffW3, dA, h2, z2, g + kOffW3, g + kOffB3, dB);
  layer_bwd<kH, kH>(theta + kOffW2, dB, h1, z1, g + kOffW2, g + kOffB2, dA);
  layer_bwd<kIn, kH>(theta + kOffW1, dA, xs, nullptr, g + kOffW1, g + kOffB1, nullptr);
}

__global__ void mlp_reduce_kernel(const float* __restrict__ gpart, int nwg,
                                  float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParams) return;
  float acc = 0.0f;
  for (int w = 0; w < nwg; ++w) acc += gpart[(int64_t)w * kNParams + p];
  grad[p] = acc;
}

__global__ void mlp_adam_kernel(float* __restrict__ theta, const float* __restrict__ grad,
                                float* __restrict__ m1, float* __restrict__ m2, float lr,
                                float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParams) return;
  const float g = grad[p];
  const float a = 0.9f * m1[p] + 0.1f * g;
  const float b = 0.999f * m2[p] + 0.001f * g * g;
  m1[p] = a;
  m2[p] = b;
  theta[p] -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
}

__global__ void mlp_init_kernel(uint64_t seed, float* __restrict__ theta) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParams) return;
  float scale = 0.0f;  // biases start at zero
  if (p < kOffB1) scale = sqrtf(2.0f / kIn);
  else if (p >= kOffW2 && p < kOffB2) scale = sqrtf(2.0f / kH);
  else if (p >= kOffW3 && p < kOffB3) scale = sqrtf(2.0f / kH);
  else if (p >= kOffW4 && p < kOffB4) scale = sqrtf(1.0f / kH);
  theta[p] = scale == 0.0f ? 0.0f : scale * (float)gaussian_at(seed, (uint64_t)p);
}

// 64 features from the energy a 64-column sketch leaves behind: with lam_0 >=
// ... >= lam_63 the squared singular values of B = Q^T A (Q an orthonormal
// basis of the sketch) and a2 = ||A||_F^2, x_j = log10(t_j / a2) / 8 with
// t_j = a2 - (lam_0 + ... + lam_j) the squared residual of the best rank
// j + 1 approximation inside range(Q), floored at 1e-16 a2 (x_j in [-2, 0]).
__global__ void sketch_features_kernel(const double* __restrict__ lam, int k,
                                       const double* __restrict__ a2p,
                                       float* __restrict__ feat) {
  const int j = threadIdx.x;
  if (j >= k) return;
  const double a2 = *a2p;
  double cap = 0.0;
  for (int i = 0; i <= j; ++i) cap += fmax(lam[i], 0.0);
  const double t = fmax(a2 - cap, 1e-16 * a2);
  feat[j] = (a2 > 0.0) ? (float)(log10(t / a2) / 8.0) : -2.0f;
}

}  // namespace

int mlp_num_params() { return kNParams; }

int mlp_init(Context* c, uint64_t seed, float* theta) {
  hipLaunchKernelGGL(mlp_init_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, c->stream, seed,
                     theta);
  RSVD_LAUNCH_CHECK();
  return 0;
}

int mlp_forward(Context* c, const float* theta, const float* X, int B, float* out) {
  if (B <= 0) return set_error("mlp_forward: batch must be positive");
  hipLaunchKernelGGL(mlp_kernel<false>, dim3((B + kSPB - 1) / kSPB), dim3(256), 0, c->stream,
                     theta, X, (const int*)nullptr, (const int*)nullptr, (const float*)nullptr,
                     0.0f, B, out, (float*)nullptr, (float*)nullptr);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// gradient of the mean loss over rows idx[0..B) (or 0..B) of X into grad;
// per-sample losses into loss (B floats)
int mlp_grad(Context* c, const float* theta, const float* X, const int* idx, const int* q,
             const float* r, float lam, int B, float* grad, float* loss) {
  if (B <= 0) return set_error("mlp_grad: batch must be positive");
  const int nwg = (B + kSPB - 1) / kSPB;
  DBuf<float> gpart;
  RSVD_TRY(gpart.alloc((size_t)nwg * kNParams));
  hipLaunchKernelGGL(mlp_kernel<true>, dim3(nwg), dim3(256), 0, c->stream, theta, X, idx, q, r,
                     lam, B, (float*)nullptr, gpart.p, loss);
  RSVD_LAUNCH_CHECK();
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, c->stream,
                     (const float*)gpart.p, nwg, grad);
  RSVD_LAUNCH_CHECK();
  return 0;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) over shuffled mini-batches of the
// n samples (X n x 64, labels q, r); epochs x ceil(n / batch) steps; the
// mean loss of each epoch into loss_hist (host, epochs entries, optional).
int mlp_train(Context* c, float* theta, const float* X, const int* q, const float* r, int n,
              float lam, int epochs, int batch, float lr, uint64_t seed, float* loss_hist) {
  if (n <= 0 || batch <= 0 || epochs < 0) return set_error("mlp_train: bad sizes");
  DBuf<float> grad, m1, m2, loss;
  DBuf<int> idx;
  RSVD_TRY(grad.alloc(kNParams));
  RSVD_TRY(m1.alloc(kNParams));
  RSVD_TRY(m2.alloc(kNParams));
  RSVD_TRY(loss.alloc((size_t)n));
  RSVD_TRY(idx.alloc((size_t)n));
  RSVD_HIP(hipMemsetAsync(m1.p, 0, sizeof(float) * kNParams, c->stream));
  RSVD_HIP(hipMemsetAsync(m2.p, 0, sizeof(float) * kNParams, c->stream));
  std::vector<int> perm((size_t)n);
  std::vector<float> lh((size_t)n);
  uint64_t state = seed * 0x9E3779B97F4A7C15ULL + 1;
  int t = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < n; ++i) perm[(size_t)i] = i;
    for (int i = n - 1; i > 0; --i) {  // Fisher-Yates with a 64-bit LCG
      state = state * 6364136223846793005ULL + 1442695040888963407ULL;
      const int j = (int)((state >> 33) % (uint64_t)(i + 1));
      std::swap(perm[(size_t)i], perm[(size_t)j]);
    }
    RSVD_HIP(hipMemcpyAsync(idx.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                            c->stream));
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int B = std::min(batch, n - b0);
      RSVD_TRY(mlp_grad(c, theta, X, idx.p + b0, q, r, lam, B, grad.p, loss.p + b0));
      ++t;
      const float c1 = 1.0f / (1.0f - powf(0.9f, (float)t));
      const float c2 = 1.0f / (1.0f - powf(0.999f, (float)t));
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, c->stream,
                         theta, (const float*)grad.p, m1.p, m2.p, lr, c1, c2);
      RSVD_LAUNCH_CHECK();
    }
    if (loss_hist) {
      RSVD_HIP(hipMemcpyAsync(lh.data(), loss.p, sizeof(float) * n, hipMemcpyDeviceToHost,
                              c->stream));
      RSVD_HIP(hipStreamSynchronize(c->stream));
      double s = 0.0;
      for (float v : lh) s += v;
      loss_hist[ep] = (float)(s / n);
    }
  }
  return 0;
}

// Features of one fp32 matrix: Y = A Omega (n x 64 Gaussian, seeded stream)
// on the bf16 matrix cores, Y^T Y in FP64 from fp32 Y, its eigenvalues, 64
// features into feat (device).
int sketch_features(Context* c, const double* A, const float* A32, int64_t m, int64_t n,
                    int64_t lda, uint64_t seed, float* feat) {
  const int k = kIn;
  if (m < k || n < k) return set_error("sketch features: need min(m, n) >= %d", k);
  DBuf<float> Om, Y;
  DBuf<double> Y64, B, G, lam, a2;
  RSVD_TRY(Om.alloc((size_t)(n * k)));
  RSVD_TRY(Y.alloc((size_t)(m * k)));
  RSVD_TRY(Y64.alloc((size_t)(m * k)));
  RSVD_TRY(B.alloc((size_t)(k * n)));
  RSVD_TRY(G.alloc((size_t)(k * k)));
  RSVD_TRY(lam.alloc((size_t)k));
  RSVD_TRY(a2.alloc(1));
  // the sketch on the bf16 matrix cores (split-bf16 products: the basis is
  // that of the float matrix to ~1e-7), its orthonormal basis in FP64, and
  // the captured energy of the FP64 matrix (B = Q^T A, G = B B^T)
  RSVD_TRY(gaussian_f32(c, seed, n, k, 0, n, Om.p, n));
  RSVD_TRY(gemm_bf16(c, "sketch_bf16", 'N', 'N', m, k, n, 1.0f, A32, lda, Om.p, n, 0.0f, Y.p, m,
                     3));
  RSVD_TRY(f32_to_f64(c, m, k, Y.p, m, Y64.p, m));
  int fb = 0;
  RSVD_TRY(orth_f64(c, m, k, Y64.p, m, nullptr, &fb));
  RSVD_TRY(gemm_f64(c, 'T', 'N', k, n, m, 1.0, Y64.p, m, A, lda, 0.0, B.p, k));
  RSVD_TRY(gemm_f64(c, 'N', 'T', k, k, n, 1.0, B.p, k, B.p, k, 0.0, G.p, k));
  RSVD_TRY(symmetrize_upper_f64(c, k, G.p));
  int sw = 0;
  RSVD_TRY(symeig_jacobi_f64(c, k, G.p, k, nullptr, k, lam.p, 60, &sw, 1));
  RSVD_TRY(sort_pairs_f64(c, k, lam.p, 1, nullptr, 0, 0, nullptr, 0, 0));
  RSVD_TRY(sumsq_f64(c, m, n, A, lda, a2.p));
  hipLaunchKernelGGL(sketch_features_kernel, dim3(1), dim3(64), 0, c->stream,
                     (const double*)lam.p, k, (const double*)a2.p, feat);
  RSVD_LAUNCH_CHECK();
  return 0;
}

}  // namespace rsvd
