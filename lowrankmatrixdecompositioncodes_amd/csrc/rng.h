// rng.h -- the seeded Gaussian stream on the device (shared by every kernel
// that draws Omega or test matrices).
//
// Element e of the stream for `seed` is lane e % 4 of the Philox4x64-10 block
// with counter (e / 4, 0, 0, 0) and key (seed, 0); the 64-bit word u maps to
// ((u >> 11) + 0.5) * 2^-53, then through the AS241 inverse normal CDF, then
// is rounded to float and widened.  The arithmetic is written contraction-free
// so the device produces the same bits as a plain IEEE double evaluation on
// the host.
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace rsvd {

static __device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t lo0 = 0xD2E7470EE14C6C93ULL * c[0];
    uint64_t hi0 = __umul64hi(0xD2E7470EE14C6C93ULL, c[0]);
    uint64_t lo1 = 0xCA5A826395121157ULL * c[2];
    uint64_t hi1 = __umul64hi(0xCA5A826395121157ULL, c[2]);
    uint64_t n0 = hi1 ^ c[1] ^ k0;
    uint64_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B97F4A7C15ULL;
    k1 += 0xBB67AE8584CAA73BULL;
  }
}

// log(x), x > 0 finite: x = f*2^e, f in [sqrt(1/2), sqrt(2)), 2*atanh series.
static __device__ double dev_log(double x) {
#pragma clang fp contract(off)
  int e;
  double f = frexp(x, &e);
  if (f < 0.70710678118654752440) {
    f = f * 2.0;
    e -= 1;
  }
  double t = (f - 1.0) / (f + 1.0);
  double t2 = t * t;
  double s = 1.0 / 27.0;
  for (int j = 25; j >= 1; j -= 2) s = s * t2 + 1.0 / (double)j;
  double lf = 2.0 * t * s;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  return ((double)e * ln2_lo + lf) + (double)e * ln2_hi;
}

// Wichura (1988) AS241 PPND16 inverse standard normal CDF.
static __device__ double dev_ndtri(double p) {
#pragma clang fp contract(off)
  double q = p - 0.5, r, x;
  if (fabs(q) <= 0.425) {
    r = 0.180625 - q * q;
    x = q *
        (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
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
  r = sqrt(-dev_log(r));
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

static __device__ __forceinline__ double gaussian_at(uint64_t seed, uint64_t e) {
  uint64_t c[4] = {e >> 2, 0, 0, 0};
  philox4x64(c, seed, 0);
  uint64_t w = c[e & 3];
  double u = ((double)(w >> 11) + 0.5) * 0x1.0p-53;
  return (double)(float)dev_ndtri(u);
}

// Elements 4 b .. 4 b + 3 of the stream from one Philox block: g[w] equals
// gaussian_at(seed, 4 b + w) bit for bit.
static __device__ __forceinline__ void gaussian_block4(uint64_t seed, uint64_t b, double (&g)[4]) {
  uint64_t c[4] = {b, 0, 0, 0};
  philox4x64(c, seed, 0);
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const double u = ((double)(c[w] >> 11) + 0.5) * 0x1.0p-53;
    g[w] = (double)(float)dev_ndtri(u);
  }
}

}  // namespace rsvd
