// jacobi_rot.h -- the 2x2 symmetric Jacobi rotation shared by the device
// eigensolvers (symeig.hip, batched.hip).
#pragma once
#include <cmath>

#include <hip/hip_runtime.h>

namespace rsvd {

// no rotation below this relative size of the off-diagonal
constexpr double kRotTol = 1e-15;
// a sweep that met no off-diagonal above this relative size ends the iteration
constexpr double kStopTol = 1e-10;
// the same for callers that need eigenvectors to fp32 level only, on the
// rotation sines |s| rather than the off-diagonal ratio: near convergence a
// pair with relative eigenvalue gap d and ratio r turns by |s| ~ r / d, and
// the next sweep leaves ratios ~ r^2 / d = s^2 d.  A sweep whose rotations
// all had |s| <= 1e-4 therefore leaves off-diagonals below ~1e-8 of
// sqrt(g_pp g_qq) for every gap (a ratio test of 1e-4 alone could stop with
// ratios of 1e-4 inside a cluster of nearly equal eigenvalues, i.e.
// first-order eigenvalue errors of 1e-4 there).  A rotation of ratio at most
// kLooseRatioTol does not count, however large its angle: inside a cluster
// (gap below 1e-4) the angles stay large until the splitting the cluster's
// members carry -- for a plateau of equal singular values, the rounding of
// the preceding stages -- is resolved to rounding, while such a pair already
// perturbs its eigenvalues by at most 1e-8 relative and leaves the columns
// of B^T V Sigma^-1 orthogonal to 1e-8.  Either way the eigenvalues are good
// to ~1e-8 relative or better and the eigenvectors of separated eigenvalues,
// and the orthogonality of B^T V Sigma^-1, to ~1e-8, below the fp32
// resolution of the outputs.
constexpr double kLooseStopTol = 1e-4;
constexpr double kLooseRatioTol = 1e-8;

// 1/x and 1/sqrt(x) from the hardware estimates plus two Newton steps
// (error a few ulp; a rotation only needs c^2 + s^2 = 1 to rounding).
static __device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
static __device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  return fma(y, e, y);
}

// Rotation (c, s) with [x' y'] = [x y] [[c, s], [-s, c]] zeroing the (p, q)
// entry of a symmetric 2x2 [[a, g], [g, b]]: t = tan theta is the root of
// g t^2 + (b - a) t - g = 0 of smaller modulus, t = sgn(d) 2g / (|d| + r),
// d = b - a, r = sqrt(d^2 + 4g^2) (Rutishauser's choice).  With den = |d| + r,
// 1 + t^2 = 2r / den, so c = 1 / sqrt(1 + t^2) = den / sqrt(2 r den) and
// s = c t = sgn(d) 2g / sqrt(2 r den): two reciprocal square roots on the
// serial path and c^2 + s^2 = (den^2 + 4g^2) / (2 r den) = 1 identically.
// Evaluated on the entries scaled by a power of two so that the squares stay
// in range.  Branch-free.  No rotation when |g| <= flo or |g| <= kRotTol
// sqrt|ab|; *sig bit 0 reports |g| > kStopTol sqrt|ab|, bit 1 a rotation
// with |s| > kLooseStopTol and |g| > kLooseRatioTol sqrt|ab| (bit 1 implies
// bit 0).
static __device__ __forceinline__ void rotation(double a, double b, double g, double flo, double* cs,
                                         double* sn, int* sig) {
  const double ag = fabs(g);
  int e;
  (void)frexp(fmax(fmax(fabs(a), fabs(b)), ag), &e);
  const double sa = ldexp(a, -e), sb = ldexp(b, -e), sg = ldexp(g, -e);
  const double ab = fabs(sa * sb), g2 = sg * sg;
  const bool big = ag > flo;
  const bool rot = big & (g2 > kRotTol * kRotTol * ab);
  const double d = sb - sa;
  const double r2 = rot ? fma(d, d, 4.0 * g2) : 1.0;
  const double r = r2 * rsq_nr(r2);
  const double den = fabs(d) + r;
  const double q = rsq_nr(2.0 * r * den);
  const double t2g = (d < 0.0) ? -2.0 * sg : 2.0 * sg;
  *cs = rot ? den * q : 1.0;
  *sn = rot ? t2g * q : 0.0;
  *sig = ((big & (g2 > kStopTol * kStopTol * ab)) ? 1 : 0) |
         ((rot & (fabs(*sn) > kLooseStopTol) & (g2 > kLooseRatioTol * kLooseRatioTol * ab)) ? 2
                                                                                           : 0);
}

}  // namespace rsvd
