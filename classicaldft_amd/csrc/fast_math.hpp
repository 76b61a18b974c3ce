// FP64 logarithm and reciprocal for the pointwise kernels.
//
// The library log (about 90 instructions on gfx950, most of them for
// argument classes these kernels never see) was the largest single cost of
// the pointwise FMT step: two logs per point, log(1 - eta) and log(rho).
// flog reaches the same accuracy (within 2 ulp, tested against numpy's log)
// in about 35:
//   x = 2^e m with m in [sqrt(1/2), sqrt(2)),
//   log m = 2 atanh(s),  s = (m - 1) / (m + 1),  |s| <= 0.1716,
//   atanh(s) / s = sum_k s^2k / (2k + 1), truncated after k = 9 (the
//   remainder is below 2.3e-17 relative, since s^2 <= 0.0295),
//   log x = e ln2 + log m, with ln2 split into a head of 42 significant bits
//   (so e * head is exact for every binary64 exponent) and a tail.
// x = 0 gives -inf, x < 0 and NaN give NaN, +inf gives +inf; subnormal x is
// decomposed exactly by frexp.
//
// frcp: 1/d by the hardware estimate refined with two Newton steps and a
// final residual correction (within 1 ulp), in place of the IEEE division
// sequence; valid for normal d whose reciprocal is normal.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace cdft {

__host__ __device__ inline double frcp(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / d;
#endif
}

// q = n / d for d in [1, 4): the quotient by frcp, then one residual step.
__host__ __device__ inline double fdiv_small(double n, double d) {
  const double r = frcp(d);
  const double q = n * r;
  return fma(fma(-q, d, n), r, q);
}

__host__ __device__ inline double flog(double x) {
  int e;
  double m;
#if defined(__HIP_DEVICE_COMPILE__)
  m = __builtin_amdgcn_frexp_mant(x);
  e = __builtin_amdgcn_frexp_exp(x);
#else
  m = std::frexp(x, &e);
#endif
  // m in [0.5, 1): move [0.5, sqrt(1/2)) up by one octave
  const bool lo = m < 0.70710678118654752440;
  m = lo ? 2.0 * m : m;
  e = lo ? e - 1 : e;
  const double f = m - 1.0;  // exact (Sterbenz)
  const double s = fdiv_small(f, f + 2.0);
  const double z = s * s;
  // (atanh(s)/s - 1)/z = 1/3 + z/5 + ... + z^8/19
  double q = 1.0 / 19.0;
  q = fma(q, z, 1.0 / 17.0);
  q = fma(q, z, 1.0 / 15.0);
  q = fma(q, z, 1.0 / 13.0);
  q = fma(q, z, 1.0 / 11.0);
  q = fma(q, z, 1.0 / 9.0);
  q = fma(q, z, 1.0 / 7.0);
  q = fma(q, z, 1.0 / 5.0);
  q = fma(q, z, 1.0 / 3.0);
  const double s2 = s + s;
  const double lm = fma(s2 * z, q, s2);
  const double de = static_cast<double>(e);
  const double r = fma(de, 0x1.62e42fefa3800p-1, fma(de, 0x1.ef35793c76730p-45, lm));
  // special arguments: 0 -> -inf, negative or NaN -> NaN, +inf -> +inf
  const double inf = __builtin_huge_val();
  if (x > 0.0 && x < inf) return r;
  return x == 0.0 ? -inf : (x == inf ? inf : __builtin_nan(""));
}

}  // namespace cdft
