// FP64 tanh shared by the FP64 embedding and fitting kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace dpmi {
namespace {

// E = expm1(2|x|) from a Cody-Waite reduction and a degree-13 Taylor
// polynomial, then tanh(x) = sign(x) E / (E + 2) with a Newton-refined reciprocal;
// |x| >= 22 gives +-1.  Relative error about 1e-16.
__device__ inline double tanh64(double x) {
  const double y = fmin(2.0 * fabs(x), 44.0);
  const double n = rint(y * 1.4426950408889634);
  double r = fma(-n, 6.93147180369123816490e-01, y);
  r = fma(-n, 1.90821492927058770002e-10, r);
  // expm1(r) = r + r^2 (1/2 + r/3! + ... + r^11/13!),  |r| <= 0.347
  double q = 1.0 / 6227020800.0;
  q = fma(q, r, 1.0 / 479001600.0);
  q = fma(q, r, 1.0 / 39916800.0);
  q = fma(q, r, 1.0 / 3628800.0);
  q = fma(q, r, 1.0 / 362880.0);
  q = fma(q, r, 1.0 / 40320.0);
  q = fma(q, r, 1.0 / 5040.0);
  q = fma(q, r, 1.0 / 720.0);
  q = fma(q, r, 1.0 / 120.0);
  q = fma(q, r, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  const double em1r = fma(r * r, q, r);
  const int ni = (int)n;
  // expm1(y) = 2^n expm1(r) + (2^n - 1)
  const double e = ldexp(em1r, ni) + (ldexp(1.0, ni) - 1.0);
  const double d = e + 2.0;
  double rc = __builtin_amdgcn_rcp(d);
  rc = fma(rc, fma(-d, rc, 1.0), rc);
  rc = fma(rc, fma(-d, rc, 1.0), rc);
  const double t0 = e * rc;
  const double t = fma(rc, fma(-d, t0, e), t0);
  return copysign(t, x);
}


}  // namespace
}  // namespace dpmi
