// Correctly rounded division by a divisor known in advance.
//
// The Thomas recurrences divide by data-independent pivots d_k.  With
// yb = RN(1/b) prepared once (an IEEE division), the quotient RN(a/b) is
//   q0 = RN(a*yb);  r0 = a - b*q0 (fma);  q1 = RN(q0 + r0*yb)  -> faithful
//   r1 = a - b*q1 (exact, fma);  q2 = RN(q1 + r1*yb)            -> RN(a/b)
// (Markstein's theorem: if yb = RN(1/b) and q1 is within one ulp of a/b, the
// remainder r1 is exact and RN(q1 + r1*yb) is the correctly rounded quotient),
// valid when no intermediate underflows or overflows, which the callers'
// magnitudes guarantee.  It replaces the hardware division sequence, whose
// numerator-dependent part is twice as long, on the serial critical path of
// each line solve; results are bit-identical to a/b (tested on the CPU
// against division and end to end on the GPU against the oracle), up to the
// sign of a zero quotient.
#pragma once

#ifdef __HIPCC__
#define VLS_HD __host__ __device__
#else
#define VLS_HD
#endif

#include <cmath>

namespace vls {

VLS_HD inline double div_known(double a, double b, double yb) {
  double q0 = a * yb;
  double r0 = std::fma(-b, q0, a);
  double q1 = std::fma(r0, yb, q0);
  double r1 = std::fma(-b, q1, a);
  return std::fma(r1, yb, q1);
}

}  // namespace vls
