// divrn.h - IEEE double division a / d by a precomputed reciprocal.
//
// The triangular solves divide every component by a diagonal entry of the
// factor (dtrsm), one division per component and right-hand side.  With
// dinv = 1.0 / d computed once per row (itself a correctly rounded
// division), q = RN(a dinv) is within an ulp of a / d, the remainder
// a - d q is exact as one fused multiply-add, and RN(q + r dinv) is the
// correctly rounded quotient (Markstein, "Computation of elementary
// functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34,
// 1990, theorem on quotients from a correctly rounded reciprocal).  The
// result equals a / d bit for bit while the quotient and the remainder stay
// in the normal range; a zero dividend returns itself (d > 0 here: d is a
// diagonal entry of a Cholesky factor), tiny, huge and non-finite
// quotients take the division itself.
#pragma once

#include <hip/hip_runtime.h>

namespace hpgl_amd {

__device__ inline double div_rn(double a, double d, double dinv) {
    const double q = a * dinv;
    const double r = fma(-q, d, a);
    const double c = fma(r, dinv, q);
    const double m = fabs(c);
    if (m > 0x1p-960 && m < 0x1p960) return c;
    if (a == 0.0) return a;
    return a / d;
}

}  // namespace hpgl_amd
