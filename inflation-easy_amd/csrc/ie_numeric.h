// ie_numeric.h -- correctly rounded division by a value with a known reciprocal.
//
// x / b rounded to nearest, given rb = RN(1 / b): q0 = RN(x rb) lies within one
// ulp of x / b, the residual x - b q0 is exact in one FMA, and one FMA correction
// q0 + r rb gives the correctly rounded quotient (Markstein's division theorem,
// for operands and quotients away from overflow and underflow).  Three
// operations against about ten for an FP64 division, with the same result, so
// code that must keep a reference's divisions bit for bit can divide by a
// per-run or per-interval constant this way.  Plain host/device inline code, so
// that the host test compiles the same arithmetic.
#pragma once

#include <math.h>

#ifndef IE_HD
#if defined(__HIPCC__)
#define IE_HD __host__ __device__
#else
#define IE_HD
#endif
#endif

namespace ie {

IE_HD inline double div_by(double x, double b, double rb)
{
    const double q = x * rb;
    const double r = fma(-q, b, x);
    return fma(r, rb, q);
}

}  // namespace ie
