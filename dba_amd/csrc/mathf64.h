// mathf64.h - FP64 natural logarithm for the Model B chemical potentials.
//
// The device library's log carries double-double corrections (about 90 VALU
// instructions on gfx950); the potentials take four logarithms per cell, which
// made them the largest single cost of the Model B step. This one follows the
// classical argument reduction
//
//     x = 2^k (1 + f),  sqrt(2)/2 <= 1 + f < sqrt(2),  s = f / (2 + f),
//     log(1 + f) = 2 atanh(s) = 2s + 2s^3/3 + 2s^5/5 + ...
//               = f - s f + s R,   R = s^2 (2/3 + 2s^2/5 + ... + 2s^18/21),
//
// with |s| <= 0.1716 (s^2 <= 0.0295), so the series truncated after s^21 leaves a
// relative error below 2^-56. The result is assembled as
//     k ln2_hi - ((hf - (s (hf + R) + k ln2_lo)) - f),  hf = f^2 / 2,
// (f - s f = f - hf + s hf) so that the large terms are added last; k ln2_hi is
// exact for |k| < 2^11. The error stays within 1 ulp of the exact logarithm
// (checked against the C library on the host by the tests, where the same
// source compiles as plain C++). Special cases follow IEEE log: NaN for x < 0
// and NaN input, -inf at +-0, +inf at +inf.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define DBA_HD __host__ __device__
#else
#define DBA_HD
#endif

namespace dba {

DBA_HD inline double log_f64(double x) {
    // straight-line code: the special cases are applied by one select at the end
    int k;
    double m = std::frexp(x, &k);  // m in [0.5, 1), exact for subnormal x too
    const bool low = m < 0x1.6a09e667f3bcdp-1;  // sqrt(2)/2
    m = low ? m + m : m;
    k = low ? k - 1 : k;
    const double f = m - 1.0;  // exact (Sterbenz)
    const double s = f / (2.0 + f);
    const double z = s * s;
    double r = 2.0 / 21.0;
    r = std::fma(r, z, 2.0 / 19.0);
    r = std::fma(r, z, 2.0 / 17.0);
    r = std::fma(r, z, 2.0 / 15.0);
    r = std::fma(r, z, 2.0 / 13.0);
    r = std::fma(r, z, 2.0 / 11.0);
    r = std::fma(r, z, 2.0 / 9.0);
    r = std::fma(r, z, 2.0 / 7.0);
    r = std::fma(r, z, 2.0 / 5.0);
    r = std::fma(r, z, 2.0 / 3.0);
    const double R = z * r;
    const double hf = 0.5 * f * f;
    const double dk = (double)k;
    // ln 2 = ln2_hi + ln2_lo, ln2_hi with 21 trailing zero bits
    const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
    const double v = dk * ln2_hi - ((hf - (s * (hf + R) + dk * ln2_lo)) - f);
    const bool finite_pos = x > 0.0 && x < INFINITY;
    return finite_pos ? v : (x == 0.0 ? -INFINITY : (x == INFINITY ? x : NAN));
}

}  // namespace dba
