// Elementary functions of the fast arithmetic mode, written out so that they
// cost a fraction of the general library routines. Host and device builds
// run the same IEEE operations (frexp, correctly rounded add / multiply /
// divide, fused multiply-add), so a host build of this header gives the
// device's bits: the accuracy test compiles it with the host compiler.
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define PPES_HD __host__ __device__
#else
#define PPES_HD
#endif

namespace ppes {

// ln 2 split for the argument reduction (Cody and Waite): LN2_HEAD is ln 2
// rounded to 32 significant bits, so e * LN2_HEAD is exact for every binary
// exponent e of a double; LN2_TAIL = ln 2 - LN2_HEAD rounded to double.
constexpr double LN2_HEAD = 0.6931471806019545;       // 0x3FE62E42FF000000
constexpr double LN2_TAIL = -4.2009150726810846e-11;  // 0xBDC718432A1B0E26

// Natural logarithm, within about 1 ulp for positive finite x (normal or
// subnormal). x = 2^e m with m in [1/sqrt(2), sqrt(2)) (frexp, then one
// doubling) and f = m - 1 (exact). With s = f / (2 + f), z = s^2 < 0.02944,
//   ln m = 2 atanh(s) = 2s + s R(z),  R(z) = 2 (z/3 + z^2/5 + ... + z^9/19),
// and since 2s = f - s f = f - f^2/2 + s f^2/2,
//   ln m = f - (f^2/2 - s (f^2/2 + R(z))),
// so the exact f leads and s, rounded, only enters the small correction. The
// first omitted series term, 2 z^10 / 21 relative to 2s, is below 2.4e-17.
// ln x = e ln 2 + ln m, with e ln 2 split as above. Zero, negative, infinite
// and NaN arguments go to the library log (a branch no wavefront of valid
// states takes).
PPES_HD inline double ln_pos(double x)
{
    int e;
    double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752) {
        m = m + m;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 2.0 / 19.0;
    p = fma(p, z, 2.0 / 17.0);
    p = fma(p, z, 2.0 / 15.0);
    p = fma(p, z, 2.0 / 13.0);
    p = fma(p, z, 2.0 / 11.0);
    p = fma(p, z, 2.0 / 9.0);
    p = fma(p, z, 2.0 / 7.0);
    p = fma(p, z, 2.0 / 5.0);
    p = fma(p, z, 2.0 / 3.0);
    const double hfsq = 0.5 * f * f;
    const double lnm = f - fma(-s, fma(z, p, hfsq), hfsq);
    const double de = (double)e;
    double r = fma(de, LN2_HEAD, fma(de, LN2_TAIL, lnm));
    if (!(x > 0.0) || !(x < HUGE_VAL)) r = log(x);
    return r;
}

}  // namespace ppes
