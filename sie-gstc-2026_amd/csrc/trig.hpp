// sin and cos of a moderate FP64 argument, for the oscillatory factor
// e^{-jkR} of the Green's function (kR is at most a few thousand here).
//
// Cody-Waite reduction by pi/2 (W. J. Cody and W. Waite, Software Manual for
// the Elementary Functions, 1980): pi/2 = P1 + P2 + P3, each a double, and
// r = ((x - n P1) - n P2) - n P3 with fused multiply-adds, so that the first
// step is exact to one rounding; then the Taylor series of sin to degree 17
// and of cos to degree 18 on |r| <= pi/4 (truncation error below 1e-19),
// evaluated by Horner's rule in r^2. Absolute error ~2e-16 for |x| < 2^20
// (checked against long-double sin/cos).
//
// sincos_tab: the same for the inner loop of the far-pair quadrature, from a
// table of e^{j 2 pi i / 2048} (2048 entries, filled on the host in long
// double by sincos_table_fill): reduction by pi/1024 in two parts, with
// t = x 1024/pi rounded to an integer by adding 1.5 2^52 (the sum's low bits
// are t, so the table index i = t mod 2048 needs no conversion),
// r = (x - t H1) - t H2 with |r| <= pi/2048, sin r = r - r^3/6 (truncation
// below 7.2e-17) and cos r - 1 = -r^2/2 + r^4/24 (below 2e-20); then
// sin x = s_i (1 + (cos r - 1)) + c_i sin r, cos x = c_i (1 + (cos r - 1)) - s_i sin r.
// 15 operations against about 45 for sincos_moderate.
#pragma once
#include <cmath>

// Qualifiers for the device translation units (__host__ __device__); empty
// for a plain C++ build of the same function.
#ifndef SIE_HD
#define SIE_HD
#endif

namespace sie {

SIE_HD inline void sincos_moderate(double x, double* s_out, double* c_out) {
    constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
    constexpr double P1 = 0x1.921fb54442d18p+0;
    constexpr double P2 = 0x1.1a62633145c07p-54;
    constexpr double P3 = -0x1.f1976b7ed8fbcp-110;
    const double n = rint(x * kTwoOverPi);
    double r = fma(-n, P1, x);
    r = fma(-n, P2, r);
    r = fma(-n, P3, r);
    const double z = r * r;
    // sin r = r + r z t(z),  t = -1/3! + z/5! - z^2/7! + ... + z^7/17!
    double t = 1.0 / 355687428096000.0;
    t = fma(t, z, -1.0 / 1307674368000.0);
    t = fma(t, z, 1.0 / 6227020800.0);
    t = fma(t, z, -1.0 / 39916800.0);
    t = fma(t, z, 1.0 / 362880.0);
    t = fma(t, z, -1.0 / 5040.0);
    t = fma(t, z, 1.0 / 120.0);
    t = fma(t, z, -1.0 / 6.0);
    const double sr = fma(r * z, t, r);
    // cos r = 1 + z w(z),  w = -1/2! + z/4! - z^2/6! + ... - z^8/18!
    double w = -1.0 / 6402373705728000.0;
    w = fma(w, z, 1.0 / 20922789888000.0);
    w = fma(w, z, -1.0 / 87178291200.0);
    w = fma(w, z, 1.0 / 479001600.0);
    w = fma(w, z, -1.0 / 3628800.0);
    w = fma(w, z, 1.0 / 40320.0);
    w = fma(w, z, -1.0 / 720.0);
    w = fma(w, z, 1.0 / 24.0);
    w = fma(w, z, -0.5);
    const double cr = fma(z, w, 1.0);
    // quadrant n mod 4: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s)
    const int q = (int)n & 3;  // |n| < 2^31 for |x| < 2^30
    const double a = (q & 1) ? cr : sr;
    const double b = (q & 1) ? sr : cr;
    *s_out = (q & 2) ? -a : a;
    *c_out = ((q + 1) & 2) ? -b : b;
}

constexpr int kSinCosTabN = 2048;

// tab[2 i] = cos(2 pi i / 2048), tab[2 i + 1] = sin(2 pi i / 2048), correctly
// rounded to double in practice (long-double evaluation). Host only.
inline void sincos_table_fill(double* tab) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < kSinCosTabN; ++i) {
        const long double a = 2.0L * pi * (long double)i / (long double)kSinCosTabN;
        tab[2 * i] = (double)cosl(a);
        tab[2 * i + 1] = (double)sinl(a);
    }
}

SIE_HD inline void sincos_tab(double x, const double* tab, double* s_out, double* c_out) {
    constexpr double kInvH = 0x1.45f306dc9c883p+8;  // 1024 / pi
    constexpr double H1 = 0x1.921fb54442d18p-9;     // pi/1024 = H1 + H2 to about 2^-119
    constexpr double H2 = 0x1.1a62633145c07p-63;
    constexpr double kRound = 0x1.8p+52;            // 1.5 2^52: sums round to integers
    // |x 1024/pi| < 2^51, so td - kRound = t exactly and the low word of td is
    // t modulo 2^32 (two's complement for t < 0)
    const double td = fma(x, kInvH, kRound);
    const double t = td - kRound;
    double r = fma(-t, H1, x);
    r = fma(-t, H2, r);
    const int i = (int)__builtin_bit_cast(unsigned long long, td) & (kSinCosTabN - 1);  // t mod 2048
    const double ci = tab[2 * i], si = tab[2 * i + 1];
    const double z = r * r;
    const double sr = fma(r * z, -1.0 / 6.0, r);
    const double cm1 = z * fma(z, 1.0 / 24.0, -0.5);
    *s_out = fma(ci, sr, fma(si, cm1, si));
    *c_out = fma(-si, sr, fma(ci, cm1, ci));
}

}  // namespace sie
