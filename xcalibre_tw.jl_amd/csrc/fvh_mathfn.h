// Deterministic elementary functions used by the k-omega wall functions.
//
// The same source is compiled for the host and for gfx950 with
// -ffp-contract=off; it uses only IEEE +, -, *, / and comparisons, so the
// device result is bit-identical to the host one.  Natural logarithm by the
// classical decomposition x = m * 2^e, m in [sqrt(1/2), sqrt(2)), and
//     log(m) = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ...),  s = (m - 1)/(m + 1),
// |s| <= 0.1716, summed to s^25 (truncation < 1e-19 relative), with ln 2 split
// into a head exactly representable times e and a tail.  Accuracy is about
// 1 ulp; it is a deterministic stand-in for the platform log, not a copy of any.
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define FVH_HD __host__ __device__
#else
#define FVH_HD
#endif

FVH_HD inline double fvh_jmax(double x, double y) {
    // Julia's max for floats: NaN propagates, max(-0.0, 0.0) = 0.0
    bool take_y = (y > x) || ((y == y) && (x == x) && (y == 0.0) && (x == 0.0) &&
                              (__builtin_signbit(x) && !__builtin_signbit(y)));
    if (x != x) return x;
    if (y != y) return y;
    return take_y ? y : x;
}

FVH_HD inline double fvh_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    int64_t e = 0;
    if (x < 2.2250738585072014e-308) {  // subnormal: scale by 2^54
        x = x * 18014398509481984.0;
        e = -54;
    }
    uint64_t bits;
    __builtin_memcpy(&bits, &x, sizeof bits);
    int64_t ex = (int64_t)((bits >> 52) & 0x7ff) - 1023;
    uint64_t mb = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    __builtin_memcpy(&m, &mb, sizeof m);  // m in [1, 2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        ex += 1;
    }
    e += ex;
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    // Horner form of 1 + s2/3 + s2^2/5 + ... + s2^12/25
    double p = 1.0 / 25.0;
    p = p * s2 + 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    const double lm = 2.0 * s + 2.0 * s * (s2 * p);
    const double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000, 11 trailing zero bits
    const double ln2_lo = 1.90821492927058770002e-10;  // ln 2 - ln2_hi
    const double de = (double)e;
    return (de * ln2_hi + (lm + de * ln2_lo));
}

// y_plus_laminar(E, kappa)  [src/RANSModels/RANSModels_kOmega.jl:200-203]
FVH_HD inline double fvh_y_plus_laminar(double E, double kappa) {
    double yL = 11.0;
    for (int i = 0; i < 10; ++i) yL = fvh_log(fvh_jmax(yL * E, 1.0)) / kappa;
    return yL;
}

// nut_wall(nu, yplus, kappa, E)  [src/RANSModels/RANSModels_kOmega.jl:220-222]
FVH_HD inline double fvh_nut_wall(double nu, double yplus, double kappa, double E) {
    return fvh_jmax(nu * (yplus * kappa / fvh_log(fvh_jmax(E * yplus, 1.0 + 1e-4)) - 1.0), 0.0);
}
