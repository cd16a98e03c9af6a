// Deterministic elementary functions used by the k-omega wall functions.
//
// The same source is compiled for the host and for gfx950 with
// -ffp-contract=off; it uses only IEEE +, -, *, / and comparisons, so the
// device result is bit-identical to the host one.  The natural logarithm is
// evaluated in double-double arithmetic and rounded once, which makes it the
// correctly rounded value outside extremely rare hard cases; it is written
// from the atanh series and Dekker's error-free products, not copied from
// any library.
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

// Double-double helpers (Dekker / Knuth error-free transformations; exact
// only without contraction to FMA, which the build forbids).
struct fvh_dd {
    double hi, lo;
};
FVH_HD inline fvh_dd fvh_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
FVH_HD inline fvh_dd fvh_fast_two_sum(double a, double b) {  // |a| >= |b|
    const double s = a + b;
    return {s, b - (s - a)};
}
FVH_HD inline fvh_dd fvh_two_prod(double a, double b) {
    const double p = a * b;
    const double ca = 134217729.0 * a, cb = 134217729.0 * b;  // 2^27 + 1: Veltkamp split
    const double ah = ca - (ca - a), al = a - ah;
    const double bh = cb - (cb - b), bl = b - bh;
    return {p, ((ah * bh - p) + ah * bl + al * bh) + al * bl};
}
FVH_HD inline fvh_dd fvh_dd_add(fvh_dd a, fvh_dd b) {
    fvh_dd s = fvh_two_sum(a.hi, b.hi);
    const double t = s.lo + (a.lo + b.lo);
    return fvh_fast_two_sum(s.hi, t);
}
FVH_HD inline fvh_dd fvh_dd_mul(fvh_dd a, fvh_dd b) {
    fvh_dd p = fvh_two_prod(a.hi, b.hi);
    const double t = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return fvh_fast_two_sum(p.hi, t);
}
FVH_HD inline fvh_dd fvh_dd_div(fvh_dd a, fvh_dd b) {
    const double q1 = a.hi / b.hi;
    fvh_dd r = fvh_dd_add(a, fvh_dd_mul({-q1, 0.0}, b));
    const double q2 = r.hi / b.hi;
    r = fvh_dd_add(r, fvh_dd_mul({-q2, 0.0}, b));
    const double q3 = r.hi / b.hi;
    return fvh_dd_add(fvh_fast_two_sum(q1, q2), {q3, 0.0});
}
// 1 / k as a double-double: hi the rounded quotient, lo the rounded rest.
FVH_HD inline fvh_dd fvh_dd_recip(double k) {
    const double hi = 1.0 / k;
    const fvh_dd p = fvh_two_prod(hi, k);  // hi * k = p.hi + p.lo exactly
    return {hi, ((1.0 - p.hi) - p.lo) / k};
}

// Natural logarithm.  x = 2^e m with m in [sqrt(1/2), sqrt(2)), and
//     log(m) = 2 atanh(s) = 2 (s + s^3/3 + ... + s^25/25),  s = (m - 1)/(m + 1),
// |s| <= 0.1716 (the omitted terms are below 2^-60 relative), evaluated in
// double-double arithmetic (about 100 significant bits), and e ln 2 added in
// double-double before the one final rounding.  The result is the correctly
// rounded logarithm except where the exact value lies within about 2^-45 ulp
// of a rounding boundary.
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
        m = m * 0.5;  // exact
        ex += 1;
    }
    e += ex;
    // s = (m - 1) / (m + 1): m - 1 is exact (Sterbenz), m + 1 exact as a double-double
    const fvh_dd s = fvh_dd_div({m - 1.0, 0.0}, fvh_two_sum(m, 1.0));
    const fvh_dd s2 = fvh_dd_mul(s, s);
    fvh_dd p = fvh_dd_recip(25.0);
    for (int k = 23; k >= 3; k -= 2) p = fvh_dd_add(fvh_dd_mul(p, s2), fvh_dd_recip((double)k));
    p = fvh_dd_add(fvh_dd_mul(p, s2), {1.0, 0.0});
    fvh_dd lm = fvh_dd_mul(s, p);
    lm = {2.0 * lm.hi, 2.0 * lm.lo};  // exact
    // ln 2 = L1 + L2 + L3: L1 has 11 trailing zero bits, so e * L1 is exact for |e| < 2048
    const double L1 = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
    const double L2 = 1.90821492927058770002e-10;  // 0x3DEA39EF35793C76
    const double L3 = 1.1612227229362532e-26;      // ln 2 - L1 - L2, rounded
    const double de = (double)e;
    const fvh_dd eln2 = fvh_dd_add({de * L1, 0.0}, fvh_dd_add(fvh_two_prod(de, L2), {de * L3, 0.0}));
    const fvh_dd r = fvh_dd_add(eln2, lm);
    return r.hi + r.lo;
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
