// gs_math.hpp - FP64 elementary functions for the trajectory kernels.
//
// The library versions of div, sqrt, log, exp, sin, cos round correctly or
// handle every range (Payne-Hanek reduction, denormal scaling); the step
// kernels only need ~1 ulp on known ranges.  These versions:
//   rcp, rsqrt, sqrt, hardware seed (v_rcp_f64 / v_rsq_f64 / v_sqrt_f64)
//   rsqrt4            + one third-order correction;
//   log               exponent split, s = (m-1)/(m+1), 2s + 2s w R(w);
//   exp               k ln2 + r reduction (two-part ln2), Taylor to r^13;
//   atan              three-point argument reduction + odd polynomial;
//   sincos_turn       sin/cos(2 pi x 2^-32) for a 32-bit integer x and
//   m2log_u32         -2 ln((x + 1) 2^-32): the Box-Muller angle and radius
//                     argument, and
//   sincos_tab        sin/cos of |x| < 2^20 (Cody-Waite pi/128 reduction),
//                     table-driven (P. T. P. Tang, "Table-driven
//                     implementation of the logarithm function in IEEE
//                     floating-point arithmetic", ACM TOMS 16(4), 1990: a
//                     table entry next to the argument and a short
//                     polynomial in the remainder; the same scheme for
//                     sin/cos).  The tables are gs_bm_tables.h; the kernels
//                     keep them in LDS.
// R, A, L are Chebyshev fits made by gen_poly.py (worst relative error of the
// function value 4.5e-18, 3.2e-19, 3.7e-17 before rounding); the sin/cos
// remainder series are Taylor's, exact to 1e-17 on |a| <= pi/256.
//
// Every multiply-add is an explicit fused multiply-add (Ops::fma3 / fmac: on
// the device v_fma_f64, on the host std::fma).  Both are correctly rounded,
// so the polynomial parts give the same bits on host and device; tests check
// both against mpmath through gs_host_math and gs_debug_math.  No inline
// assembly: the compiler's hazard recognizer must see every VALU instruction
// (a v_rcp_f64 / v_rsq_f64 result read by the next instruction needs wait
// states on gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "gs_bm_tables.h"

#pragma clang fp contract(off)

namespace gs {
namespace fm {

constexpr double kLn2Hi = 6.93147180369123816490e-01;  // upper 32 bits of ln 2
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kLn2Hi
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kTwoPiOver2p32 = 1.4629180792671596e-09;   // 2 pi 2^-32

__host__ __device__ inline double bits_to_double(uint64_t b) {
    union { uint64_t u; double d; } v;
    v.u = b;
    return v.d;
}
__host__ __device__ inline uint64_t double_to_bits(double d) {
    union { uint64_t u; double d; } v;
    v.d = d;
    return v.u;
}

// rcp, rsqrt and sqrt take the hardware seed (relative error e0 of about
// 2^-24) and apply ONE correction of third order instead of two Newton steps:
// with e = 1 - x y (or 1 - x y^2), the truncated series of (1 - e)^-1 and
// (1 - e)^-1/2 leave an error of order e0^3 < 2^-70, so the result's error is
// the final fused multiply-add's rounding plus at most 2^-54 from the rounded
// y^2 or x y: within one ulp, in 4 to 7 instructions.

// 1/x for normal x: y (1 + e + e^2), e = 1 - x y.
__device__ __forceinline__ double rcp(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, y, 1.0);
    return fma(y, fma(e, e, e), y);
}

// 1/sqrt(x) for normal x > 0: y (1 + e/2 + 3e^2/8), e = 1 - x y^2.
__device__ __forceinline__ double rsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x, y * y, 1.0);
    return fma(y, e * fma(0.375, e, 0.5), y);
}

// sqrt(x) for x >= 0 (0 -> 0; negative x: undefined): g = x y times the same
// series, e = 1 - g y.  The seed is taken at max(x, DBL_MIN), so x = 0 runs
// through with g = 0 and a finite seed: no select.
__device__ __forceinline__ double sqrt_nn(double x) {
    const double y = __builtin_amdgcn_rsq(fmax(x, 2.2250738585072014e-308));
    const double g = x * y;
    const double e = fma(-g, y, 1.0);
    return fma(g, e * fma(0.375, e, 0.5), g);
}

// x^(-1/4) for normal x > 0: the seed sqrt(rsq(x)) from two hardware
// seeds, then the series of (1 - e)^(-1/4) to third order, e = 1 - x q^4,
// written in f = e/4 = 1/4 - (x/4) q^4: q (1 + f + 5/2 f^2).  (Each fused
// multiply-add then has at most one non-inline constant.)
__device__ __forceinline__ double rsqrt4(double x) {
    const double q = __builtin_amdgcn_sqrt(__builtin_amdgcn_rsq(x));
    const double q2 = q * q;
    const double f = fma(-0.25 * x, q2 * q2, 0.25);
    return fma(q, f * fma(2.5, f, 1.0), q);
}

// sqrt(x), 0 for x <= 0.
__device__ __forceinline__ double sqrt_pos(double x) { return x > 0.0 ? sqrt_nn(x) : 0.0; }

// Ops policy: fma3 (general), fmac (addend a constant: Horner steps),
// inv (reciprocal).
struct DevOps {
    __device__ static double fma3(double a, double b, double c) { return fma(a, b, c); }
    __device__ static double fmac(double a, double b, double c) { return fma(a, b, c); }
    __device__ static double inv(double v) { return rcp(v); }
    __device__ static double root(double v) { return sqrt_nn(v); }
};
struct HostOps {
    static double fma3(double a, double b, double c) { return std::fma(a, b, c); }
    static double fmac(double a, double b, double c) { return std::fma(a, b, c); }
    static double inv(double v) { return 1.0 / v; }
    static double root(double v) { return std::sqrt(v); }
};

// ln((1+s)/(1-s)) = 2s + 2s w R(w), w = s^2 <= 0.0295.
template <class Ops>
__host__ __device__ inline double log_series(double s, double w) {
    double p = 0.07308224842521702930211;
    p = Ops::fmac(p, w, 0.07665860800278020634313);
    p = Ops::fmac(p, w, 0.09091444562630861040134);
    p = Ops::fmac(p, w, 0.1111110556739753992754);
    p = Ops::fmac(p, w, 0.1428571431298774246073);
    p = Ops::fmac(p, w, 0.199999999999497522449);
    p = Ops::fmac(p, w, 0.3333333333333334843082);
    const double s2 = 2.0 * s;
    return Ops::fma3(s2 * w, p, s2);
}

// ln(x 2^e_off) for a positive normal double x.  The power of two is folded
// into the exponent, so ln(2^k) is exactly k ln2 (0 for k = 0).
template <class Ops>
__host__ __device__ inline double log_e(double x, int e_off) {
    const uint64_t b = double_to_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023 + e_off;
    uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;  // m in [1, 2)
    if (mb > 0x3ff6a09e667f3bcdull) {   // m > sqrt(2): use m / 2
        mb -= 0x0010000000000000ull;
        e += 1;
    }
    const double m = bits_to_double(mb);
    const double s = (m - 1.0) * Ops::inv(m + 1.0);
    const double lm = log_series<Ops>(s, s * s);
    const double fe = (double)e;
    return Ops::fmac(fe, kLn2Hi, Ops::fma3(fe, kLn2Lo, lm));
}

// exp(x) for |x| < 700 (results are normal numbers).  Taylor 1/k!, k <= 13,
// on |r| <= ln2/2.
template <class Ops>
__host__ __device__ inline double exp_e(double x) {
    const double k = rint(x * kInvLn2);
    const double r = Ops::fma3(-k, kLn2Lo, Ops::fma3(-k, kLn2Hi, x));
    double p = 1.0 / 6227020800.0;
    p = Ops::fmac(p, r, 1.0 / 479001600.0);
    p = Ops::fmac(p, r, 1.0 / 39916800.0);
    p = Ops::fmac(p, r, 1.0 / 3628800.0);
    p = Ops::fmac(p, r, 1.0 / 362880.0);
    p = Ops::fmac(p, r, 1.0 / 40320.0);
    p = Ops::fmac(p, r, 1.0 / 5040.0);
    p = Ops::fmac(p, r, 1.0 / 720.0);
    p = Ops::fmac(p, r, 1.0 / 120.0);
    p = Ops::fmac(p, r, 1.0 / 24.0);
    p = Ops::fmac(p, r, 1.0 / 6.0);
    p = Ops::fmac(p, r, 0.5);
    p = Ops::fmac(p, r, 1.0);
    p = Ops::fmac(p, r, 1.0);
    const int64_t ki = (int64_t)k;
    return bits_to_double(double_to_bits(p) + ((uint64_t)ki << 52));
}

// atan t = t - t w A(w), w = t^2, |t| <= tan(pi/16) (A: gen_poly.py, 3.2e-19).
template <class Ops>
__host__ __device__ inline double atan_reduced(double t) {
    const double w = t * t;
    double p = -0.05115745967981170892;
    p = Ops::fmac(p, w, 0.06618732486766733679356);
    p = Ops::fmac(p, w, -0.07690724472468718292484);
    p = Ops::fmac(p, w, 0.09090879951934661083198);
    p = Ops::fmac(p, w, -0.1111111081899807722419);
    p = Ops::fmac(p, w, 0.1428571428427355230372);
    p = Ops::fmac(p, w, -0.1999999999999728936999);
    p = Ops::fmac(p, w, 0.3333333333333333249568);
    return Ops::fma3(-(t * w), p, t);
}

constexpr double kTanPi16 = 0.19891236737965801;
constexpr double kTan3Pi16 = 0.66817863791929892;
constexpr double kTanPi8 = 0.41421356237309503;
constexpr double kPi8Hi = 0.39269908169872414, kPi8Lo = 1.5308084989341915e-17;   // pi/8
constexpr double kPi4Hi = 0.78539816339744828, kPi4Lo = 3.061616997868383e-17;    // pi/4
constexpr double kPi2Hi = 1.5707963267948966, kPi2Lo = 6.123233995736766e-17;     // pi/2

// A 16-byte-aligned pair of table doubles, read as one 128-bit load.
struct alignas(16) Pair {
    double a, b;
};
__host__ __device__ inline Pair load_pair(const double *p) {
    Pair v;
    __builtin_memcpy(&v, __builtin_assume_aligned(p, 16), sizeof v);
    return v;
}

// atan x.  |x| > 1 uses pi/2 - atan(1/|x|); then a in [0, 1] is moved to
// t = (a - c) / (1 + c a) with c in {0, tan(pi/8), 1}, so that
// |t| <= tan(pi/16), and atan a = atan c + atan t.
template <class Ops>
__host__ __device__ inline double atan_e(double x) {
    const double ax = fabs(x);
    const bool big = ax > 1.0;
    const double a = big ? (ax < INFINITY ? Ops::inv(ax) : 0.0) : ax;   // device rcp(inf) is NaN
    const bool low = a < kTanPi16;
    const bool high = a >= kTan3Pi16;
    const double cp = high ? 1.0 : (low ? 0.0 : kTanPi8);
    const double base_hi = high ? kPi4Hi : (low ? 0.0 : kPi8Hi);
    const double base_lo = high ? kPi4Lo : (low ? 0.0 : kPi8Lo);
    const double t = low ? a : (a - cp) * Ops::inv(Ops::fma3(cp, a, 1.0));
    double r = atan_reduced<Ops>(t);
    r = base_hi + (r + base_lo);
    if (big) r = kPi2Hi - (r - kPi2Lo);
    return x < 0.0 ? -r : r;
}

// The three reduction intervals of atan_frac as a table: (c, atan c hi,
// lo, 0) for c = 0, tan(pi/8), 1.  The kernels keep a copy in LDS after
// the sin/cos table; a lane reads its interval's entry instead of selecting
// three constants out of three.
constexpr int kAtanRedWords = 16;   // 12 used; 16 keeps the next table 16-byte aligned
constexpr double kAtanRed[kAtanRedWords] = {0.0,     0.0,    0.0,    0.0, kTanPi8, kPi8Hi, kPi8Lo, 0.0,
                                            1.0,     kPi4Hi, kPi4Lo, 0.0, 0.0,     0.0,    0.0,    0.0};

// atan(n / d) for 0 <= n <= d, d > 0, with a single reciprocal: the same
// reduction in fraction form, t = (n - c d) / (d + c n).  The interval tests
// compare n with c d (no quotient); near a boundary either side's t stays
// within the polynomial's range.  `red` is kAtanRed (or a copy of it).
template <class Ops>
__host__ __device__ inline double atan_frac(double n, double d, const double *red) {
    const int k = (n >= kTanPi16 * d ? 1 : 0) + (n >= kTan3Pi16 * d ? 1 : 0);
    const Pair cb = load_pair(red + 4 * k);   // c, atan c hi
    const double base_lo = red[4 * k + 2];
    const double t = Ops::fma3(-cb.a, d, n) * Ops::inv(Ops::fma3(cb.a, n, d));
    return cb.b + (atan_reduced<Ops>(t) + base_lo);
}

// Neutral-sheet drift factors of the 2D models (SolarPropLikeCpuModel.cpp:
// 183-209, GeliosphereCpuModel.cpp:335-361).  With x = sy / (pi - sy),
// alphaH = acos x, arg = k tan(alphaH), k = 1 - 2 theta / pi:
//   f = atan(arg) / alphaH,  g = tan(alphaH) / (alphaH (1 + arg^2)).
// In X = sy and Y = sqrt(pi (pi - 2 sy)) = sqrt(1 - x^2) (pi - sy), with
// Y >= 1.86 > |X| for |sy| <= 1: alphaH = pi/2 - atan(X / Y),
// tan(alphaH) = Y / X, atan(arg) = atan(k Y / X) and
// tan(alphaH) / (1 + arg^2) = X Y / (X^2 + k^2 Y^2).  None of the reference's
// quotients (x, tan, arg, 1/alphaH, 1/(1 + arg^2)) is formed on its own: two
// atan_frac reciprocals and one of alphaH (X^2 + k^2 Y^2) for both results.
// At sy = 0 this gives the limits f = +-(pi/2)/alphaH, g = 0.
template <class Ops>
__host__ __device__ inline void sheet_drift(double sy, double k, const double *red, double &f, double &g) {
    constexpr double kPiD = 3.141592653589793;
    const double X = sy, ax = fabs(sy);
    const double Y = Ops::root(kPiD * (kPiD - 2.0 * sy));
    const double a0 = atan_frac<Ops>(ax, Y, red);
    const double aH = kPi2Hi - (copysign(a0, X) - kPi2Lo);   // a0 >= 0 takes X's sign
    const double kY = fabs(k) * Y;
    const bool swap = kY > ax;
    const double a1 = atan_frac<Ops>(fmin(ax, kY), fmax(ax, kY), red);   // the smaller over the larger
    const double at = swap ? kPi2Hi - (a1 - kPi2Lo) : a1;     // atan(|k| Y / |X|)
    const double D2 = Ops::fma3(X, X, kY * kY);
    const double R = Ops::inv(aH * D2);
    f = ((k < 0.0) != (X < 0.0) ? -at : at) * (D2 * R);      // atan(arg) / alphaH
    g = (X * Y) * R;                                          // (X Y / D2) / alphaH
}

// ------------------------------------------------------------ Box-Muller
constexpr int kBmTrigWords = 512;   // (sin, cos)(2 pi j/256), j = 0..255
constexpr int kBmLogWords = 512;    // 128 x (-2 invc, 2 ln invc hi, lo, 0)
constexpr uint64_t kLogOff = 0x3fe5f00000000000ull;   // bits of 0.685546875

// sin/cos(A + a) from a table entry sA = sin A, cA = cos A and |a| <= pi/256
// (w = a^2): sin a = a + a w (-1/6 + w/120), cos a - 1 = w (-1/2 + w/24 -
// w^2/720) (next terms below 1e-17), then sin(A + a) = sA + (sA (cos a - 1) +
// cA sin a) and cos(A + a) = cA + (cA (cos a - 1) - sA sin a).
template <class Ops>
__host__ __device__ inline void sincos_entry(double a, double sA, double cA, double &s, double &c) {
    const double w = a * a;
    const double sa = Ops::fma3(a * w, Ops::fmac(1.0 / 120.0, w, -1.0 / 6.0), a);
    const double cm = w * Ops::fmac(Ops::fmac(-1.0 / 720.0, w, 1.0 / 24.0), w, -0.5);
    s = Ops::fma3(cA, sa, Ops::fma3(sA, cm, sA));
    c = Ops::fma3(-sA, sa, Ops::fma3(cA, cm, cA));
}

// sin/cos(2 pi x 2^-32): j = round(x 2^-24) mod 256, d = x - j 2^24 in
// [-2^23, 2^23) exactly, a = d 2 pi 2^-32, entry j of `trig`.
template <class Ops>
__host__ __device__ inline void sincos_turn(uint32_t x, const double *trig, double &s, double &c) {
    const uint32_t xj = x + 0x800000u;
    const uint32_t j = xj >> 24;
    const int32_t d = (int32_t)(x - (xj & 0xff000000u));
    const Pair t = load_pair(trig + 2 * j);
    sincos_entry<Ops>((double)d * kTwoPiOver2p32, t.a, t.b, s, c);
}

// sin and cos of x for |x| < 2^20 from the same table: k = rint(x 128/pi),
// a = x - k pi/128 with pi/128 in two parts (the first has 27 significant
// bits, so k times it is exact for |k| < 2^26), |a| <= pi/256, entry k mod 256.
constexpr double k128OverPi = 40.74366543152521;
constexpr double kPi128Hi = 0x1.921fb54p-6;
constexpr double kPi128Lo = 0x1.10b4611a62633p-36;
template <class Ops>
__host__ __device__ inline void sincos_tab(double x, const double *trig, double &s, double &c) {
    const double k = rint(x * k128OverPi);
    const double a = Ops::fma3(-k, kPi128Lo, Ops::fma3(-k, kPi128Hi, x));
    const uint32_t j = (uint32_t)(int32_t)k & 255u;
    const Pair t = load_pair(trig + 2 * j);
    sincos_entry<Ops>(a, t.a, t.b, s, c);
}

// -2 ln((x + 1) 2^-32) for a 32-bit integer x, in [0, 44.37]: the
// Box-Muller radius squared.  v = x + 1 (exact) = 2^e z: tmp = bits(v) -
// kLogOff, e = tmp >> 52 (0 .. 32), z = bits(v) - e 2^52 in [0.6855, 1.3711),
// interval i = bits 45..51 of tmp.  Since kLogOff's low word is zero, only
// the high words take part.  ln v = e ln2 + ln(1/invc_i) + ln(1 + r) with
// r = z invc_i - 1, |r| <= 2^-8, and ln(1 + r) = r + r^2 L(r).  Every term
// is formed times -2: the table holds -2 invc and 2 ln(invc) (hi, lo), r2 =
// z (-2 invc) + 2 = -2 r, and -2 r^2 L(r) = r2^2 M(r2) with M's coefficients
// those of L times (-1)^(k+1) 2^(-k-1).  Scaling by a power of two commutes
// with rounding, so the result is -2 times the plain evaluation bit for bit,
// without a final multiply.  The interval around 1 has invc = 1 and
// ln(1/invc) = 0, so near v = 2^32 (e = 32) the result is -2 ln(1 + r) with
// r2 = 2 - 2z exact: relative accuracy where the result is tiny.
constexpr double kM2Ln2Hi = -2.0 * kLn2Hi;
constexpr double kM2Ln2Lo = -2.0 * kLn2Lo;

template <class Ops>
__host__ __device__ inline double m2log_u32(uint32_t x, const double *logt) {
    const uint64_t ix = double_to_bits((double)x + 1.0);   // (double)x + 1 is exact
    const uint32_t hi = (uint32_t)(ix >> 32);
    // the high word of tmp - 32 2^52: its exponent field is e - 32 directly,
    // and bits 0..19 (z's mantissa bits and i) are those of tmp
    const uint32_t thi = hi - (uint32_t)((kLogOff + (32ull << 52)) >> 32);
    const int e32 = (int32_t)thi >> 20;
    const uint32_t i = (thi >> 13) & 127u;
    const uint32_t zhi = (thi & 0xfffffu) + (uint32_t)(kLogOff >> 32);
    const double z = bits_to_double(((uint64_t)zhi << 32) | (ix & 0xffffffffull));
    const Pair ent = load_pair(logt + 4 * i);   // -2 invc, 2 ln(invc) hi
    const double l2_lo = logt[4 * i + 2];
    const double r2 = Ops::fma3(z, ent.a, 2.0);
    double p = -0.1666691666986670666717 * (-1.0 / 32.0);
    p = Ops::fmac(p, r2, 0.2000028571784131347818 * (1.0 / 16.0));
    p = Ops::fmac(p, r2, -0.2499999999899998399979 * (-1.0 / 8.0));
    p = Ops::fmac(p, r2, 0.3333333333219045841247 * (1.0 / 4.0));
    p = Ops::fmac(p, r2, -0.5 * (-1.0 / 2.0));
    const double lp = Ops::fma3(r2 * r2, p, r2);
    const double fe = (double)e32;
    return Ops::fmac(fe, kM2Ln2Hi, ent.b) + (lp + Ops::fmac(fe, kM2Ln2Lo, l2_lo));
}

// ln((x + 1) 2^-32) in [-22.18, 0]: -1/2 of m2log_u32, exactly.
template <class Ops>
__host__ __device__ inline double log_u32(uint32_t x, const double *logt) {
    return -0.5 * m2log_u32<Ops>(x, logt);
}

// The device forms used by the kernels.
__device__ __forceinline__ double log_fast(double x) { return log_e<DevOps>(x, 0); }
__device__ __forceinline__ double exp_fast(double x) { return exp_e<DevOps>(x); }
__device__ __forceinline__ double log_u32d(uint32_t x, const double *logt) { return log_u32<DevOps>(x, logt); }
__device__ __forceinline__ double m2log_u32d(uint32_t x, const double *logt) { return m2log_u32<DevOps>(x, logt); }
__device__ __forceinline__ void sincos_turnd(uint32_t x, const double *trig, double &s, double &c) {
    sincos_turn<DevOps>(x, trig, s, c);
}
__device__ __forceinline__ void sincosd(double x, const double *trig, double &s, double &c) {
    sincos_tab<DevOps>(x, trig, s, c);
}
__device__ __forceinline__ double sind(double x, const double *trig) {
    double s, c;
    sincos_tab<DevOps>(x, trig, s, c);
    return s;
}
__device__ __forceinline__ double atand(double x) { return atan_e<DevOps>(x); }
__device__ __forceinline__ double atan_fracd(double n, double d, const double *red) {
    return atan_frac<DevOps>(n, d, red);
}

// 1/a, 1/b, 1/c, 1/d from one reciprocal of abcd (a few ulp each; the
// product must stay a normal number).
__device__ __forceinline__ void rcp4(double a, double b, double c, double d, double &ia, double &ib, double &ic,
                                     double &id) {
    const double ab = a * b, abc = ab * c;
    const double R = rcp(abc * d);
    id = abc * R;
    const double t = d * R;    // 1/(abc)
    ic = ab * t;
    const double u = c * t;    // 1/(ab)
    ib = a * u;
    ia = b * u;
}

}  // namespace fm
}  // namespace gs
