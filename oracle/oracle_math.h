// Test infrastructure only: the oracle's scalar floating-point helpers.
//
// The oracle restates APBF's GLSL shaders on the CPU. GLSL leaves exp/pow,
// float->int conversion of out-of-range values and min/max of NaN to the
// implementation, so the oracle fixes each of them here with a documented,
// published method:
//
//  * exp_f   : Cody & Waite argument reduction (x = k*ln2 + r, ln2 split in a
//              high part with trailing zero bits and a low part) followed by
//              the degree-7 Taylor polynomial of e^r evaluated by Horner with
//              fused multiply-adds, scaled by 2^k built from the exponent bits.
//              Arguments <= -86 return 0 (no subnormal results); > 88 returns +inf.
//  * pow_f   : pow(x, y) = exp(y * ln x) evaluated in binary64: ln x from the
//              exponent/mantissa split and 2*atanh((m-1)/(m+1)) (series to t^21),
//              exp by Cody & Waite reduction and a degree-13 Taylor polynomial;
//              the result is rounded once to binary32. Special cases follow C99
//              pow for x >= 0; x < 0 returns NaN (GLSL: undefined).
//  * to_i32 / to_u32 : truncation toward zero with saturation, NaN -> 0
//              (the behaviour of the v_cvt_{i32,u32}_f32 conversions the shaders
//              get on AMD hardware; C++ leaves these cases undefined).
//  * gmax / gmin : GLSL max/min as "b if a<b else a" (resp. b<a), with a NaN
//              first operand yielding the second operand.
//
// Every operation is a single IEEE-754 binary32/binary64 operation (+,-,*,/,
// sqrt, fma, conversions), so the results are reproducible on any IEEE
// platform when the file is compiled without floating-point contraction
// (-ffp-contract=off) and without fast-math.
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>

namespace om {

static inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline double bits_d(uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }
static inline uint64_t d_bits(double d) { uint64_t u; std::memcpy(&u, &d, 8); return u; }

static inline float gmax(float a, float b) { return (a < b) ? b : ((a == a) ? a : b); }
static inline float gmin(float a, float b) { return (b < a) ? b : ((a == a) ? a : b); }

static inline int32_t to_i32(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static inline uint32_t to_u32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// e^x in binary32 (Cody-Waite + Taylor degree 7).
static inline float exp_f(float x) {
    if (x != x) return x;
    if (x <= -86.0f) return 0.0f;
    if (x > 88.0f) return INFINITY;
    const float log2e  = bits_f(0x3FB8AA3Bu);   // 1.44269502
    const float ln2_hi = bits_f(0x3F317200u);   // 0.693145751953125
    const float ln2_lo = bits_f(0x35BFBE8Eu);   // 1.42860677e-06
    float k = rintf(x * log2e);
    float r = fmaf(-k, ln2_hi, x);
    r = fmaf(-k, ln2_lo, r);
    float p = 1.98412698e-04f;          // 1/7!
    p = fmaf(p, r, 1.38888889e-03f);    // 1/6!
    p = fmaf(p, r, 8.33333377e-03f);    // 1/5!
    p = fmaf(p, r, 4.16666679e-02f);    // 1/4!
    p = fmaf(p, r, 1.66666672e-01f);    // 1/3!
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int ki = (int)k;
    return p * bits_f((uint32_t)(ki + 127) << 23);
}

// ln x in binary64 for finite x > 0.
static inline double log_d(double x) {
    const double ln2_hi = bits_d(0x3FE62E42FEE00000ull);  // 6.93147180369123816490e-01
    const double ln2_lo = bits_d(0x3DEA39EF35793C76ull);  // 1.90821492927058770002e-10
    uint64_t u = d_bits(x);
    int e = (int)((u >> 52) & 0x7FF);
    if (e == 0) {  // subnormal: normalise first
        x = x * bits_d(0x4350000000000000ull);  // 2^54
        u = d_bits(x);
        e = (int)((u >> 52) & 0x7FF) - 54;
    }
    e -= 1023;
    double m = bits_d((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);  // [1,2)
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double t = (m - 1.0) / (m + 1.0);
    double t2 = t * t;
    double s = 1.0 / 21.0;
    s = fma(s, t2, 1.0 / 19.0);
    s = fma(s, t2, 1.0 / 17.0);
    s = fma(s, t2, 1.0 / 15.0);
    s = fma(s, t2, 1.0 / 13.0);
    s = fma(s, t2, 1.0 / 11.0);
    s = fma(s, t2, 1.0 / 9.0);
    s = fma(s, t2, 1.0 / 7.0);
    s = fma(s, t2, 1.0 / 5.0);
    s = fma(s, t2, 1.0 / 3.0);
    s = fma(s, t2, 1.0);
    double lnm = 2.0 * t * s;
    double de = (double)e;
    return fma(de, ln2_hi, fma(de, ln2_lo, lnm));
}

// e^z in binary64 for |z| < 700.
static inline double exp_d(double z) {
    const double inv_ln2 = bits_d(0x3FF71547652B82FEull);  // 1.44269504088896338700e+00
    const double ln2_hi  = bits_d(0x3FE62E42FEE00000ull);
    const double ln2_lo  = bits_d(0x3DEA39EF35793C76ull);
    double k = rint(z * inv_ln2);
    double r = fma(-k, ln2_hi, z);
    r = fma(-k, ln2_lo, r);
    double p = 1.0 / 6227020800.0;        // 1/13!
    p = fma(p, r, 1.0 / 479001600.0);     // 1/12!
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int ki = (int)k;
    return p * bits_d((uint64_t)(ki + 1023) << 52);
}

// pow(x, y) in binary32 through binary64.
static inline float pow_f(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x != x || y != y) return NAN;
    if (x < 0.0f) return NAN;
    if (x == 1.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : INFINITY;
    if (std::isinf(x)) return (y > 0.0f) ? INFINITY : 0.0f;
    double z = (double)y * log_d((double)x);
    if (z > 89.0) return INFINITY;
    if (z < -87.0) return 0.0f;
    return (float)exp_d(z);
}

}  // namespace om
