// Scalar math shared by the host code and the gfx950 kernels.
//
// GLSL leaves exp/pow precision, float->int conversion of out-of-range values
// and min/max with NaN to the implementation. This file fixes each with a
// published method built only from IEEE-754 single operations (+ - * / sqrt,
// fused multiply-add, conversions), so a build without floating-point
// contraction gives the same bits on the CPU and on the GPU:
//
//  * apbf_exp  : Cody & Waite reduction x = k ln2 + r (ln2 split into a
//                high part with trailing zero bits and a low part), degree-7
//                Taylor polynomial of e^r by Horner with fma, times 2^k from the
//                exponent bits. x <= -86 gives 0 (no subnormal results), x > 88
//                gives +inf.
//  * apbf_pow  : exp(y ln x) in binary64: ln x from the exponent/mantissa split
//                and 2 atanh((m-1)/(m+1)) summed to t^21; exp by Cody & Waite
//                reduction and a degree-13 Taylor polynomial; one final rounding
//                to binary32. x < 0 gives NaN (GLSL: undefined).
//  * apbf_to_i32 / apbf_to_u32 : truncation toward zero, saturating, NaN -> 0.
//  * apbf_max / apbf_min : "b if a < b else a" (resp. b < a); a NaN first
//                operand yields the second operand.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define APBF_HD __host__ __device__ __forceinline__

namespace apbf {

APBF_HD float bits_f(uint32_t u) { return __builtin_bit_cast(float, u); }
APBF_HD double bits_d(uint64_t u) { return __builtin_bit_cast(double, u); }
APBF_HD uint64_t d_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

APBF_HD float fmax_g(float a, float b) { return (a < b) ? b : ((a == a) ? a : b); }
APBF_HD float fmin_g(float a, float b) { return (b < a) ? b : ((a == a) ? a : b); }

APBF_HD int32_t to_i32_ref(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}
APBF_HD uint32_t to_u32_ref(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
// On the device, gfx950's V_CVT_I32_F32 / V_CVT_U32_F32 truncate toward
// zero, saturate and map NaN to 0, which is the definition above in one
// instruction (the compiler cannot know that and keeps the range tests).
// Checked against to_i32_ref / to_u32_ref over all 2^32 inputs by
// apbf_check_conversions.
APBF_HD int32_t to_i32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
#else
    return to_i32_ref(x);
#endif
}
APBF_HD uint32_t to_u32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
#else
    return to_u32_ref(x);
#endif
}

// Largest float t with sqrt(t) <= r (IEEE square root), so that for every
// float d2 >= 0 the test sqrt(d2) <= r holds exactly when d2 <= t: a distance
// comparison without the square root. -inf when no d2 qualifies (r < 0, NaN).
// t starts at r * r, within a few units in the last place of the bound: each
// float r is the rounded root of at most a handful of consecutive floats.
APBF_HD float sqrt_le_bound(float r) {
    if (!(r >= 0.0f)) return -__builtin_inff();
    if (r == __builtin_inff()) return r;
    float t = r * r;
    for (int k = 0; k < 8 && __builtin_sqrtf(t) > r; ++k)
        t = bits_f(__builtin_bit_cast(uint32_t, t) - 1u);  // t > 0 here: sqrt(0) <= r
    for (int k = 0; k < 8; ++k) {
        const float u = bits_f(__builtin_bit_cast(uint32_t, t) + 1u);
        if (!(__builtin_sqrtf(u) <= r)) break;
        t = u;
    }
    return t;
}

// apbf_exp on an argument in (-86, 88]: Cody & Waite reduction and the
// degree-7 polynomial.
APBF_HD float exp_core(float x) {
    const float log2e = bits_f(0x3FB8AA3Bu);
    const float ln2_hi = bits_f(0x3F317200u);
    const float ln2_lo = bits_f(0x35BFBE8Eu);
    float k = __builtin_rintf(x * log2e);
    float r = __builtin_fmaf(-k, ln2_hi, x);
    r = __builtin_fmaf(-k, ln2_lo, r);
    float p = 1.98412698e-04f;
    p = __builtin_fmaf(p, r, 1.38888889e-03f);
    p = __builtin_fmaf(p, r, 8.33333377e-03f);
    p = __builtin_fmaf(p, r, 4.16666679e-02f);
    p = __builtin_fmaf(p, r, 1.66666672e-01f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    int ki = (int)k;
    return p * bits_f((uint32_t)(ki + 127) << 23);
}
APBF_HD float exp_f(float x) {
    if (x != x) return x;
    if (x <= -86.0f) return 0.0f;
    if (x > 88.0f) return __builtin_inff();
    return exp_core(x);
}

// IEEE square root and reciprocal for arguments in a known range. On the
// device the square root is v_sqrt_f32 with its one-ulp correction (the full
// sequence without its scaling of tiny arguments and its special cases) and
// the reciprocal v_rcp_f32 with one Newton step on fma; both give the IEEE
// bits for x = 0 or 2^-96 <= x <= 2^120 (square root) and for
// 2^-120 <= |x| <= 2^120 (reciprocal), checked on every such float by
// apbf_check_math. Callers keep their arguments in those ranges.
APBF_HD float sqrt_in_range(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = bits_f(__builtin_bit_cast(uint32_t, s) - 1u);
    const float sp = bits_f(__builtin_bit_cast(uint32_t, s) + 1u);
    const float rm = __builtin_fmaf(-sm, s, x);
    const float rp = __builtin_fmaf(-sp, s, x);
    const float r = (rm <= 0.0f) ? sm : s;
    return (rp > 0.0f) ? sp : r;
#else
    return __builtin_sqrtf(x);
#endif
}
APBF_HD float recip_in_range(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(e, r, r);
#else
    return 1.0f / x;
#endif
}

// x / b for a divisor 2^-40 <= |b| <= 2^40, given y = 1/b (IEEE, as from
// recip_in_range): q0 = x * y, then one correction by the remainder
// fma(-b, q0, x), which is exact (Markstein's correction step). The result
// is the IEEE quotient whenever |x| >= 2^-80: the steps are invariant under
// scaling x and b by powers of two inside those ranges (no intermediate
// underflows or overflows), and the quotient was compared with the IEEE one
// over every significand of x for 100 000 divisor significands and over every
// significand of b for 65 536 dividend significands, without a difference
// (apbf_check_math repeats 512 divisors on the device). For |x| < 2^-80 both
// quotients are below 2^-39 in magnitude. Three operations per quotient
// after the reciprocal against about eleven for a full IEEE division.
APBF_HD float quotient_by_recip(float x, float b, float y) {
    const float q0 = x * y;
    return __builtin_fmaf(__builtin_fmaf(-b, q0, x), y, q0);
}

APBF_HD double log_d(double x) {
    const double ln2_hi = bits_d(0x3FE62E42FEE00000ull);
    const double ln2_lo = bits_d(0x3DEA39EF35793C76ull);
    uint64_t u = d_bits(x);
    int e = (int)((u >> 52) & 0x7FF);
    if (e == 0) {
        x = x * bits_d(0x4350000000000000ull);
        u = d_bits(x);
        e = (int)((u >> 52) & 0x7FF) - 54;
    }
    e -= 1023;
    double m = bits_d((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double t = (m - 1.0) / (m + 1.0);
    double t2 = t * t;
    double s = 1.0 / 21.0;
    s = __builtin_fma(s, t2, 1.0 / 19.0);
    s = __builtin_fma(s, t2, 1.0 / 17.0);
    s = __builtin_fma(s, t2, 1.0 / 15.0);
    s = __builtin_fma(s, t2, 1.0 / 13.0);
    s = __builtin_fma(s, t2, 1.0 / 11.0);
    s = __builtin_fma(s, t2, 1.0 / 9.0);
    s = __builtin_fma(s, t2, 1.0 / 7.0);
    s = __builtin_fma(s, t2, 1.0 / 5.0);
    s = __builtin_fma(s, t2, 1.0 / 3.0);
    s = __builtin_fma(s, t2, 1.0);
    double lnm = 2.0 * t * s;
    double de = (double)e;
    return __builtin_fma(de, ln2_hi, __builtin_fma(de, ln2_lo, lnm));
}

APBF_HD double exp_d(double z) {
    const double inv_ln2 = bits_d(0x3FF71547652B82FEull);
    const double ln2_hi = bits_d(0x3FE62E42FEE00000ull);
    const double ln2_lo = bits_d(0x3DEA39EF35793C76ull);
    double k = __builtin_rint(z * inv_ln2);
    double r = __builtin_fma(-k, ln2_hi, z);
    r = __builtin_fma(-k, ln2_lo, r);
    double p = 1.0 / 6227020800.0;
    p = __builtin_fma(p, r, 1.0 / 479001600.0);
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    int ki = (int)k;
    return p * bits_d((uint64_t)(ki + 1023) << 52);
}

APBF_HD float pow_f(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x != x || y != y) return __builtin_nanf("");
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 1.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : __builtin_inff();
    if (x == __builtin_inff()) return (y > 0.0f) ? __builtin_inff() : 0.0f;
    double z = (double)y * log_d((double)x);
    if (z > 89.0) return __builtin_inff();
    if (z < -87.0) return 0.0f;
    return (float)exp_d(z);
}

// ------------------------------------------------------------------ vectors
struct v3f {
    float x, y, z;
};
APBF_HD v3f mk(float x, float y, float z) { return v3f{x, y, z}; }
APBF_HD v3f operator+(v3f a, v3f b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
APBF_HD v3f operator-(v3f a, v3f b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
APBF_HD v3f operator*(v3f a, v3f b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
APBF_HD v3f operator*(v3f a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
APBF_HD v3f operator/(v3f a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
APBF_HD v3f operator-(v3f a) { return mk(-a.x, -a.y, -a.z); }
APBF_HD float dot(v3f a, v3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
APBF_HD float length(v3f a) { return __builtin_sqrtf(dot(a, a)); }
APBF_HD v3f normalize(v3f a) { return a * (1.0f / length(a)); }
APBF_HD float fract(float x) { return x - __builtin_floorf(x); }
APBF_HD float sign_f(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
APBF_HD int32_t iadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
APBF_HD int32_t isub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
APBF_HD v3f vec_i(int32_t x, int32_t y, int32_t z) { return mk((float)x, (float)y, (float)z); }

}  // namespace apbf
