/* The solver's quotient by an inverse mass through the reciprocal and one
 * correction (quotient_by_recip in apbf_amd/csrc/apbf_math.h: q0 = x * y,
 * q = fma(fma(-b, q0, x), y, q0) with y = 1 / b) against the IEEE quotient
 * x / b: every dividend significand for `argv[1]` divisors (all-ones and 1.5
 * significands first, then hashed ones), with the divisor's exponent spread
 * over -40..40 and the dividend's over -80..80. Prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char** argv) {
    const uint32_t nd = argc > 1 ? (uint32_t)atoi(argv[1]) : 64u;
    unsigned long long bad = 0, total = 0;
    for (uint32_t di = 0; di < nd; ++di) {
        uint32_t h = di * 0x9E3779B1u;
        h ^= h >> 15;
        h *= 0x85EBCA77u;
        h ^= h >> 13;
        const uint32_t bm = di == 0u ? 0x7FFFFFu : (di == 1u ? 0x400000u : (h & 0x7FFFFFu));
        const float b = bits_f(((127u + di % 81u - 40u) << 23) | bm);
        const float y = 1.0f / b;
        for (uint32_t m = 0; m < (1u << 23); ++m) {
            const uint32_t xe = 127u + (di * 7u + (m >> 11)) % 161u - 80u;
            const float x = bits_f(((m & 1u) << 31) | (xe << 23) | m);
            const float q0 = x * y;
            const float q = fmaf(fmaf(-b, q0, x), y, q0);
            bad += f_bits(q) != f_bits(x / b);
            ++total;
        }
    }
    printf("%llu %llu\n", bad, total);
    return 0;
}
