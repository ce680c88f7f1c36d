// SPH smoothing kernels of the reference (shaders/kernels.glsl:3-123). `dim`
// is DIMENSIONS (cpu_gpu_shared_config.h:2), 2.0f or 3.0f; the cubic, poly6
// and spiky kernels carry their 3D constants in either (as the reference).
//
// The reference evaluates kernel_height / kernel_gradient per pair and
// recomputes the Gaussian constants from the kernel width on every call; they
// depend on the width only, so the kernels here take them precomputed per
// particle (GaussConst), which yields the same values.
#pragma once
#include "apbf_math.h"

namespace apbf {

struct GaussConst {
    float ivd;   // pow(height, 2/D) * PI       (kernels.glsl:86,91)
    float norm;  // pow(pow(height, 2/D), D/2)  (kernels.glsl:88)
};

APBF_HD GaussConst gauss_const(float kw, float dim) {
    const float pi = 3.14159265f;
    float height = 0.6f / pow_f(kw / 2.0f, dim);  // kernels.glsl:104,117
    float ivdwp = pow_f(height, 2.0f / dim);
    return GaussConst{ivdwp * pi, pow_f(ivdwp, dim / 2.0f)};
}

APBF_HD float gauss_height(v3f r, GaussConst g) {  // kernels.glsl:84-89
    return exp_f(-dot(r, r) * g.ivd) * g.norm;
}
// kernels.glsl:91-97. r is a difference of fixed-point positions over 2^18,
// so dot(r, r) is 0 or lies in [2^-36, 2^31], and the length is >= 1e-4 when
// it is inverted: the in-range square root and reciprocal give the IEEE
// results of length(r) and normalize(r) = r * (1 / length(r)).
APBF_HD v3f gauss_gradient(v3f r, GaussConst g, float height_value) {
    const float dist = sqrt_in_range(dot(r, r));
    if (dist < 0.0001f) return mk(0, 0, 0);
    const float s = -height_value * 2.0f * dist * g.ivd;
    return (r * recip_in_range(dist)) * s;
}

APBF_HD float poly6_height(v3f r, float h) {  // kernels.glsl:3-9
    const float pi = 3.14159265f;
    float r2 = dot(r, r), h2 = h * h;
    if (r2 > h2) return 0.0f;
    return 315.0f * pow_f(h2 - r2, 3.0f) / (64.0f * pi * pow_f(h, 9.0f));
}
APBF_HD v3f spiky_gradient(v3f r, float h) {  // kernels.glsl:11-16
    const float pi = 3.14159265f;
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return mk(0, 0, 0);
    float s = -45.0f * pow_f(h - dist, 2.0f) / (pi * pow_f(h, 6.0f));
    return (r / dist) * s;
}
APBF_HD float cubic_height(v3f r, float h) {  // kernels.glsl:18-31
    const float pi = 3.14159265f;
    float dist = length(r);
    if (dist > h) return 0.0f;
    float q = dist / h;
    float k = 8.0f / (pi * h * h * h);
    if (q <= 0.5f) {
        float q2 = q * q, q3 = q * q2;
        return k * (6.0f * q3 - 6.0f * q2 + 1.0f);
    }
    return k * (2.0f * pow_f(1.0f - q, 3.0f));
}
APBF_HD v3f cubic_gradient(v3f r, float h) {  // kernels.glsl:33-46
    const float pi = 3.14159265f;
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return mk(0, 0, 0);
    float q = dist / h;
    v3f gradq = r * (1.0f / (dist * h));
    float l = 48.0f / (pi * h * h * h);
    if (q <= 0.5f) return gradq * (l * q * (3.0f * q - 2.0f));
    float factor = 1.0f - q;
    return gradq * (l * (-factor * factor));
}
APBF_HD float cone_height(v3f r, float h, float dim) {  // kernels.glsl:48-52
    const float pi = 3.14159265f;
    float height = 3.0f / (pi * pow_f(h, dim));
    return fmax_g(0.0f, (1.0f - length(r) / h) * height);
}
APBF_HD v3f cone_gradient(v3f r, float h, float dim) {  // kernels.glsl:54-60
    const float pi = 3.14159265f;
    float steep = 3.0f / (pi * pow_f(h, dim + 1.0f));
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return mk(0, 0, 0);
    return (-r) * (steep / dist);
}
// kernels.glsl:64-68,76-80: the 3D and 2D normalisations
APBF_HD float quadratic_spike_a(float h, float dim) {
    const float pi = 3.14159265f;
    return dim == 3.0f ? 15.0f / (2.0f * pi * pow_f(h, 5.0f)) : 6.0f / (pi * pow_f(h, 4.0f));
}
APBF_HD float quadratic_spike_height(v3f r, float h, float dim) {  // kernels.glsl:62-70
    float a = quadratic_spike_a(h, dim);
    return a * pow_f(fmin_g(0.0f, length(r) - h), 2.0f);
}
APBF_HD v3f quadratic_spike_gradient(v3f r, float h, float dim) {  // kernels.glsl:72-82
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return mk(0, 0, 0);
    float a = quadratic_spike_a(h, dim);
    return r * (-2.0f * a * fmax_g(0.0f, h - dist) / dist);
}

// kernel_height / kernel_gradient dispatch (kernels.glsl:99-123). Returns W
// and/or grad W for the pair offset r and kernel width kw.
APBF_HD void kernel_eval(int hid, int gid, float dim, v3f r, float kw, GaussConst g, bool need_w, bool need_g,
                         float* w, v3f* grad) {
    float gh = 0.0f;
    bool have_gh = false;
    if (need_w) {
        if (hid == 1) {
            gh = gauss_height(r, g);
            have_gh = true;
            *w = gh;
        } else if (hid == 0) {
            *w = cubic_height(r, kw);
        } else if (hid == 2) {
            *w = poly6_height(r, kw);
        } else if (hid == 3) {
            *w = cone_height(r, kw, dim);
        } else if (hid == 4) {
            *w = quadratic_spike_height(r, kw, dim);
        } else {
            *w = 0.0f;
        }
    }
    if (need_g) {
        if (gid == 1) {
            if (!have_gh) gh = gauss_height(r, g);
            *grad = gauss_gradient(r, g, gh);
        } else if (gid == 0) {
            *grad = cubic_gradient(r, kw);
        } else if (gid == 2) {
            *grad = spiky_gradient(r, kw);
        } else if (gid == 3) {
            *grad = cone_gradient(r, kw, dim);
        } else if (gid == 4) {
            *grad = quadratic_spike_gradient(r, kw, dim);
        } else {
            *grad = mk(0, 0, 0);
        }
    }
}

}  // namespace apbf
