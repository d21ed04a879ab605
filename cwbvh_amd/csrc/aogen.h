// AO ray generation from surfels, one definition for host and device.
//
// Semantics of the reference's generator (src/ValidationKernels.cu:41-101):
// for surfel i (position p, normal n, texel radius r) and sample k,
//   t1 = n x (0,0,1), or n x (0,1,0) when |t1| < 0.1;  t1 = normalize(t1)
//   t2 = normalize(t1 x n)
//   (x, y) = concentric map of (u1, u2) onto the unit disk,  z = sqrt(1 - x^2 - y^2)
//   dir = normalize(t1 x + t2 y + n z),  origin = p + (n r) 0.5
//
// Here the arithmetic is fixed so that the host (g++) and the device (hipcc)
// produce the same bits: only IEEE +, -, *, / and sqrt (both compilers keep
// one rounding per operation: -ffp-contract=off, correctly rounded division
// and square root), explicit fmaf in the polynomials, and a counter-based
// SplitMix64 stream instead of curand (u1, u2 = the two 24-bit halves of one
// hash of (seed, ray index)).  sin and cos are evaluated on [-pi/4, pi/4]
// only (the second branch of the concentric map uses
// cos(pi/2 - phi) = sin(phi)), by their Taylor series to x^9 / x^10, whose
// truncation error there is below 1e-8.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define CWAO_HD __host__ __device__ __forceinline__
#define CWAO_FMA(a, b, c) __builtin_fmaf(a, b, c)
#define CWAO_SQRT(x) __builtin_sqrtf(x)
#else
#include <cmath>
#define CWAO_HD inline
#define CWAO_FMA(a, b, c) std::fma(a, b, c)
#define CWAO_SQRT(x) std::sqrt(x)
#endif

namespace cwao {

struct F3 {
    float x, y, z;
};

CWAO_HD uint64_t mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// two uniforms in [0, 1) with 24 random bits each
CWAO_HD void uniforms(uint64_t seed, uint64_t ray, float& u1, float& u2) {
    const uint64_t h = mix(seed ^ mix(ray));
    u1 = (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
    u2 = (float)(uint32_t)((h >> 16) & 0xFFFFFFu) * (1.0f / 16777216.0f);
}

// sin and cos of x in [-pi/4, pi/4]
CWAO_HD void sincos_quarter(float x, float& s, float& c) {
    const float x2 = x * x;
    float ps = CWAO_FMA(2.7557319e-6f, x2, -1.9841270e-4f);
    ps = CWAO_FMA(ps, x2, 8.3333333e-3f);
    ps = CWAO_FMA(ps, x2, -1.6666667e-1f);
    s = CWAO_FMA(x * x2, ps, x);
    float pc = CWAO_FMA(-2.7557319e-7f, x2, 2.4801587e-5f);
    pc = CWAO_FMA(pc, x2, -1.3888889e-3f);
    pc = CWAO_FMA(pc, x2, 4.1666667e-2f);
    pc = CWAO_FMA(pc, x2, -0.5f);
    c = CWAO_FMA(pc, x2, 1.0f);
}

CWAO_HD F3 cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

CWAO_HD float length(F3 a) { return CWAO_SQRT(a.x * a.x + a.y * a.y + a.z * a.z); }

CWAO_HD F3 normalize(F3 a) {
    const float l = length(a);
    if (!(l > 0.0f)) return a;
    const float inv = 1.0f / l;
    return F3{a.x * inv, a.y * inv, a.z * inv};
}

// Concentric square-to-disk map (Shirley-Chiu), lifted to the hemisphere.
CWAO_HD F3 cosine_hemisphere(float u1, float u2) {
    const float a = 2.0f * u1 - 1.0f, b = 2.0f * u2 - 1.0f;
    float x = 0.0f, y = 0.0f;
    if (a != 0.0f || b != 0.0f) {
        const float quarter_pi = 0.78539816f;
        float s, c;
        if (a * a > b * b) {
            sincos_quarter(quarter_pi * (b / a), s, c);
            x = a * c;
            y = a * s;
        } else {
            // theta = pi/2 - phi: cos(theta) = sin(phi), sin(theta) = cos(phi)
            sincos_quarter(quarter_pi * (a / b), s, c);
            x = b * s;
            y = b * c;
        }
    }
    const float zz = 1.0f - x * x - y * y;
    return F3{x, y, zz > 0.0f ? CWAO_SQRT(zz) : 0.0f};
}

// Ray `k` of surfel (p, n, r); writes origin and direction.
CWAO_HD void surfel_ray(F3 p, F3 n, float r, uint64_t seed, uint64_t ray, F3& o, F3& d) {
    F3 t1 = cross(n, F3{0.0f, 0.0f, 1.0f});
    if (length(t1) < 0.1f) t1 = cross(n, F3{0.0f, 1.0f, 0.0f});
    t1 = normalize(t1);
    const F3 t2 = normalize(cross(t1, n));
    float u1, u2;
    uniforms(seed, ray, u1, u2);
    const F3 l = cosine_hemisphere(u1, u2);
    d = normalize(F3{t1.x * l.x + t2.x * l.y + n.x * l.z, t1.y * l.x + t2.y * l.y + n.y * l.z,
                     t1.z * l.x + t2.z * l.y + n.z * l.z});
    o = F3{p.x + (n.x * r) * 0.5f, p.y + (n.y * r) * 0.5f, p.z + (n.z * r) * 0.5f};
}

}  // namespace cwao
