// Internal host-side definitions shared by the C++ translation units.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cwbvh_amd.h"

namespace cw {

// ---------------------------------------------------------------- errors

void set_error(const std::string& msg);
const char* get_error();

#define CW_CHECK_ARG(cond, msg)                      \
    do {                                             \
        if (!(cond)) {                               \
            ::cw::set_error(msg);                    \
            return CWBVH_ERR_INVALID_ARGUMENT;       \
        }                                            \
    } while (0)

// ------------------------------------------------------------ vector math

struct V3 {
    float x, y, z;
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    float l = length(a);
    return l > 0.f ? a * (1.0f / l) : a;
}
// Component-wise min/max that keep `a` when `b` is NaN (as fmin/fmax do for a
// NaN operand), written as compares so that they compile to minss/maxss
// instead of calls into libm; callers pass the accumulator as `a`.
inline float fmin_acc(float a, float b) { return b < a ? b : a; }
inline float fmax_acc(float a, float b) { return b > a ? b : a; }
inline V3 vmin(V3 a, V3 b) { return {fmin_acc(a.x, b.x), fmin_acc(a.y, b.y), fmin_acc(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {fmax_acc(a.x, b.x), fmax_acc(a.y, b.y), fmax_acc(a.z, b.z)}; }
inline float comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

struct Box {
    V3 lo{INFINITY, INFINITY, INFINITY};
    V3 hi{-INFINITY, -INFINITY, -INFINITY};
    void grow(V3 p) { lo = vmin(lo, p); hi = vmax(hi, p); }
    void grow(const Box& b) { lo = vmin(lo, b.lo); hi = vmax(hi, b.hi); }
    bool empty() const { return !(lo.x <= hi.x); }
    // in double: the surface area of a box with float corners can exceed the
    // float range (extents above ~1e19), and SAH costs must stay finite
    double area() const {
        if (empty()) return 0.0;
        const double dx = (double)hi.x - lo.x, dy = (double)hi.y - lo.y, dz = (double)hi.z - lo.z;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
    V3 center() const { return (lo + hi) * 0.5f; }
};

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline float i2f(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

// ------------------------------------------------------------------- RNG
// SplitMix64 (Steele, Lea, Flood 2014) as a counter-based generator:
// value(seed, index) = mix(seed + (index+1) * golden gamma).

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t state;
    explicit Rng(uint64_t seed, uint64_t stream = 0) : state(mix64(seed ^ mix64(stream + 0x632be59bd9b4e019ULL))) {}
    uint64_t next() {
        state += 0x9e3779b97f4a7c15ULL;
        return mix64(state);
    }
    // uniform in [0,1) with 24 random bits
    float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    double uniform_d() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    float normal() {
        // Box-Muller on doubles, one value per call
        double u1 = uniform_d(), u2 = uniform_d();
        if (u1 < 1e-300) u1 = 1e-300;
        return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
};

// ------------------------------------------------------------------ mesh

}  // namespace cw

struct cwbvh_mesh {
    std::vector<float> vertices;   // 3 per vertex
    std::vector<int32_t> indices;  // 3 per triangle
    std::vector<int32_t> materials;  // one per triangle when read from a mesh file, else empty
    int64_t nv() const { return (int64_t)vertices.size() / 3; }
    int64_t nt() const { return (int64_t)indices.size() / 3; }
    cw::V3 vert(int64_t i) const { return cw::V3{vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]}; }
    void tri(int64_t t, cw::V3& a, cw::V3& b, cw::V3& c) const {
        a = vert(indices[3 * t]);
        b = vert(indices[3 * t + 1]);
        c = vert(indices[3 * t + 2]);
    }
};

struct cwbvh_layout {
    int32_t kind = CWBVH_KIND_CWBVH;
    int32_t node_stride = 5;
    std::vector<cwbvh_float4> nodes;
    std::vector<cwbvh_float4> woop;
    std::vector<int32_t> prim_idx;
    int32_t max_depth = 0;
    int32_t stack_need = 0;
    int64_t num_nodes = 0;
    int64_t num_leaves = 0;
    int64_t num_tri_refs = 0;
    double sah_cost = 0.0;
};

namespace cw {

int num_threads_or_default(int requested);

// Run fn(begin, end) over [0, n) split across threads.
template <class F>
void parallel_for(int64_t n, int threads, F&& fn);

}  // namespace cw

#include <thread>

template <class F>
void cw::parallel_for(int64_t n, int threads, F&& fn) {
    if (n <= 0) return;
    int t = threads;
    if (t < 1) t = 1;
    if (n < 4096 || t == 1) {
        fn((int64_t)0, n);
        return;
    }
    if ((int64_t)t > n) t = (int)n;
    std::vector<std::thread> pool;
    int64_t chunk = (n + t - 1) / t;
    for (int i = 0; i < t; ++i) {
        int64_t b = i * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto& th : pool) th.join();
}
