// Device-side building blocks shared by the traversal kernels (gfx950, wave64).
//
// Arithmetic follows one contract, stated in full beside each function, so
// that every kernel variant returns the same bits for the same ray:
// one binary32 rounding per operation, fused multiply-add only where
// __builtin_fmaf is written (the sources are compiled with -ffp-contract=off),
// IEEE division.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cwd {

struct Ray4 {
    float4 o;  // origin.xyz, tmin
    float4 d;  // dir.xyz, tmax
};

constexpr int kStackCapacity = 64;      // CWBVH entries (uint2)
constexpr int kBvh2StackCapacity = 64;  // BVH2 entries (int), sentinel included

__device__ __forceinline__ int f2i(float f) { return __float_as_int(f); }
__device__ __forceinline__ float i2f(int i) { return __int_as_float(i); }
__device__ __forceinline__ unsigned f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(unsigned u) { return __uint_as_float(u); }

// Order of IEEE bit patterns read as int32 (reference vmin/vmax.s32 on float
// bits, src/FastDeviceMinMax.h:44-51); compiles to v_min3_i32 / v_max3_i32.
__device__ __forceinline__ float imin2(float a, float b) { return i2f(min(f2i(a), f2i(b))); }
__device__ __forceinline__ float imax2(float a, float b) { return i2f(max(f2i(a), f2i(b))); }
__device__ __forceinline__ float imin3(float a, float b, float c) { return i2f(min(min(f2i(a), f2i(b)), f2i(c))); }
__device__ __forceinline__ float imax3(float a, float b, float c) { return i2f(max(max(f2i(a), f2i(b)), f2i(c))); }

// |d| clamped to 2^-80 (reference src/TraversalKernelCWBVH.cu:78-80); a NaN
// becomes +-2^-80 as well.
__device__ __forceinline__ float clamp_dir(float d) {
    const float ooeps = 0x1p-80f;
    return fabsf(d) > ooeps ? d : copysignf(ooeps, d);
}

// 1 / x as v_rcp_f32 refined by one fma Newton step.  For every binary32 x
// with |x| in [2^-125, 2^125] this equals the IEEE quotient 1.0f / x (all
// such values compared on gfx950); outside that range it may not.
__device__ __forceinline__ float rcp_refined(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
}

// 1 / d per component with |d| clamped to 2^-80, i.e. the IEEE quotients of
// the reference's ray setup (src/TraversalKernelCWBVH.cu:78-80, and
// src/TraversalKernelBVH2.cu:52-55).  Clamped components are at least
// 2^-80 in magnitude, so the refined reciprocal (3 instructions) returns the
// quotient's bits unless a component exceeds 2^125 or is infinite; the lanes
// calling together take the IEEE division (11 instructions each) when any of
// them holds such a component.
__device__ __forceinline__ void safe_rcp3(const float3 d, float& ix, float& iy, float& iz) {
    const float cx = clamp_dir(d.x), cy = clamp_dir(d.y), cz = clamp_dir(d.z);
    const bool in_range = fabsf(cx) <= 0x1p125f && fabsf(cy) <= 0x1p125f && fabsf(cz) <= 0x1p125f;
    if (__all(in_range)) {
        ix = rcp_refined(cx);
        iy = rcp_refined(cy);
        iz = rcp_refined(cz);
    } else {
        ix = 1.0f / cx;
        iy = 1.0f / cy;
        iz = 1.0f / cz;
    }
}

// Woop ray/triangle test (reference src/TraversalKernelCWBVH.cu:318-347).
//   Oz = v0.w - o.v0   Dz = d.v0   t = Oz * (1/Dz)
//   u  = (v1.w + o.v1) + t*(d.v1)   v likewise with v2
// Accept: tmin < t, 0 <= u <= 1, 0 <= v, u + v <= 1, and
//   ANY:     t < tmax
//   closest: t < tmax, or t == tmax and tri_addr < hit_addr (deterministic ties)
// 1/Dz is the IEEE quotient: computed as rcp_refined (3 instructions against
// the division's 11) when every calling lane's |Dz| lies in [2^-125, 2^125],
// where the two agree, and by the division otherwise.
template <bool ANY>
__device__ __forceinline__ bool woop_test(const float4 v0, const float4 v1, const float4 v2, const float3 o,
                                          const float3 d, float tmin, float tmax, int tri_addr, int hit_addr,
                                          float& t_out, float& u_out, float& v_out) {
    const float Oz = __builtin_fmaf(-o.z, v0.z, __builtin_fmaf(-o.y, v0.y, __builtin_fmaf(-o.x, v0.x, v0.w)));
    const float Dz = __builtin_fmaf(d.z, v0.z, __builtin_fmaf(d.y, v0.y, d.x * v0.x));
    float invDz;
    if (__all(fabsf(Dz) >= 0x1p-125f && fabsf(Dz) <= 0x1p125f)) invDz = rcp_refined(Dz);
    else invDz = 1.0f / Dz;
    const float t = Oz * invDz;
    bool ok;
    if (ANY) ok = t > tmin && t < tmax;
    else ok = t > tmin && (t < tmax || (t == tmax && hit_addr >= 0 && tri_addr < hit_addr));
    const float Ox = __builtin_fmaf(o.z, v1.z, __builtin_fmaf(o.y, v1.y, __builtin_fmaf(o.x, v1.x, v1.w)));
    const float Dx = __builtin_fmaf(d.z, v1.z, __builtin_fmaf(d.y, v1.y, d.x * v1.x));
    const float u = __builtin_fmaf(t, Dx, Ox);
    const float Oy = __builtin_fmaf(o.z, v2.z, __builtin_fmaf(o.y, v2.y, __builtin_fmaf(o.x, v2.x, v2.w)));
    const float Dy = __builtin_fmaf(d.z, v2.z, __builtin_fmaf(d.y, v2.y, d.x * v2.x));
    const float v = __builtin_fmaf(t, Dy, Oy);
    ok = ok && u >= 0.0f && u <= 1.0f && v >= 0.0f && u + v <= 1.0f;
    t_out = t;
    u_out = u;
    v_out = v;
    return ok;
}

// Traversal stack: the first `depth` entries of every lane live in LDS,
// interleaved by lane (entry k of thread tid at k * blockDim + tid, so a wave
// touches consecutive words); deeper entries spill to a per-lane scratch array.
// The LDS array and the spill array are passed separately so that the index
// state stays in registers and LDS is addressed with ds_* instructions.
// LDS part of a traversal stack: each wave owns depth x 64 consecutive
// entries, entry k of lane l at (k << 6) + l inside them, so a lane's entries
// are one shift and one add away (no multiply) and the 64 lanes of a wave
// touch 64 consecutive words per push or pop.
struct StackIndex {
    int col;    // (wave in block) x depth x 64 + lane
    int depth;  // entries held in LDS
    __device__ StackIndex(int thread, int depth_) : col(((thread >> 6) * depth_ << 6) + (thread & 63)), depth(depth_) {}
};

#define CWD_LDS __attribute__((address_space(3)))

template <class T>
__device__ __forceinline__ CWD_LDS T* lds_ptr(T* p) {
    return (CWD_LDS T*)p;
}

template <class T, int CAP>
__device__ __forceinline__ void stack_put(CWD_LDS T* lds, T (&spill)[CAP], const StackIndex& s, int sp, T v) {
    if (sp < s.depth) lds[(sp << 6) + s.col] = v;
    else spill[sp - s.depth] = v;
}

// The LDS pointer keeps its address space in the type, so the two loads below
// cannot be merged into one generic (flat) load through a selected pointer.
template <class T, int CAP>
__device__ __forceinline__ T stack_get(const CWD_LDS T* lds, const T (&spill)[CAP], const StackIndex& s, int sp) {
    if (sp < s.depth) return lds[(sp << 6) + s.col];
    return spill[sp - s.depth];
}

// Streaming (read-once / write-once) 16-byte accesses for rays and hits.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 nt_load4(const float4* p) {
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void nt_store4(float4 v, float4* p) {
    const f4v w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<f4v*>(p));
}

// CWBVH stack entries (node or triangle group) travel as one 64-bit word.
__device__ __forceinline__ unsigned long long pack2(uint2 v) {
    return ((unsigned long long)v.y << 32) | (unsigned long long)v.x;
}
__device__ __forceinline__ uint2 unpack2(unsigned long long v) {
    return make_uint2((unsigned)(v & 0xFFFFFFFFull), (unsigned)(v >> 32));
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

}  // namespace cwd
