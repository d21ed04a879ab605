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

// 1 / d with |d| clamped to 2^-80 (reference src/TraversalKernelCWBVH.cu:78-80).
__device__ __forceinline__ float safe_rcp(float d) {
    const float ooeps = 0x1p-80f;
    return 1.0f / (fabsf(d) > ooeps ? d : copysignf(ooeps, d));
}

// Woop ray/triangle test (reference src/TraversalKernelCWBVH.cu:318-347).
//   Oz = v0.w - o.v0   Dz = d.v0   t = Oz * (1/Dz)
//   u  = (v1.w + o.v1) + t*(d.v1)   v likewise with v2
// Accept: tmin < t, 0 <= u <= 1, 0 <= v, u + v <= 1, and
//   ANY:     t < tmax
//   closest: t < tmax, or t == tmax and tri_addr < hit_addr (deterministic ties)
template <bool ANY>
__device__ __forceinline__ bool woop_test(const float4 v0, const float4 v1, const float4 v2, const float3 o,
                                          const float3 d, float tmin, float tmax, int tri_addr, int hit_addr,
                                          float& t_out, float& u_out, float& v_out) {
    const float Oz = __builtin_fmaf(-o.z, v0.z, __builtin_fmaf(-o.y, v0.y, __builtin_fmaf(-o.x, v0.x, v0.w)));
    const float Dz = __builtin_fmaf(d.z, v0.z, __builtin_fmaf(d.y, v0.y, d.x * v0.x));
    const float invDz = 1.0f / Dz;
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
template <class T, int CAP>
struct Stack {
    T* lds;        // this thread's column base: lds_base + tid
    int stride;    // blockDim.x
    int depth;     // entries held in LDS
    T spill[CAP];

    __device__ __forceinline__ void init(T* base, int tid, int block, int d) {
        lds = base + tid;
        stride = block;
        depth = d;
    }
    __device__ __forceinline__ void put(int sp, T v) {
        if (sp < depth) lds[sp * stride] = v;
        else spill[sp - depth] = v;
    }
    __device__ __forceinline__ T get(int sp) const {
        if (sp < depth) return lds[sp * stride];
        return spill[sp - depth];
    }
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

}  // namespace cwd
