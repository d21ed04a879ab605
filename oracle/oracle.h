/*
 * oracle.h -- CPU restatement of the reference traversal (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
 * library, and only as the checker; the product never links or calls it.
 *
 * Arithmetic contract shared with the HIP kernels (each step is one IEEE-754
 * binary32 operation, FMA where written, no other contraction):
 *   idir_k  = 1 / (|d_k| > 2^-80 ? d_k : copysign(2^-80, d_k))
 *   CWBVH   adj_k  = 2^e_k * idir_k ; orig_k = (-(o_k - p_k)) * idir_k
 *           slab   = fma(q, adj_k, orig_k)
 *           cmin   = max_int(max3_int(near slabs), tmin); cmax = min_int(min3_int(far slabs), tmax)
 *           (the reference combines with tmin/tmax through fmaxf/fminf; for
 *            tmin, tmax >= +0 and non-NaN slabs both orders give the same value)
 *   BVH2    ood_k  = o_k * idir_k ; slab = fma(bound, idir_k, -ood_k)
 *           tmin'  = max_int(min_int(lo,hi) per axis, tmin); tmax' = min_int(max_int(lo,hi) per axis, tmax)
 *   (max/min_int = order of the IEEE bit patterns read as int32, the reference's
 *    vmin/vmax.s32 semantics, src/FastDeviceMinMax.h:44-79)
 *   Woop    Oz = fma(-o.z,v0.z, fma(-o.y,v0.y, fma(-o.x,v0.x, v0.w)))
 *           Dz = fma(d.z,v0.z, fma(d.y,v0.y, d.x*v0.x)); t = Oz * (1/Dz)
 *           Ox = fma(o.z,v1.z, fma(o.y,v1.y, fma(o.x,v1.x, v1.w)))
 *           Dx = fma(d.z,v1.z, fma(d.y,v1.y, d.x*v1.x)); u = fma(t, Dx, Ox)   (v likewise with v2)
 *           accept: t > tmin, u >= 0, u <= 1, v >= 0, u + v <= 1 and
 *                   closest: t < tmax, or t == tmax with a smaller triAddr than the current hit
 *                   any:     t < tmax (ray tmax), then stop.
 * Hit record = (t, triAddr bits, u, v); miss = (ray tmax, -1, 0, 0).
 */
#ifndef CWBVH_ORACLE_H
#define CWBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z, w; } or_float4;
typedef struct { or_float4 origin_tmin, dir_tmax; } or_ray;
typedef struct { or_float4 t_triId_u_v; } or_hit;

/* One Woop ray/triangle test; returns 1 when the triangle is hit inside
 * (tmin, tmax) (strict on both ends) with inclusive edge tests. */
int oracle_woop_test(const or_float4* woop3, const or_ray* ray, float tmin, float tmax,
                     float* t, float* u, float* v);

/* Scalar CWBVH traversal (reference src/TraversalKernelCWBVH.cu:56-371 without
 * the warp-level dynamic fetch and triangle postponing).  any_hit != 0 stops at
 * the first accepted triangle.  counters (nullable) += {nodes, triangles} per
 * ray.  threads <= 1 runs on the calling thread. Returns 0, or -1 when a
 * ray's stack would exceed stack_capacity entries. */
int oracle_trace_cwbvh(const or_float4* nodes, int node_stride, const or_float4* woop,
                       const or_ray* rays, or_hit* hits, int64_t n, int any_hit,
                       uint64_t* counters, int threads, int stack_capacity);

/* Scalar BVH2 while-while traversal (reference src/TraversalKernelBVH2.cu:26-178). */
int oracle_trace_bvh2(const or_float4* nodes, const or_float4* woop, const or_ray* rays,
                      or_hit* hits, int64_t n, int any_hit, uint64_t* counters,
                      int threads, int stack_capacity);

/* Brute force: test the triangles at Woop float4 addresses tri_addrs[0..ntri)
 * for every ray (closest with the same tie rule, or any). */
int oracle_trace_brute(const or_float4* woop, const int32_t* tri_addrs, int64_t ntri,
                       const or_ray* rays, or_hit* hits, int64_t n, int any_hit, int threads);

/* Reference Woop transform (src/GPUBVHConverter.cpp:380-403 with
 * src/WoopTriangleHelper.h:318-354): invert the column-major 4x4 matrix
 * [v0-v2 | v1-v2 | (v0-v2)x(v1-v2) | (v2,1)] by cofactors in binary32;
 * row2 (w negated) -> out[0], row0 -> out[1], row1 -> out[2]; det == 0 gives
 * the identity matrix as the reference does. */
void oracle_woopify_ref(const float* v0, const float* v1, const float* v2, or_float4* out3);

/* Decode one CWBVH node's 8 child boxes in double precision:
 * lo/hi[c*3 + axis] = p + q * 2^e. */
void oracle_cwbvh_decode(const or_float4* node, double* lo24, double* hi24);

#ifdef __cplusplus
}
#endif

#endif
