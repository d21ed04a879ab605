/*
 * Drop-in replacement for the reference header src/TraversalKernelCWBVH.h:3-11.
 *
 * Same function names, argument lists and meaning; the definitions live in
 * libcwbvh_amd.so with C linkage (implemented over the C ABI of cwbvh_amd.h).
 * As in the reference, the includer provides float4, Ray and Hit (Ray = two
 * float4 {origin.xyz, tmin}, {dir.xyz, tmax}; Hit = one float4 {t, triAddr
 * bits, u, v}), e.g. from <hip/hip_vector_types.h> plus its own Ray/Hit.
 * rtTraceCWBVH runs on the null stream and returns after the kernel finished;
 * cwbvh_last_status() / cwbvh_last_trace_ms() (cwbvh_amd.h) report the outcome.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

void rtBindCWBVHData(const float4* InBVHTreeNodes, const float4* InTriangleWoopCoordinates,
                     const int* InMappingFromTriangleAddressToIndex);

void rtTraceCWBVH(Ray* rayBuffer, Hit* rayResultBuffer, int rayCount);

#ifdef __cplusplus
}
#endif
