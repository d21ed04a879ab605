/*
 * Drop-in replacement for the reference header src/TraversalKernelBVH2.h:3-11.
 * Same conventions as TraversalKernelCWBVH.h: the includer provides float4,
 * Ray and Hit; the definitions live in libcwbvh_amd.so with C linkage.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

void rtBindBVH2Data(const float4* InBVHTreeNodes, const float4* InTriangleWoopCoordinates,
                    const int* InMappingFromTriangleAddressToIndex);

void rtTraceBVH2(Ray* rayBuffer, Hit* rayResultBuffer, int rayCount);

#ifdef __cplusplus
}
#endif
