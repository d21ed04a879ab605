/*
 * Drop-in replacement for the reference header src/TraversalKernelCWBVH.h:3-11.
 *
 * Same function names, argument lists, meaning and linkage (C++, as in the
 * reference): libcwbvh_amd.so exports these symbols, mangled for HIP's float4
 * and global-namespace structs named Ray and Hit.  As in the reference, the
 * includer provides float4 (e.g. <hip/hip_runtime.h>), Ray = two float4
 * {origin.xyz, tmin}, {dir.xyz, tmax} and Hit = one float4 {t, triAddr bits,
 * u, v}, with those names, before including this header.
 * rtTraceCWBVH runs on the null stream and returns after the kernel finished; it
 * logs "<ms>MS, <rate>MRays/s (...)" like the reference, and
 * cwbvh_last_status() / cwbvh_last_trace_ms() (cwbvh_amd.h) report the outcome
 * instead of an abort.  C callers use the C-linkage entry points of the same
 * names in cwbvh_amd.h.
 */
#pragma once

void rtBindCWBVHData(const float4* InBVHTreeNodes, const float4* InTriangleWoopCoordinates,
                     const int* InMappingFromTriangleAddressToIndex);

void rtTraceCWBVH(Ray* rayBuffer, Hit* rayResultBuffer, int rayCount);
