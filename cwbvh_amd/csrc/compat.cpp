// Reference entry points (drop-in): rtBindCWBVHData / rtTraceCWBVH and
// rtBindBVH2Data / rtTraceBVH2, same names and argument meaning as
// src/TraversalKernelCWBVH.h:3-11 and src/TraversalKernelBVH2.h:3-11.
//
// Semantics kept from the reference (src/TraversalKernelCWBVH.cu:377-426,
// src/TraversalKernelBVH2.cu:180-227): bind stores device pointers without
// taking ownership and replaces any earlier binding; trace runs on the null
// stream, waits for completion, writes one hit per ray and logs
// "<ms>MS, <rate>MRays/s (...)".  Instead of assert(0) on a failure, the
// status is kept for cwbvh_last_status().
//
// Each entry point is exported twice:
//  * with C++ linkage and the reference's own parameter types (HIP float4,
//    global-namespace Ray and Hit), so the symbols carry the mangled names a
//    caller compiled against src/TraversalKernelCWBVH.h / TraversalKernelBVH2.h
//    links to (include/TraversalKernel*.h declare exactly these);
//  * with C linkage over cwbvh_float4 / cwbvh_ray / cwbvh_hit (cwbvh_amd.h),
//    for FFI callers that cannot bind mangled C++ names.
#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdio>
#include <mutex>

#include "internal.h"

namespace {

std::mutex g_mu;
cwbvh_scene* g_cwbvh = nullptr;
cwbvh_scene* g_bvh2 = nullptr;
thread_local int32_t g_status = CWBVH_OK;
thread_local float g_ms = 0.f;
int g_log = 1;

void bind(cwbvh_scene*& slot, int32_t kind, int32_t stride, const cwbvh_float4* nodes, const cwbvh_float4* woop,
          const int* map) {
    std::lock_guard<std::mutex> g(g_mu);
    if (slot) {
        cwbvh_scene_destroy(slot);
        slot = nullptr;
    }
    g_status = cwbvh_scene_wrap_device(kind, stride, nodes, woop, map, 0, &slot);
}

void trace(cwbvh_scene* scene, cwbvh_ray* rays, cwbvh_hit* hits, int count, const char* label) {
    std::lock_guard<std::mutex> g(g_mu);
    g_ms = 0.f;
    if (!scene) {
        cw::set_error("trace before bind");
        g_status = CWBVH_ERR_NOT_BOUND;
        return;
    }
    if (count <= 0) {
        g_status = count == 0 ? CWBVH_OK : CWBVH_ERR_INVALID_ARGUMENT;
        return;
    }
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t ce = hipEventCreate(&a);
    if (ce == hipSuccess) ce = hipEventCreate(&b);
    if (ce != hipSuccess) {
        if (a) hipEventDestroy(a);
        g_status = cw::hip_fail(ce, "rtTrace event");
        return;
    }
    // the reference brackets each trace with cudaProfilerStart / Stop
    // (src/TraversalKernelCWBVH.cu:404,423): the roctx counterparts scope
    // rocprofv3 --selected-regions --marker-trace to the launch
    roctxProfilerResume(0);
    roctxRangePush(label);
    hipEventRecord(a, nullptr);
    int32_t rc = cwbvh_trace(scene, rays, hits, count, CWBVH_TRACE_CLOSEST, nullptr, nullptr, nullptr);
    hipEventRecord(b, nullptr);
    hipError_t e = hipEventSynchronize(b);
    roctxRangePop();
    roctxProfilerPause(0);
    if (rc == CWBVH_OK && e != hipSuccess) rc = cw::hip_fail(e, "rtTrace");
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc == CWBVH_OK) {
        int32_t of = 0;
        rc = cwbvh_stack_overflowed(nullptr, &of);
        if (rc == CWBVH_OK && of) {
            cw::set_error("traversal stack overflow");
            rc = CWBVH_ERR_STACK_DEPTH;
        }
    }
    g_status = rc;
    g_ms = ms;
    if (g_log && rc == CWBVH_OK)
        std::printf("%.3fMS, %.2fMRays/s (%s)\n", ms, (double)count / 1e6 / (ms / 1e3), label);
}

}  // namespace

extern "C" {

void rtBindCWBVHData(const cwbvh_float4* InBVHTreeNodes, const cwbvh_float4* InTriangleWoopCoordinates,
                     const int* InMappingFromTriangleAddressToIndex) {
    bind(g_cwbvh, CWBVH_KIND_CWBVH, 5, InBVHTreeNodes, InTriangleWoopCoordinates, InMappingFromTriangleAddressToIndex);
}

void rtTraceCWBVH(cwbvh_ray* rayBuffer, cwbvh_hit* rayResultBuffer, int rayCount) {
    trace(g_cwbvh, rayBuffer, rayResultBuffer, rayCount, "rtTraceCWBVH Dynamic Fetch");
}

void rtBindBVH2Data(const cwbvh_float4* InBVHTreeNodes, const cwbvh_float4* InTriangleWoopCoordinates,
                    const int* InMappingFromTriangleAddressToIndex) {
    bind(g_bvh2, CWBVH_KIND_BVH2, 4, InBVHTreeNodes, InTriangleWoopCoordinates, InMappingFromTriangleAddressToIndex);
}

void rtTraceBVH2(cwbvh_ray* rayBuffer, cwbvh_hit* rayResultBuffer, int rayCount) {
    trace(g_bvh2, rayBuffer, rayResultBuffer, rayCount, "rtTraceBVH2 No Dynamic Fetch");
}

int32_t cwbvh_profiler_start(void) {
    roctxProfilerResume(0);
    return CWBVH_OK;
}

int32_t cwbvh_profiler_stop(void) {
    roctxProfilerPause(0);
    return CWBVH_OK;
}

int32_t cwbvh_last_status(void) { return g_status; }
float cwbvh_last_trace_ms(void) { return g_ms; }
void cwbvh_set_log(int32_t enabled) { g_log = enabled ? 1 : 0; }

}  // extern "C"

// The reference's ray and hit records (src/helper_math.h:1853-1865), named as
// there so that the C++ entry points below mangle to the reference's symbols.
struct Ray {
    float4 origin_tmin;
    float4 dir_tmax;
};
struct Hit {
    float4 t_triId_u_v;
};
static_assert(sizeof(float4) == sizeof(cwbvh_float4) && sizeof(Ray) == sizeof(cwbvh_ray) &&
                  sizeof(Hit) == sizeof(cwbvh_hit),
              "the reference records and the C ABI records have one layout");

void rtBindCWBVHData(const float4* InBVHTreeNodes, const float4* InTriangleWoopCoordinates,
                     const int* InMappingFromTriangleAddressToIndex) {
    rtBindCWBVHData(reinterpret_cast<const cwbvh_float4*>(InBVHTreeNodes),
                    reinterpret_cast<const cwbvh_float4*>(InTriangleWoopCoordinates),
                    InMappingFromTriangleAddressToIndex);
}

void rtTraceCWBVH(Ray* rayBuffer, Hit* rayResultBuffer, int rayCount) {
    rtTraceCWBVH(reinterpret_cast<cwbvh_ray*>(rayBuffer), reinterpret_cast<cwbvh_hit*>(rayResultBuffer), rayCount);
}

void rtBindBVH2Data(const float4* InBVHTreeNodes, const float4* InTriangleWoopCoordinates,
                    const int* InMappingFromTriangleAddressToIndex) {
    rtBindBVH2Data(reinterpret_cast<const cwbvh_float4*>(InBVHTreeNodes),
                   reinterpret_cast<const cwbvh_float4*>(InTriangleWoopCoordinates),
                   InMappingFromTriangleAddressToIndex);
}

void rtTraceBVH2(Ray* rayBuffer, Hit* rayResultBuffer, int rayCount) {
    rtTraceBVH2(reinterpret_cast<cwbvh_ray*>(rayBuffer), reinterpret_cast<cwbvh_hit*>(rayResultBuffer), rayCount);
}
