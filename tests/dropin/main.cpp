// A caller of the reference's four entry points, written the way the
// reference's driver uses them (src/main.cpp:88-245): build the scene, convert
// it to the BVH2 and CWBVH layouts, allocate and upload device buffers, bind
// both layouts (:138-142, :195-199), trace the same rays with both kernels
// (:207-208) and read the hits back.  It includes the drop-in headers
// include/TraversalKernelBVH2.h and include/TraversalKernelCWBVH.h with its
// own float4 / Ray / Hit, as the reference's main.cpp does through
// helper_math.h, and links libcwbvh_amd.so.  Scene building goes through the
// library's C ABI (the reference used Embree and its converter here).
//
// Unlike the reference, which traces both kernels into one hit buffer, each
// kernel gets its own, and everything needed to check the hits is written to
// the output directory as raw little-endian arrays:
//   rays.f32 (n x 8), bvh2_nodes.f32, bvh2_woop.f32, cwbvh_nodes.f32,
//   cwbvh_woop.f32 (x 4 floats), hits_bvh2.f32, hits_cwbvh.f32 (n x 4),
//   meta.txt ("key value" lines).
//
// usage: dropin_main <out_dir> [num_points] [rays_per_point]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cwbvh_amd.h"

struct Ray {
    float4 origin_tmin;
    float4 dir_tmax;
};

struct Hit {
    float4 t_triId_u_v;
};

#include "TraversalKernelBVH2.h"
#include "TraversalKernelCWBVH.h"

#define CHECK_HIP(x)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                   \
            std::exit(2);                                                                         \
        }                                                                                         \
    } while (0)
#define CHECK_CW(x)                                                                               \
    do {                                                                                          \
        int32_t s_ = (x);                                                                         \
        if (s_ != CWBVH_OK) {                                                                     \
            std::fprintf(stderr, "%s failed: %s (%s)\n", #x, cwbvh_status_string(s_), cwbvh_last_error()); \
            std::exit(3);                                                                         \
        }                                                                                         \
    } while (0)

template <class T>
static void write_file(const std::string& path, const T* data, size_t count) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(data, sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(4);
    }
    std::fclose(f);
}

// Device copies of one layout's arrays, as the reference's driver makes them.
struct DeviceLayout {
    float4* nodes = nullptr;
    float4* woop = nullptr;
    int* prim = nullptr;
};

static DeviceLayout upload(const cwbvh_layout_view& v) {
    DeviceLayout d;
    CHECK_HIP(hipMalloc(&d.nodes, v.nodes_f4 * sizeof(float4)));
    CHECK_HIP(hipMalloc(&d.woop, v.woop_f4 * sizeof(float4)));
    CHECK_HIP(hipMalloc(&d.prim, v.prim_idx_len * sizeof(int)));
    CHECK_HIP(hipMemcpy(d.nodes, v.nodes, v.nodes_f4 * sizeof(float4), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d.woop, v.woop, v.woop_f4 * sizeof(float4), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d.prim, v.prim_idx, v.prim_idx_len * sizeof(int), hipMemcpyHostToDevice));
    return d;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <out_dir> [num_points] [rays_per_point]\n", argv[0]);
        return 1;
    }
    const std::string out = argv[1];
    const int64_t num_points = argc > 2 ? std::atoll(argv[2]) : 131072;
    const int rays_per_point = argc > 3 ? std::atoi(argv[3]) : 8;

    // the 256 x 128 heightfield with 8 sine octaves and its AO rays
    cwbvh_mesh* mesh = nullptr;
    CHECK_CW(cwbvh_mesh_heightfield(256, 128, 8, 0.1f, 0xC0C00001ull, &mesh));
    float lo[3], hi[3];
    CHECK_CW(cwbvh_mesh_bounds(mesh, lo, hi));
    double diag2 = 0.0;
    for (int k = 0; k < 3; ++k) diag2 += double(hi[k] - lo[k]) * double(hi[k] - lo[k]);
    const float diag = (float)std::sqrt(diag2);
    const int numRays = (int)(num_points * rays_per_point);
    std::vector<Ray> hostRays(numRays);
    CHECK_CW(cwbvh_rays_ao_surface(mesh, num_points, rays_per_point, 1e-4f * diag, 0.0f, 0.1f * diag,
                                   0xC0C00002ull, reinterpret_cast<cwbvh_ray*>(hostRays.data())));

    std::printf("Building BVH2 and CWBVH\n");
    cwbvh_layout *bvh2 = nullptr, *cwbvh = nullptr;
    cwbvh_build_params params;
    CHECK_CW(cwbvh_build_params_default(&params));
    CHECK_CW(cwbvh_build_layouts(mesh, &params, &bvh2, &cwbvh));
    cwbvh_layout_view v2, v8;
    CHECK_CW(cwbvh_layout_get_view(bvh2, &v2));
    CHECK_CW(cwbvh_layout_get_view(cwbvh, &v8));

    Ray* cudaRays = nullptr;
    CHECK_HIP(hipMalloc(&cudaRays, sizeof(Ray) * numRays));
    CHECK_HIP(hipMemcpy(cudaRays, hostRays.data(), sizeof(Ray) * numRays, hipMemcpyHostToDevice));

    const DeviceLayout d2 = upload(v2);
    rtBindBVH2Data(d2.nodes, d2.woop, d2.prim);
    CHECK_CW(cwbvh_last_status());
    const DeviceLayout d8 = upload(v8);
    rtBindCWBVHData(d8.nodes, d8.woop, d8.prim);
    CHECK_CW(cwbvh_last_status());

    Hit *hitsBVH2 = nullptr, *hitsCWBVH = nullptr;
    CHECK_HIP(hipMalloc(&hitsBVH2, sizeof(Hit) * numRays));
    CHECK_HIP(hipMalloc(&hitsCWBVH, sizeof(Hit) * numRays));

    std::printf("Launching kernels\n");
    rtTraceBVH2(cudaRays, hitsBVH2, numRays);
    CHECK_CW(cwbvh_last_status());
    const float msBVH2 = cwbvh_last_trace_ms();
    rtTraceCWBVH(cudaRays, hitsCWBVH, numRays);
    CHECK_CW(cwbvh_last_status());
    const float msCWBVH = cwbvh_last_trace_ms();

    std::vector<Hit> hostBVH2(numRays), hostCWBVH(numRays);
    CHECK_HIP(hipMemcpy(hostBVH2.data(), hitsBVH2, sizeof(Hit) * numRays, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(hostCWBVH.data(), hitsCWBVH, sizeof(Hit) * numRays, hipMemcpyDeviceToHost));

    // the first 10 results, as the reference prints them
    for (int i = 0; i < 10 && i < numRays; ++i) {
        const float4 h = hostCWBVH[i].t_triId_u_v;
        int id;
        std::memcpy(&id, &h.y, sizeof(id));
        std::printf("%.2f %d\t", h.x, id);
    }
    std::printf("\n");

    write_file(out + "/rays.f32", reinterpret_cast<const float*>(hostRays.data()), (size_t)numRays * 8);
    write_file(out + "/bvh2_nodes.f32", reinterpret_cast<const float*>(v2.nodes), (size_t)v2.nodes_f4 * 4);
    write_file(out + "/bvh2_woop.f32", reinterpret_cast<const float*>(v2.woop), (size_t)v2.woop_f4 * 4);
    write_file(out + "/cwbvh_nodes.f32", reinterpret_cast<const float*>(v8.nodes), (size_t)v8.nodes_f4 * 4);
    write_file(out + "/cwbvh_woop.f32", reinterpret_cast<const float*>(v8.woop), (size_t)v8.woop_f4 * 4);
    write_file(out + "/hits_bvh2.f32", reinterpret_cast<const float*>(hostBVH2.data()), (size_t)numRays * 4);
    write_file(out + "/hits_cwbvh.f32", reinterpret_cast<const float*>(hostCWBVH.data()), (size_t)numRays * 4);
    FILE* meta = std::fopen((out + "/meta.txt").c_str(), "w");
    if (!meta) return 4;
    std::fprintf(meta, "rays %d\ncwbvh_node_stride %d\nbvh2_ms %.4f\ncwbvh_ms %.4f\n", numRays, v8.node_stride,
                 msBVH2, msCWBVH);
    std::fclose(meta);

    (void)hipFree(cudaRays);
    (void)hipFree(hitsBVH2);
    (void)hipFree(hitsCWBVH);
    (void)hipFree(d2.nodes);
    (void)hipFree(d2.woop);
    (void)hipFree(d2.prim);
    (void)hipFree(d8.nodes);
    (void)hipFree(d8.woop);
    (void)hipFree(d8.prim);
    cwbvh_layout_free(bvh2);
    cwbvh_layout_free(cwbvh);
    cwbvh_mesh_free(mesh);
    std::printf("dropin ok: %d rays through rtTraceBVH2 and rtTraceCWBVH\n", numRays);
    return 0;
}
