// Host-side internal interfaces between the C ABI (api.cpp, compat.cpp,
// batch/serving code) and the kernel launchers (*.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <mutex>
#include <string>

#include "common.h"

struct cwbvh_scene {
    int32_t kind = CWBVH_KIND_CWBVH;
    int32_t node_stride = 5;
    const void* nodes = nullptr;   // device
    const void* woop = nullptr;    // device
    const int32_t* prim_idx = nullptr;
    int64_t woop_f4 = 0;           // valid float4 count of the Woop array (0 = unknown)
    int32_t stack_need = 0;
    bool owning = false;
    int64_t bytes = 0;
    int32_t device = 0;
    float lo[3] = {0.f, 0.f, 0.f};  // scene bounds (Morton ordering); unit cube if unknown
    float hi[3] = {1.f, 1.f, 1.f};
    bool bounds_known = false;      // lo/hi come from the root node's child boxes
};

namespace cw {

// HIP error -> status + message
int32_t hip_fail(hipError_t e, const char* what);
#define CW_HIP(call)                                      \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return ::cw::hip_fail(e_, #call); \
    } while (0)

// Per-stream work counters for persistent launches: uint32[kWorkWords] =
// {fetch counters [0, kWorkSlices), exited block groups, stack overflow flag,
// loop profile, drained-slice mask, exited blocks per group [0, kWorkSlices)}.
// The kernel resets the fetch counters, the exit counts and the mask when its
// last block exits, so launches on one stream reuse the same counters (and a
// captured graph replays cleanly).  The overflow flag stays set until read.
// Each counter sits on its own 256-byte line so that the slices' atomics are
// served by different memory channels instead of serializing on one line.
// Scenes larger than the MI355X Infinity Cache (256 MB) are traversed in the
// HBM-bound regime (see cwbvh_trace_config_for_scene).
constexpr int64_t kCacheResidentBytes = 256ll << 20;

// cwbvh_trace_config.persistent == 2: persistent grid from this many rays up
// (MI355X, AO rays on a 65 k-triangle heightfield: one ray per thread 0.142
// against 0.164 ms at 2^19 rays, 0.230 against 0.200 ms at 2^20).
constexpr int64_t kAutoPersistentRays = 3ll << 18;

constexpr int kWorkSlices = 8;
constexpr int kWorkStride = 64;                       // uint32 per counter line
constexpr int kWorkExited = kWorkSlices * kWorkStride;
constexpr int kWorkOverflow = kWorkExited + kWorkStride;
constexpr int kWorkProfile = kWorkOverflow + kWorkStride;   // uint64[6] loop-occupancy counters
constexpr int kWorkDrained = kWorkProfile + kWorkStride;    // bit s: slice s has no items left
constexpr int kWorkGroupExit = kWorkDrained + kWorkStride;  // per home slice: blocks exited
constexpr int kWorkWords = kWorkGroupExit + kWorkSlices * kWorkStride;
int32_t stream_counters(hipStream_t stream, unsigned** out);
// Frees the Morton workspace kept for `stream` on the current device.
void release_stream_scratch(hipStream_t stream);
// The lock that guards every stream's workspace: hold it from a workspace
// lookup until the kernels that use the workspace are queued.
std::unique_lock<std::mutex> lock_stream_scratch();
// Ray order for coherent traversal: rays sorted stably by a kRayKeyBits-bit
// key, their direction octant above the Morton code of their origin
// quantized to 7 bits per axis inside the scene bounds.
constexpr int kRayKeyBits = 24;
// With that lock held: the ray order of n rays (n >= 1) into the stream's
// workspace; *order points into it until the stream's next use.
int32_t morton_order_scratch(const void* d_rays, int64_t n, const float lo[3], const float hi[3],
                             hipStream_t stream, const int32_t** order);
// The ray order of n rays into `order`, with caller buffers: keys k0 and k1
// and values va and vb (n entries each) and the sort's counts ws
// (radix_sort_workspace(n, kRayKeyBits) bytes).  Stream-ordered.
int32_t morton_order_into(const void* d_rays, int64_t n, const float lo3[3], const float hi3[3], uint32_t* k0,
                          uint32_t* k1, int32_t* va, int32_t* vb, int32_t* order, void* ws, size_t ws_bytes,
                          hipStream_t st);
// cwbvh_trace_config.ray_order = 1 sorts launches of at least this many rays
// (below it the sort's fixed cost outweighs the gain)
constexpr int64_t kMortonMinRays = int64_t(1) << 20;
// ... and sorts and traces larger launches in pieces of at most this many
// rays, which bounds the kept workspace (about 1.1 GB)
constexpr int64_t kMortonMaxRays = int64_t(1) << 26;

struct TraceLaunch {
    const cwbvh_scene* scene;
    const void* rays;
    const int32_t* order;
    void* hits;
    int64_t n;
    uint32_t flags;
    cwbvh_trace_config cfg;
    uint64_t* stats;
    unsigned* counters;   // from stream_counters (or a dedicated buffer)
    hipStream_t stream;
    // rays generated from surfels instead of read (rays == nullptr): ray r is
    // sample r % rays_per_texel of surfel r / rays_per_texel (aogen.h)
    const cwbvh_float4* surf_pos = nullptr;
    const cwbvh_float4* surf_nrm = nullptr;
    const float* surf_rad = nullptr;
    int32_t rays_per_texel = 0;
    float gen_tmin = 0.f, gen_tmax = 0.f;
    uint64_t seed = 0;
    uint32_t* texel_hits = nullptr;   // per-surfel hit counts instead of hits
};

// Argument checks of cwbvh_trace / cwbvh_trace_ordered and the launch record
// they describe (counters not yet assigned; cfg NULL = the scene's default).
int32_t prepare_trace(const cwbvh_scene* scene, const cwbvh_ray* d_rays, const int32_t* d_order, cwbvh_hit* d_hits,
                      int64_t n, uint32_t flags, const cwbvh_trace_config* cfg, uint64_t* d_counters, void* stream,
                      TraceLaunch* out);
int32_t launch_cwbvh(const TraceLaunch& L);
int32_t launch_cwbvh_surfels(const TraceLaunch& L);
int32_t launch_surfel_rays(const cwbvh_float4* pos, const cwbvh_float4* nrm, const float* rad, int64_t num_texels,
                           int32_t rays_per_texel, float tmin, float tmax, uint64_t seed, cwbvh_ray* out,
                           hipStream_t stream);
int32_t launch_bvh2(const TraceLaunch& L);

// Fill the device grid for a persistent kernel: resident blocks per CU * CUs.
int device_compute_units();

// One multi-scene job for the persistent batched kernel.
struct BatchJob {
    const void* nodes;
    const void* woop;
    const void* rays;
    const int32_t* order;   // may be null
    void* hits;
    const int32_t* prim_idx;
    int32_t n;
    int32_t node_stride;
    uint32_t chunk_begin;   // first global chunk of this job
    int32_t spec_limit;     // speculative pushes allowed while the stack holds fewer entries
};
static_assert(sizeof(BatchJob) == 64, "BatchJob is 64 bytes");

// Stack room a speculating lane may fill with parked triangle groups: the
// node-group entries of a tree never exceed its stack_need, so speculation
// below (capacity - stack_need) cannot overflow.  0 (no speculation) for a
// tree of unknown depth.
inline int32_t speculation_limit(int32_t stack_need) {
    return stack_need > 0 && stack_need < CWBVH_STACK_CAPACITY ? CWBVH_STACK_CAPACITY - stack_need : 0;
}

struct BatchLaunch {
    const BatchJob* d_jobs;
    int32_t num_jobs;
    uint32_t total_chunks;
    uint32_t flags;
    cwbvh_trace_config cfg;   // fetch_chunk is the job chunk size
    uint64_t* stats;
    unsigned* counters;
    hipStream_t stream;
    int32_t lds_only;         // every job's tree needs at most cfg.lds_depth stack entries
};
// Rays per job chunk of a multi-scene launch when the configuration asks for
// an automatic chunk size (fetch_chunk == 0); batch plans and the kernel must
// agree on it.
constexpr int32_t kBatchChunk = 256;
int32_t launch_cwbvh_batch(const BatchLaunch& L);

int32_t validate_config(int32_t kind, const cwbvh_trace_config& c);

// Stable LSD radix sort of (key, value) pairs (sort.hip), 8 bits per pass
// over bits [0, key_bits).  k0 holds the keys and is overwritten; k1 is the
// other key buffer.  Values start as vals_init (nullptr: the input index),
// move through va / vb and end in dst, which receives the values in key
// order: va != vals_init, vb may be vals_init, dst differs from va and vb.
// ws holds radix_sort_workspace(n, key_bits) bytes.  Stream-ordered.
size_t radix_sort_workspace(int64_t n, int key_bits);
int32_t radix_sort_u32(void* ws, size_t ws_bytes, uint32_t* k0, uint32_t* k1, const int32_t* vals_init, int32_t* va,
                       int32_t* vb, int32_t* dst, int64_t n, int key_bits, hipStream_t st);
int32_t radix_sort_u64(void* ws, size_t ws_bytes, uint64_t* k0, uint64_t* k1, const int32_t* vals_init, int32_t* va,
                       int32_t* vb, int32_t* dst, int64_t n, int key_bits, hipStream_t st);

}  // namespace cw
