// Multi-scene batches, Morton ray ordering, hipGraph capture and the
// autotuner.
//
// No counterpart in the reference (which binds one scene per kernel in
// module globals, src/TraversalKernelCWBVH.cu:14-16, and traces it with one
// synchronous launch).  Here one persistent launch of the CWBVH kernel walks a
// device job table: the global chunk counter hands out `fetch_chunk`-ray
// chunks of any job to any wave (see cwbvh_trace<.., MULTI=true> in
// trace.hip), so batches of very uneven size share one launch without tails
// per scene.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "device_common.h"
#include "internal.h"

using namespace cwd;

namespace {

// Ray-order key (cw::kRayKeyBits bits): the ray's direction octant (the
// traversal's octant bits: x sign highest) above the Morton code of its origin
// quantized to kOriginBits per axis inside [lo, hi] (x bit highest in each
// triple).  Octant-major keys let a wave's rays share their child order as well
// as their nodes.  Measured on MI355X, C3 (20 M-triangle soup, 67 M shuffled
// diffuse rays, traversal only): 30-bit origin codes 51.96 ms CWBVH / 212.7 ms
// BVH2, 24-bit origin codes 52.06 ms, octant + 21-bit origin 50.92 / 195.4 ms,
// octant + 27-bit origin 50.77 / 195.0 ms; the 24-bit key also sorts in three
// 8-bit passes instead of four.
constexpr int kOriginBits = 7;
static_assert(3 + 3 * kOriginBits == cw::kRayKeyBits, "key width");

// bits 0..9 of v spread to every third bit
__device__ __forceinline__ unsigned spread10(unsigned v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ unsigned ray_key(float4 o, float4 d, float3 lo, float3 scale) {
    constexpr float top = (float)((1 << kOriginBits) - 1);
    const float fx = fminf(fmaxf((o.x - lo.x) * scale.x, 0.f), top);
    const float fy = fminf(fmaxf((o.y - lo.y) * scale.y, 0.f), top);
    const float fz = fminf(fmaxf((o.z - lo.z) * scale.z, 0.f), top);
    const unsigned m = (spread10((unsigned)fx) << 2) | (spread10((unsigned)fy) << 1) | spread10((unsigned)fz);
    const unsigned oct = ((d.x < 0.f ? 1u : 0u) << 2) | ((d.y < 0.f ? 1u : 0u) << 1) | (d.z < 0.f ? 1u : 0u);
    return (oct << (3 * kOriginBits)) | m;
}

__global__ void morton_keys_single(const Ray4* __restrict__ rays, int n, float3 lo, float3 scale,
                                   unsigned* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ray_key(rays[i].o, rays[i].d, lo, scale);
}

struct MortonJob {
    const Ray4* rays;
    long long first;    // global index of the job's first ray
    int n;
    float lo[3];
    float scale[3];
};

// One thread per ray of the whole batch; key = job index (high bits) : ray key.
__global__ void morton_keys_batch(const MortonJob* __restrict__ jobs, int num_jobs, long long total,
                                  unsigned long long* __restrict__ keys, int* __restrict__ vals) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int lo = 0, hi = num_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= g) lo = mid;
        else hi = mid - 1;
    }
    const MortonJob& J = jobs[lo];
    const int i = (int)(g - J.first);
    const unsigned k = ray_key(J.rays[i].o, J.rays[i].d, make_float3(J.lo[0], J.lo[1], J.lo[2]),
                               make_float3(J.scale[0], J.scale[1], J.scale[2]));
    keys[g] = ((unsigned long long)lo << cw::kRayKeyBits) | k;
    vals[g] = i;
}

float3 scale_of(const float* lo, const float* hi) {
    float s[3];
    for (int k = 0; k < 3; ++k) {
        const float ext = hi[k] - lo[k];
        s[k] = ext > 0.f ? (float)(1 << kOriginBits) / ext : 0.f;
    }
    return make_float3(s[0], s[1], s[2]);
}

int bits_for(int count) {
    int b = 0;
    while ((1 << b) < count) ++b;
    return b;
}

}  // namespace

struct cwbvh_batch {
    std::vector<cw::BatchJob> jobs;
    std::vector<int64_t> ray_first;   // per job, in the concatenated ray space
    cw::BatchJob* d_jobs = nullptr;
    unsigned* counters = nullptr;
    uint32_t total_chunks = 0;
    int64_t total_rays = 0;
    uint32_t flags = 0;
    cwbvh_trace_config cfg{};
    int32_t morton = 0;
    // Morton pre-pass buffers: keys (two), values (the keys kernel's job-local
    // ray indices, a ping-pong buffer, the sorted order) and the sort's counts
    MortonJob* d_mjobs = nullptr;
    unsigned long long* d_keys[2] = {nullptr, nullptr};
    int32_t* d_vals[3] = {nullptr, nullptr, nullptr};
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int key_bits = cw::kRayKeyBits;
    // deepest stack need over the jobs' trees; -1 when a tree's is unknown
    int32_t max_need = 0;
    int64_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Per (device, stream) device workspace of the Morton pre-pass, grown on
// demand and kept for later calls: a call on a stream reuses it once the
// stream's earlier work is ordered before it, so only growth allocates.
// Released by cwbvh_stream_destroy (or at process end).  Callers hold
// g_scratch_mu from lookup until their kernels are enqueued, so a buffer is
// never freed between a call's lookup and its launches.
std::mutex g_scratch_mu;
std::map<std::pair<int, hipStream_t>, std::pair<char*, size_t>> g_scratch;
// cwbvh_set_workspace_limit: largest workspace one stream may hold (0 = no limit)
size_t g_scratch_limit = 0;

int32_t stream_scratch(hipStream_t st, size_t bytes, char** out) {
    if (st == hipStreamPerThread) {
        cw::set_error("hipStreamPerThread is not supported: pass a stream handle or NULL");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    if (g_scratch_limit > 0 && bytes > g_scratch_limit) {
        cw::set_error("sort workspace of " + std::to_string(bytes) + " bytes exceeds the limit of " +
                  std::to_string(g_scratch_limit));
        return CWBVH_ERR_OUT_OF_MEMORY;
    }
    int dev = 0;
    CW_HIP(hipGetDevice(&dev));
    auto& slot = g_scratch[std::make_pair(dev, st)];
    if (slot.second < bytes) {
        if (slot.first) {
            (void)hipFree(slot.first);   // waits for the device, so no queued call still uses it
            slot = {nullptr, 0};
        }
        // 1/8 headroom so slowly growing calls do not reallocate every time
        size_t want = bytes + bytes / 8;
        if (g_scratch_limit > 0) want = std::min(want, g_scratch_limit);
        CW_HIP(hipMalloc(&slot.first, want));
        slot.second = want;
    }
    *out = slot.first;
    return CWBVH_OK;
}

// Origin Morton order of n rays on stream st, in the stream's workspace
// (g_scratch_mu held by the caller): keys, the other key buffer, two value
// buffers, the sort's counts, and, when d_order is null, the order itself
// (*in_scratch).
int32_t morton_sort(const void* d_rays, int64_t n, const float* lo3, const float* hi3, hipStream_t st,
                    int32_t* d_order, int32_t** in_scratch) {
    const size_t col = align256((size_t)n * 4);
    const int cols = d_order ? 4 : 5;
    const size_t sort_ws = align256(cw::radix_sort_workspace(n, cw::kRayKeyBits));
    const size_t need = cols * col + sort_ws;
    char* base = nullptr;
    int32_t rc = stream_scratch(st, need, &base);
    if (rc) return rc;
    if (!d_order) {
        d_order = reinterpret_cast<int32_t*>(base + 4 * col);
        *in_scratch = d_order;
    }
    return cw::morton_order_into(d_rays, n, lo3, hi3, reinterpret_cast<uint32_t*>(base),
                                 reinterpret_cast<uint32_t*>(base + col), reinterpret_cast<int32_t*>(base + 2 * col),
                                 reinterpret_cast<int32_t*>(base + 3 * col), d_order, base + cols * col, sort_ws, st);
}

int32_t batch_sort(cwbvh_batch* b, hipStream_t st) {
    if (!b->morton || b->total_rays == 0) return CWBVH_OK;
    const int block = 256;
    const long long grid = (b->total_rays + block - 1) / block;
    hipLaunchKernelGGL(morton_keys_batch, dim3((unsigned)grid), dim3(block), 0, st, b->d_mjobs, (int)b->jobs.size(),
                       (long long)b->total_rays, b->d_keys[0], b->d_vals[0]);
    CW_HIP(hipGetLastError());
    // values start as the job-local ray indices the keys kernel wrote
    return cw::radix_sort_u64(b->d_sort_tmp, b->sort_tmp_bytes, (uint64_t*)b->d_keys[0], (uint64_t*)b->d_keys[1],
                              b->d_vals[0], b->d_vals[2], b->d_vals[0], b->d_vals[1], b->total_rays, b->key_bits, st);
}

int32_t batch_trace(cwbvh_batch* b, hipStream_t st) {
    if (b->total_chunks == 0) return CWBVH_OK;
    cw::BatchLaunch L{};
    L.d_jobs = b->d_jobs;
    L.num_jobs = (int32_t)b->jobs.size();
    L.total_chunks = b->total_chunks;
    L.flags = b->flags;
    L.cfg = b->cfg;
    L.stats = nullptr;
    L.counters = b->counters;
    L.stream = st;
    L.lds_only = b->max_need > 0 && b->max_need <= b->cfg.lds_depth && b->cfg.lds_depth <= CWBVH_STACK_CAPACITY;
    return cw::launch_cwbvh_batch(L);
}

}  // namespace

using namespace cw;

namespace cw {

int32_t morton_order_into(const void* d_rays, int64_t n, const float lo3[3], const float hi3[3], uint32_t* k0,
                          uint32_t* k1, int32_t* va, int32_t* vb, int32_t* order, void* ws, size_t ws_bytes,
                          hipStream_t st) {
    if (n <= 0) return CWBVH_OK;
    const int block = 256;
    hipLaunchKernelGGL(morton_keys_single, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, st,
                       (const Ray4*)d_rays, (int)n, make_float3(lo3[0], lo3[1], lo3[2]), scale_of(lo3, hi3), k0);
    CW_HIP(hipGetLastError());
    return radix_sort_u32(ws, ws_bytes, k0, k1, nullptr, va, vb, order, n, kRayKeyBits, st);
}

std::unique_lock<std::mutex> lock_stream_scratch() { return std::unique_lock<std::mutex>(g_scratch_mu); }

int32_t morton_order_scratch(const void* d_rays, int64_t n, const float lo[3], const float hi[3], hipStream_t st,
                             const int32_t** order) {
    int32_t* out = nullptr;
    int32_t rc = morton_sort(d_rays, n, lo, hi, st, nullptr, &out);
    *order = out;
    return rc;
}

void release_stream_scratch(hipStream_t st) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch.find(std::make_pair(dev, st));
    if (it == g_scratch.end()) return;
    if (it->second.first) (void)hipFree(it->second.first);
    g_scratch.erase(it);
}

}  // namespace cw

extern "C" {

int32_t cwbvh_scene_bounds(const cwbvh_scene* s, float* lo3, float* hi3) {
    CW_CHECK_ARG(s && lo3 && hi3, "NULL argument");
    for (int k = 0; k < 3; ++k) {
        lo3[k] = s->lo[k];
        hi3[k] = s->hi[k];
    }
    return CWBVH_OK;
}

int32_t cwbvh_set_workspace_limit(int64_t bytes) {
    CW_CHECK_ARG(bytes >= 0, "limit must be >= 0");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    g_scratch_limit = (size_t)bytes;
    // workspaces already above the new limit are released (hipFree waits for
    // the work queued on them)
    for (auto it = g_scratch.begin(); it != g_scratch.end();) {
        if (bytes > 0 && it->second.second > (size_t)bytes) {
            if (it->second.first) (void)hipFree(it->second.first);
            it = g_scratch.erase(it);
        } else {
            ++it;
        }
    }
    return CWBVH_OK;
}

int32_t cwbvh_morton_order(const cwbvh_ray* d_rays, int64_t n, const float* lo3, const float* hi3, int32_t* d_order,
                           void* stream) {
    CW_CHECK_ARG(n >= 0 && n <= 0x7FFFFFFF, "ray count must be in [0, 2^31)");
    CW_CHECK_ARG(n == 0 || (d_rays && d_order && lo3 && hi3), "NULL argument");
    if (n == 0) return CWBVH_OK;
    auto lock = lock_stream_scratch();
    return morton_sort(d_rays, n, lo3, hi3, (hipStream_t)stream, d_order, nullptr);
}

int32_t cwbvh_batch_create(const cwbvh_batch_item* items, int32_t count, uint32_t flags,
                           const cwbvh_trace_config* cfg, int32_t morton, cwbvh_batch** out) {
    CW_CHECK_ARG(out != nullptr, "out is NULL");
    CW_CHECK_ARG(count >= 0 && (count == 0 || items), "bad item array");
    CW_CHECK_ARG(count < (1 << 24), "too many items");
    CW_CHECK_ARG((flags & ~(uint32_t)(CWBVH_TRACE_ANY | CWBVH_TRACE_PRIM_ID)) == 0, "unsupported flags for a batch");
    cwbvh_trace_config c;
    if (cfg) c = *cfg;
    else cwbvh_trace_config_default(CWBVH_KIND_CWBVH, &c);
    int32_t rc = validate_config(CWBVH_KIND_CWBVH, c);
    if (rc) return rc;
    // the job table is planned in chunks: fix the chunk size here so that the
    // plan and the kernel use the same value
    if (c.fetch_chunk == 0) c.fetch_chunk = kBatchChunk;
    auto* b = new cwbvh_batch();
    b->flags = flags;
    b->cfg = c;
    b->morton = morton ? 1 : 0;
    uint64_t chunks = 0;
    std::vector<MortonJob> mj;
    for (int32_t i = 0; i < count; ++i) {
        const cwbvh_batch_item& it = items[i];
        if (!it.scene || it.scene->kind != CWBVH_KIND_CWBVH || it.n < 0 || it.n > 0x7FFFFFFF ||
            (it.n > 0 && (!it.d_rays || !it.d_hits)) ||
            (((uintptr_t)it.d_rays | (uintptr_t)it.d_hits) & 15u) != 0 ||
            ((flags & CWBVH_TRACE_PRIM_ID) && !it.scene->prim_idx)) {
            delete b;
            set_error("batch item " + std::to_string(i) +
                      " is invalid (needs a CWBVH scene and 16-byte aligned rays and hits)");
            return CWBVH_ERR_INVALID_ARGUMENT;
        }
        if (it.n == 0) continue;
        BatchJob J{};
        J.nodes = it.scene->nodes;
        J.woop = it.scene->woop;
        J.rays = it.d_rays;
        J.order = nullptr;
        J.hits = it.d_hits;
        J.prim_idx = it.scene->prim_idx;
        J.n = (int32_t)it.n;
        J.node_stride = it.scene->node_stride;
        J.chunk_begin = (uint32_t)chunks;
        J.spec_limit = speculation_limit(it.scene->stack_need);
        if (it.scene->stack_need <= 0) b->max_need = -1;
        else if (b->max_need >= 0) b->max_need = std::max(b->max_need, it.scene->stack_need);
        chunks += ((uint64_t)it.n + c.fetch_chunk - 1) / c.fetch_chunk;
        if (chunks > 0xFFFFFFF0ull) {
            delete b;
            set_error("batch too large for 32-bit chunk indices");
            return CWBVH_ERR_INVALID_ARGUMENT;
        }
        MortonJob m{};
        m.rays = (const Ray4*)it.d_rays;
        m.first = b->total_rays;
        m.n = (int)it.n;
        float3 sc = scale_of(it.scene->lo, it.scene->hi);
        for (int k = 0; k < 3; ++k) m.lo[k] = it.scene->lo[k];
        m.scale[0] = sc.x;
        m.scale[1] = sc.y;
        m.scale[2] = sc.z;
        mj.push_back(m);
        b->ray_first.push_back(b->total_rays);
        b->total_rays += it.n;
        b->jobs.push_back(J);
    }
    b->total_chunks = (uint32_t)chunks;
    // as the scene default does: with the default configuration, trees
    // deeper than its LDS depth keep their whole stack in LDS
    if (!cfg && b->max_need > b->cfg.lds_depth && b->max_need <= CWBVH_STACK_CAPACITY) b->cfg.lds_depth = b->max_need;
    if (b->morton && b->total_rays > 0x7FFFFFFF) {
        delete b;
        set_error("Morton pre-pass supports < 2^31 rays per batch");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    hipError_t e = hipSuccess;
    const size_t jb = std::max<size_t>(1, b->jobs.size()) * sizeof(BatchJob);
    e = hipMalloc(&b->d_jobs, jb);
    if (e == hipSuccess) e = hipMalloc(&b->counters, kWorkWords * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(b->counters, 0, kWorkWords * sizeof(unsigned));
    b->bytes = (int64_t)(jb + kWorkWords * sizeof(unsigned));
    if (e == hipSuccess && b->morton && b->total_rays > 0) {
        const size_t n = (size_t)b->total_rays;
        b->key_bits = cw::kRayKeyBits + bits_for((int)b->jobs.size());
        e = hipMalloc(&b->d_mjobs, mj.size() * sizeof(MortonJob));
        if (e == hipSuccess) e = hipMemcpy(b->d_mjobs, mj.data(), mj.size() * sizeof(MortonJob), hipMemcpyHostToDevice);
        for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipMalloc(&b->d_keys[k], n * sizeof(unsigned long long));
        for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMalloc(&b->d_vals[k], n * sizeof(int32_t));
        b->sort_tmp_bytes = radix_sort_workspace((int64_t)n, b->key_bits);
        if (e == hipSuccess) e = hipMalloc(&b->d_sort_tmp, b->sort_tmp_bytes);
        b->bytes += (int64_t)(mj.size() * sizeof(MortonJob) + n * 28 + b->sort_tmp_bytes);
        // job j reads its rays in the order stored in its slice of the sorted values
        for (size_t j = 0; j < b->jobs.size(); ++j) b->jobs[j].order = b->d_vals[1] + b->ray_first[j];
    }
    if (e == hipSuccess && !b->jobs.empty())
        e = hipMemcpy(b->d_jobs, b->jobs.data(), b->jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        cwbvh_batch_destroy(b);
        return hip_fail(e, "cwbvh_batch_create");
    }
    *out = b;
    return CWBVH_OK;
}

int32_t cwbvh_batch_launch(cwbvh_batch* b, void* stream) {
    CW_CHECK_ARG(b != nullptr, "batch is NULL");
    hipStream_t st = (hipStream_t)stream;
    int32_t rc = batch_sort(b, st);
    if (rc) return rc;
    return batch_trace(b, st);
}

int32_t cwbvh_batch_capture(cwbvh_batch* b, void* stream) {
    CW_CHECK_ARG(b != nullptr, "batch is NULL");
    CW_CHECK_ARG(stream != nullptr, "graph capture needs a non-null stream");
    hipStream_t st = (hipStream_t)stream;
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    // the persistent kernel resets its own fetch counters when its last wave
    // exits, so the graph holds the traversal launch alone (no memset node)
    CW_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int32_t rc = batch_trace(b, st);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    CW_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return CWBVH_OK;
}

int32_t cwbvh_batch_replay(cwbvh_batch* b, void* stream) {
    CW_CHECK_ARG(b != nullptr, "batch is NULL");
    if (!b->exec) {
        set_error("batch has no captured graph");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    hipStream_t st = (hipStream_t)stream;
    int32_t rc = batch_sort(b, st);
    if (rc) return rc;
    CW_HIP(hipGraphLaunch(b->exec, st));
    return CWBVH_OK;
}

int32_t cwbvh_batch_info(const cwbvh_batch* b, int64_t* total_rays, int32_t* num_jobs, int64_t* device_bytes) {
    CW_CHECK_ARG(b != nullptr, "batch is NULL");
    if (total_rays) *total_rays = b->total_rays;
    if (num_jobs) *num_jobs = (int32_t)b->jobs.size();
    if (device_bytes) *device_bytes = b->bytes;
    return CWBVH_OK;
}

void cwbvh_batch_destroy(cwbvh_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->counters) (void)hipFree(b->counters);
    if (b->d_mjobs) (void)hipFree(b->d_mjobs);
    for (int k = 0; k < 2; ++k)
        if (b->d_keys[k]) (void)hipFree(b->d_keys[k]);
    for (int k = 0; k < 3; ++k)
        if (b->d_vals[k]) (void)hipFree(b->d_vals[k]);
    if (b->d_sort_tmp) (void)hipFree(b->d_sort_tmp);
    delete b;
}

int32_t cwbvh_autotune_batch(const cwbvh_batch_item* items, int32_t count, uint32_t flags,
                             const cwbvh_tune_space* space, int32_t reps, cwbvh_tune_record* records,
                             int32_t capacity, int32_t* n_records, cwbvh_tune_record* best, void* stream) {
    CW_CHECK_ARG(space != nullptr && best != nullptr, "NULL argument");
    CW_CHECK_ARG(space->n_block >= 1 && space->n_block <= 4 && space->n_lds >= 1 && space->n_lds <= 4 &&
                     space->n_morton >= 1 && space->n_morton <= 2 && space->n_prefetch >= 1 && space->n_prefetch <= 2 &&
                     space->n_loop_modes >= 0 && space->n_loop_modes <= 3,
                 "bad tuning space");
    CW_CHECK_ARG(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CW_HIP(hipEventCreate(&e0));
    CW_HIP(hipEventCreate(&e1));
    // lds_depths entry 0: the deepest stack need of the items' trees
    int32_t need = 0;
    for (int32_t i = 0; i < count && items; ++i) {
        const cwbvh_scene* sc = items[i].scene;
        if (!sc || sc->stack_need <= 0) {
            need = -1;
            break;
        }
        need = std::max(need, sc->stack_need);
    }
    int32_t n = 0;
    best->ms = INFINITY;
    int32_t rc = CWBVH_OK;
    for (int ib = 0; ib < space->n_block && rc == CWBVH_OK; ++ib)
        for (int il = 0; il < space->n_lds && rc == CWBVH_OK; ++il)
            for (int im = 0; im < space->n_morton && rc == CWBVH_OK; ++im)
                for (int ip = 0; ip < space->n_prefetch && rc == CWBVH_OK; ++ip)
                for (int io = 0; io < std::max(1, space->n_loop_modes) && rc == CWBVH_OK; ++io) {
                    cwbvh_trace_config c;
                    cwbvh_trace_config_default(CWBVH_KIND_CWBVH, &c);
                    c.block_size = space->block_sizes[ib];
                    c.lds_depth = space->lds_depths[il];
                    if (c.lds_depth == 0) c.lds_depth = need > 0 && need <= CWBVH_STACK_CAPACITY ? need : 10;
                    c.prefetch = space->prefetch[ip];
                    if (space->n_loop_modes > 0) c.loop_mode = space->loop_modes[io];
                    cwbvh_batch* b = nullptr;
                    rc = cwbvh_batch_create(items, count, flags, &c, space->morton[im], &b);
                    if (rc) break;
                    std::vector<float> t;
                    rc = cwbvh_batch_launch(b, st);  // warm-up
                    for (int r = 0; r < reps && rc == CWBVH_OK; ++r) {
                        (void)hipEventRecord(e0, st);
                        rc = cwbvh_batch_launch(b, st);
                        (void)hipEventRecord(e1, st);
                        hipError_t he = hipEventSynchronize(e1);
                        if (he != hipSuccess) rc = hip_fail(he, "autotune sync");
                        float ms = 0.f;
                        (void)hipEventElapsedTime(&ms, e0, e1);
                        t.push_back(ms);
                    }
                    cwbvh_batch_destroy(b);
                    if (rc) break;
                    std::sort(t.begin(), t.end());
                    cwbvh_tune_record rec{};
                    rec.cfg = c;
                    rec.morton = space->morton[im];
                    rec.ms = t[t.size() / 2];
                    if (records && n < capacity) records[n] = rec;
                    ++n;
                    if (rec.ms < best->ms) *best = rec;
                }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (n_records) *n_records = n;
    return rc;
}

}  // extern "C"
