// C ABI: errors, device plumbing, device scenes and the trace entry points.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <sched.h>

#include <thread>

#include "internal.h"

namespace cw {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }
const char* get_error() { return g_error.c_str(); }

int32_t hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    // the error is reported here: clear it from the runtime's last error so
    // that the launch check of a later call does not report it again
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? CWBVH_ERR_OUT_OF_MEMORY
                                    : (e == hipErrorNoDevice ? CWBVH_ERR_NO_DEVICE : CWBVH_ERR_HIP);
}

// Host worker threads by default: the CPUs this process may run on (its
// affinity mask, which a shared machine narrows to the job's share), at most
// 64; hardware_concurrency() when the mask cannot be read.
int num_threads_or_default(int requested) {
    if (requested > 0) return requested;
    int n = 0;
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

int device_compute_units() {
    static int cus = 0;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        else
            (void)hipGetLastError();  // tolerated: fall back to one unit
        if (cus <= 0) cus = 1;
    });
    return cus;
}

// Keyed by (current device, stream): the null stream names a different queue
// on each device, and each device needs counters in its own memory.  The
// per-thread handle names a different stream on every host thread, so one key
// cannot stand for it: it is refused.
int32_t stream_counters(hipStream_t stream, unsigned** out) {
    static std::mutex mu;
    static std::map<std::pair<int, hipStream_t>, unsigned*> table;
    if (stream == hipStreamPerThread) {
        set_error("hipStreamPerThread is not supported: pass a stream handle or NULL");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    int dev = 0;
    CW_HIP(hipGetDevice(&dev));
    const std::pair<int, hipStream_t> key(dev, stream);
    std::lock_guard<std::mutex> g(mu);
    auto it = table.find(key);
    if (it != table.end()) {
        *out = it->second;
        return CWBVH_OK;
    }
    unsigned* p = nullptr;
    CW_HIP(hipMalloc(&p, kWorkWords * sizeof(unsigned)));
    CW_HIP(hipMemset(p, 0, kWorkWords * sizeof(unsigned)));
    table[key] = p;
    *out = p;
    return CWBVH_OK;
}

}  // namespace cw

using namespace cw;

extern "C" {

int32_t cwbvh_abi_version(void) { return CWBVH_ABI_VERSION; }

const char* cwbvh_status_string(int32_t s) {
    switch (s) {
        case CWBVH_OK: return "ok";
        case CWBVH_ERR_INVALID_ARGUMENT: return "invalid argument";
        case CWBVH_ERR_HIP: return "HIP runtime error";
        case CWBVH_ERR_OUT_OF_MEMORY: return "out of device memory";
        case CWBVH_ERR_STACK_DEPTH: return "tree deeper than the traversal stack";
        case CWBVH_ERR_NOT_BOUND: return "no BVH bound";
        case CWBVH_ERR_NO_DEVICE: return "no HIP device";
        case CWBVH_ERR_BUDGET: return "request exceeds the memory budget";
        case CWBVH_ERR_CANCELLED: return "cancelled";
        case CWBVH_ERR_INTERNAL: return "internal error";
        case CWBVH_ERR_IO: return "file error";
        default: return "unknown status";
    }
}

const char* cwbvh_last_error(void) { return cw::get_error(); }

// ------------------------------------------------------------ plumbing

int32_t cwbvh_device_count(int32_t* count) {
    CW_CHECK_ARG(count != nullptr, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) n = 0;
    else if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    *count = n;
    return CWBVH_OK;
}

int32_t cwbvh_get_device_info(int32_t device, cwbvh_device_info* info) {
    CW_CHECK_ARG(info != nullptr, "info is NULL");
    hipDeviceProp_t p;
    CW_HIP(hipGetDeviceProperties(&p, device));
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->name, sizeof(info->name), "%.63s", p.name);
    std::snprintf(info->arch, sizeof(info->arch), "%.31s", p.gcnArchName);
    info->compute_units = p.multiProcessorCount;
    info->wave_size = p.warpSize;
    info->lds_per_block = (int64_t)p.sharedMemPerBlock;
    info->l2_bytes = (int64_t)p.l2CacheSize;
    info->global_mem_bytes = (int64_t)p.totalGlobalMem;
    info->clock_khz = p.clockRate;
    info->mem_clock_khz = p.memoryClockRate;
    info->mem_bus_width = p.memoryBusWidth;
    return CWBVH_OK;
}

int32_t cwbvh_malloc(void** dptr, int64_t bytes) {
    CW_CHECK_ARG(dptr != nullptr && bytes >= 0, "bad malloc arguments");
    *dptr = nullptr;
    if (bytes == 0) return CWBVH_OK;
    CW_HIP(hipMalloc(dptr, (size_t)bytes));
    return CWBVH_OK;
}

int32_t cwbvh_free(void* dptr) {
    if (dptr) CW_HIP(hipFree(dptr));
    return CWBVH_OK;
}

int32_t cwbvh_memcpy_htod(void* dst, const void* src, int64_t bytes) {
    CW_CHECK_ARG(bytes >= 0 && (bytes == 0 || (dst && src)), "bad copy arguments");
    if (bytes) CW_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return CWBVH_OK;
}

int32_t cwbvh_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
    CW_CHECK_ARG(bytes >= 0 && (bytes == 0 || (dst && src)), "bad copy arguments");
    if (bytes) CW_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return CWBVH_OK;
}

int32_t cwbvh_memcpy_dtod(void* dst, const void* src, int64_t bytes) {
    CW_CHECK_ARG(bytes >= 0 && (bytes == 0 || (dst && src)), "bad copy arguments");
    if (bytes) CW_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice));
    return CWBVH_OK;
}

int32_t cwbvh_memset(void* dptr, int32_t value, int64_t bytes) {
    CW_CHECK_ARG(bytes >= 0 && (bytes == 0 || dptr), "bad memset arguments");
    if (bytes) CW_HIP(hipMemset(dptr, value, (size_t)bytes));
    return CWBVH_OK;
}

int32_t cwbvh_stream_create(void** stream) {
    CW_CHECK_ARG(stream != nullptr, "stream is NULL");
    hipStream_t s;
    CW_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return CWBVH_OK;
}

int32_t cwbvh_stream_destroy(void* stream) {
    if (!stream) return CWBVH_OK;
    cw::release_stream_scratch((hipStream_t)stream);
    CW_HIP(hipStreamDestroy((hipStream_t)stream));
    return CWBVH_OK;
}

int32_t cwbvh_stream_synchronize(void* stream) {
    CW_HIP(hipStreamSynchronize((hipStream_t)stream));
    return CWBVH_OK;
}

int32_t cwbvh_device_synchronize(void) {
    CW_HIP(hipDeviceSynchronize());
    return CWBVH_OK;
}

int32_t cwbvh_event_create(void** event) {
    CW_CHECK_ARG(event != nullptr, "event is NULL");
    hipEvent_t e;
    CW_HIP(hipEventCreate(&e));
    *event = e;
    return CWBVH_OK;
}

int32_t cwbvh_event_destroy(void* event) {
    if (event) CW_HIP(hipEventDestroy((hipEvent_t)event));
    return CWBVH_OK;
}

int32_t cwbvh_event_record(void* event, void* stream) {
    CW_CHECK_ARG(event != nullptr, "event is NULL");
    CW_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return CWBVH_OK;
}

int32_t cwbvh_event_synchronize(void* event) {
    CW_CHECK_ARG(event != nullptr, "event is NULL");
    CW_HIP(hipEventSynchronize((hipEvent_t)event));
    return CWBVH_OK;
}

int32_t cwbvh_event_elapsed_ms(void* start, void* stop, float* ms) {
    CW_CHECK_ARG(start && stop && ms, "NULL argument");
    CW_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return CWBVH_OK;
}

// ------------------------------------------------------------ scenes

// Scene bounds from the root node (its first 5 float4 for CWBVH, 4 for
// BVH2): decoded child boxes (CWBVH) or the two child boxes (BVH2); empty or
// inverted boxes are skipped.  False (and the unit cube) when none is left.
static bool root_bounds(int32_t kind, const cwbvh_float4* root, float lo[3], float hi[3]) {
    float a[3] = {INFINITY, INFINITY, INFINITY}, b[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (kind == CWBVH_KIND_CWBVH) {
        const uint8_t* n0 = reinterpret_cast<const uint8_t*>(&root[0]);
        const uint8_t* meta = reinterpret_cast<const uint8_t*>(&root[1]) + 8;
        const uint8_t* q = reinterpret_cast<const uint8_t*>(&root[2]);
        const float p[3] = {root[0].x, root[0].y, root[0].z};
        for (int s = 0; s < 8; ++s) {
            if (!meta[s]) continue;
            for (int k = 0; k < 3; ++k) {
                const double sc = std::ldexp(1.0, (int)(int8_t)n0[12 + k]);
                a[k] = std::min(a[k], (float)(p[k] + q[k * 8 + s] * sc));
                b[k] = std::max(b[k], (float)(p[k] + q[24 + k * 8 + s] * sc));
            }
        }
    } else if (kind == CWBVH_KIND_BVH2) {
        const cwbvh_float4* n = root;
        const float blo[2][3] = {{n[0].x, n[0].z, n[2].x}, {n[1].x, n[1].z, n[2].z}};
        const float bhi[2][3] = {{n[0].y, n[0].w, n[2].y}, {n[1].y, n[1].w, n[2].w}};
        for (int c = 0; c < 2; ++c) {
            if (!(blo[c][0] <= bhi[c][0] && blo[c][1] <= bhi[c][1] && blo[c][2] <= bhi[c][2])) continue;
            for (int k = 0; k < 3; ++k) {
                a[k] = std::min(a[k], blo[c][k]);
                b[k] = std::max(b[k], bhi[c][k]);
            }
        }
    }
    bool known = true;
    for (int k = 0; k < 3; ++k) known = known && a[k] <= b[k] && std::isfinite(a[k]) && std::isfinite(b[k]);
    for (int k = 0; k < 3; ++k) {
        lo[k] = known ? a[k] : 0.f;
        hi[k] = known ? b[k] : 1.f;
    }
    return known;
}

int32_t cwbvh_scene_create(const cwbvh_layout* L, void* stream, cwbvh_scene** out) {
    return cwbvh_scene_create_ex(L, stream, CWBVH_SCENE_STRIDE_AUTO, out);
}

int32_t cwbvh_scene_create_ex(const cwbvh_layout* L, void* stream, int32_t stride_mode, cwbvh_scene** out) {
    CW_CHECK_ARG(L != nullptr && out != nullptr, "NULL argument");
    CW_CHECK_ARG(!L->nodes.empty(), "empty node array");
    CW_CHECK_ARG(stride_mode >= CWBVH_SCENE_STRIDE_AUTO && stride_mode <= CWBVH_SCENE_STRIDE_128,
                 "stride_mode must be CWBVH_SCENE_STRIDE_AUTO, _KEEP or _128");
    const int cap = L->kind == CWBVH_KIND_CWBVH ? CWBVH_STACK_CAPACITY : CWBVH_BVH2_STACK_CAPACITY;
    if (L->stack_need > cap) {
        set_error("tree needs " + std::to_string(L->stack_need) + " stack entries, kernel holds " +
                  std::to_string(cap));
        return CWBVH_ERR_STACK_DEPTH;
    }
    // 80 B nodes re-laid at 128 B (same bytes, 48 B of zeros behind each)
    const cwbvh_float4* node_src = L->nodes.data();
    size_t node_f4 = L->nodes.size();
    int32_t stride = L->node_stride;
    std::vector<cwbvh_float4> relaid;
    if (L->kind == CWBVH_KIND_CWBVH && L->node_stride == 5) {
        const int64_t streamed = (int64_t)(L->nodes.size() + L->woop.size()) * (int64_t)sizeof(cwbvh_float4);
        if (stride_mode == CWBVH_SCENE_STRIDE_128 ||
            (stride_mode == CWBVH_SCENE_STRIDE_AUTO && streamed > kCacheResidentBytes)) {
            const int64_t nn = (int64_t)L->nodes.size() / 5;
            relaid.assign((size_t)nn * 8, cwbvh_float4{0.f, 0.f, 0.f, 0.f});
            parallel_for(nn, num_threads_or_default(0), [&](int64_t b, int64_t e) {
                for (int64_t i = b; i < e; ++i)
                    std::memcpy(&relaid[(size_t)i * 8], &L->nodes[(size_t)i * 5], 5 * sizeof(cwbvh_float4));
            });
            node_src = relaid.data();
            node_f4 = relaid.size();
            stride = 8;
        }
    }
    auto* s = new cwbvh_scene();
    s->kind = L->kind;
    s->node_stride = stride;
    s->stack_need = L->stack_need;
    s->owning = true;
    if (L->nodes.size() >= (size_t)(L->kind == CWBVH_KIND_CWBVH ? 5 : 4))
        s->bounds_known = root_bounds(L->kind, L->nodes.data(), s->lo, s->hi);
    s->woop_f4 = (int64_t)L->woop.size();
    hipGetDevice(&s->device);
    const size_t nb = node_f4 * sizeof(cwbvh_float4);
    // two float4 of zero padding after the Woop rows keep every triangle-sized
    // read of the last leaf inside the allocation
    const size_t wb = (L->woop.size() + 2) * sizeof(cwbvh_float4);
    const size_t pb = std::max<size_t>(1, L->prim_idx.size()) * sizeof(int32_t);
    void *dn = nullptr, *dw = nullptr, *dp = nullptr;
    hipError_t e = hipMalloc(&dn, nb);
    if (e == hipSuccess) e = hipMalloc(&dw, wb);
    if (e == hipSuccess) e = hipMalloc(&dp, pb);
    hipStream_t st = (hipStream_t)stream;
    if (e == hipSuccess) e = hipMemcpyAsync(dn, node_src, nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(dw, 0, wb, st);
    if (e == hipSuccess && !L->woop.empty())
        e = hipMemcpyAsync(dw, L->woop.data(), L->woop.size() * sizeof(cwbvh_float4), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !L->prim_idx.empty())
        e = hipMemcpyAsync(dp, L->prim_idx.data(), L->prim_idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        // copies queued before the failing call read host vectors freed on return
        (void)hipStreamSynchronize(st);
        if (dn) hipFree(dn);
        if (dw) hipFree(dw);
        if (dp) hipFree(dp);
        delete s;
        return hip_fail(e, "scene upload");
    }
    s->nodes = dn;
    s->woop = dw;
    s->prim_idx = (const int32_t*)dp;
    s->bytes = (int64_t)(nb + wb + pb);
    *out = s;
    return CWBVH_OK;
}

int32_t cwbvh_scene_wrap_device(int32_t kind, int32_t node_stride, const void* d_nodes, const void* d_woop,
                                const int32_t* d_prim_idx, int32_t stack_need, cwbvh_scene** out) {
    CW_CHECK_ARG(out != nullptr && d_nodes != nullptr && d_woop != nullptr, "NULL argument");
    CW_CHECK_ARG(kind == CWBVH_KIND_BVH2 ? node_stride == 4
                                         : (kind == CWBVH_KIND_CWBVH && (node_stride == 5 || node_stride == 8)),
                 "bad kind / node stride");
    const int cap = kind == CWBVH_KIND_CWBVH ? CWBVH_STACK_CAPACITY : CWBVH_BVH2_STACK_CAPACITY;
    if (stack_need > cap) {
        set_error("tree deeper than the traversal stack");
        return CWBVH_ERR_STACK_DEPTH;
    }
    auto* s = new cwbvh_scene();
    s->kind = kind;
    s->node_stride = node_stride;
    s->nodes = d_nodes;
    s->woop = d_woop;
    s->prim_idx = d_prim_idx;
    s->stack_need = stack_need;
    s->owning = false;
    hipGetDevice(&s->device);
    // bounds for the Morton ray pre-pass from the root node (a failed read
    // leaves them unknown and the pre-pass off)
    cwbvh_float4 root[5];
    const int root_f4 = kind == CWBVH_KIND_CWBVH ? 5 : 4;
    if (hipMemcpy(root, d_nodes, root_f4 * sizeof(cwbvh_float4), hipMemcpyDeviceToHost) == hipSuccess)
        s->bounds_known = root_bounds(kind, root, s->lo, s->hi);
    // valid extent of the Woop array: up to the end of its allocation
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)d_woop) == hipSuccess && size > 0) {
        size_t off = (size_t)((const char*)d_woop - (const char*)base);
        s->woop_f4 = (int64_t)((size - off) / sizeof(cwbvh_float4));
        s->bytes += (int64_t)(size - off);
    }
    // node bytes: from the node pointer to the end of its allocation (an upper bound)
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)d_nodes) == hipSuccess && size > 0)
        s->bytes += (int64_t)(size - (size_t)((const char*)d_nodes - (const char*)base));
    hipGetLastError();
    *out = s;
    return CWBVH_OK;
}

int32_t cwbvh_scene_device_bytes(const cwbvh_scene* s, int64_t* bytes) {
    CW_CHECK_ARG(s != nullptr && bytes != nullptr, "NULL argument");
    *bytes = s->bytes;
    return CWBVH_OK;
}

int32_t cwbvh_scene_kind(const cwbvh_scene* s, int32_t* kind) {
    CW_CHECK_ARG(s != nullptr && kind != nullptr, "NULL argument");
    *kind = s->kind;
    return CWBVH_OK;
}

int32_t cwbvh_scene_node_stride(const cwbvh_scene* s, int32_t* stride) {
    CW_CHECK_ARG(s != nullptr && stride != nullptr, "NULL argument");
    *stride = s->node_stride;
    return CWBVH_OK;
}

void cwbvh_scene_destroy(cwbvh_scene* s) {
    if (!s) return;
    if (s->owning) {
        hipFree(const_cast<void*>(s->nodes));
        hipFree(const_cast<void*>(s->woop));
        hipFree(const_cast<int32_t*>(s->prim_idx));
    }
    delete s;
}

// ------------------------------------------------------------- trace

int32_t cwbvh_trace_config_default(int32_t kind, cwbvh_trace_config* c) {
    CW_CHECK_ARG(c != nullptr, "cfg is NULL");
    CW_CHECK_ARG(kind == CWBVH_KIND_BVH2 || kind == CWBVH_KIND_CWBVH, "bad kind");
    std::memset(c, 0, sizeof(*c));
    if (kind == CWBVH_KIND_CWBVH) {
        // measured on MI355X (C1, 33M AO rays; C0, 1M): 10 LDS entries cover
        // the stack of trees up to depth 9 at full occupancy; 128..512-ray
        // claims spread over 8 counters keep the fetch atomics off the
        // critical path (C1 +2%, C0 +46% over one counter); the interleaved
        // schedule is 1.36x the grouped one and keeps the reference's
        // per-ray order (speculation adds under 3%)
        c->block_size = 128;
        c->lds_depth = 10;
        c->grid_blocks = 0;
        c->fetch_chunk = 0;   // automatic
        c->refill_idle = 4;
        c->refill_budget = 16;
        c->postpone_pct = 0;
        c->prefetch = 0;
        c->work_slices = 8;
        c->persistent = 2;    // automatic by ray count
        c->loop_mode = 1;
        c->occupancy = 0;     // automatic by ray count
    } else {
        c->block_size = 256;
        c->lds_depth = 8;
        c->work_slices = 1;
    }
    return CWBVH_OK;
}

int32_t cwbvh_trace_config_for_scene(const cwbvh_scene* scene, cwbvh_trace_config* c) {
    CW_CHECK_ARG(scene != nullptr && c != nullptr, "NULL argument");
    int32_t rc = cwbvh_trace_config_default(scene->kind, c);
    if (rc) return rc;
    // measured on MI355X: a tree that fits the 256 MB Infinity Cache is
    // VALU-bound and gains from the interleaved schedule (C1: 1.36x).  A
    // larger one is bound by HBM traffic.  Sorting a launch's rays by
    // direction octant and origin first lets neighbouring waves share the
    // nodes they fetch and their child order (C3, 1.7 GB, 67 M shuffled rays:
    // the 2.0 ms sort and the sorted trace together are 1.23x the trace in the
    // given order), and sorted rays run best with the interleaved schedule and
    // node prefetch, whose loads of the next node overlap the triangle tests
    // (C3 closest hit on sorted rays: 49.6 ms, against 50.7 speculative with
    // prefetch, 56.0 grouped, 58.5 interleaved without prefetch; block 256
    // and deeper LDS stacks add nothing).  A tree without bounds (an empty root) cannot
    // quantize origins; its unsorted rays keep the grouped schedule, which
    // tests a leaf's triangles right after fetching the node (C3 shuffled:
    // 1.10x over interleaved).
    // a tree whose stack need exceeds the default LDS depth keeps its whole
    // stack in LDS (the kernels without the scratch part serve it)
    if (scene->kind == CWBVH_KIND_CWBVH && scene->stack_need > c->lds_depth && scene->stack_need <= CWBVH_STACK_CAPACITY)
        c->lds_depth = scene->stack_need;
    if (scene->bytes > kCacheResidentBytes) {
        if (scene->bounds_known) c->ray_order = 1;
        if (scene->kind == CWBVH_KIND_CWBVH) {
            c->loop_mode = scene->bounds_known ? 1 : 0;
            c->prefetch = scene->bounds_known ? 1 : 0;
        } else {
            // BVH2 on a tree that streams from HBM: the persistent grid keeps
            // lanes busy while the slowest rays of a wave finish (20 M-triangle
            // soup, 67 M sorted rays: 176 ms against 197 one ray per thread;
            // on a cache-resident 1 M-triangle scene it gains 4.5% any hit and
            // loses 1% closest hit)
            c->persistent = 2;
            c->refill_idle = 2;
            c->refill_budget = 8;
            c->work_slices = 8;
        }
    }
    return CWBVH_OK;
}

// Rays and hits are read and written as 16-byte vectors.
static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // extern "C"

namespace cw {

int32_t prepare_trace(const cwbvh_scene* scene, const cwbvh_ray* d_rays, const int32_t* d_order, cwbvh_hit* d_hits,
                      int64_t n, uint32_t flags, const cwbvh_trace_config* cfg, uint64_t* d_counters, void* stream,
                      TraceLaunch* out) {
    CW_CHECK_ARG(scene != nullptr, "scene is NULL");
    CW_CHECK_ARG(n >= 0 && n <= 0x7FFFFFFF, "ray count must be in [0, 2^31)");
    CW_CHECK_ARG(n == 0 || (d_rays != nullptr && d_hits != nullptr), "NULL ray or hit buffer");
    CW_CHECK_ARG(aligned16(d_rays) && aligned16(d_hits), "ray and hit buffers must be 16-byte aligned");
    CW_CHECK_ARG((flags & ~(uint32_t)(CWBVH_TRACE_ANY | CWBVH_TRACE_PRIM_ID | CWBVH_TRACE_STATS)) == 0, "unknown flags");
    CW_CHECK_ARG(!(flags & CWBVH_TRACE_STATS) || d_counters != nullptr, "STATS needs a counter buffer");
    CW_CHECK_ARG(!(flags & CWBVH_TRACE_PRIM_ID) || scene->prim_idx != nullptr, "PRIM_ID needs a prim index array");
    CW_CHECK_ARG((hipStream_t)stream != hipStreamPerThread,
                 "hipStreamPerThread is not supported: pass a stream handle or NULL");
    cwbvh_trace_config c;
    if (cfg) c = *cfg;
    else cwbvh_trace_config_for_scene(scene, &c);
    int32_t rc = validate_config(scene->kind, c);
    if (rc) return rc;
    TraceLaunch& L = *out;
    L = TraceLaunch{};
    L.scene = scene;
    L.rays = d_rays;
    L.order = d_order;
    L.hits = d_hits;
    L.n = n;
    L.flags = flags;
    L.cfg = c;
    L.stats = d_counters;
    L.stream = (hipStream_t)stream;
    return CWBVH_OK;
}

}  // namespace cw

extern "C" {

static int32_t trace_impl(const cwbvh_scene* scene, const cwbvh_ray* d_rays, const int32_t* d_order,
                          cwbvh_hit* d_hits, int64_t n, uint32_t flags, const cwbvh_trace_config* cfg,
                          uint64_t* d_counters, void* stream) {
    TraceLaunch L{};
    int32_t rc = prepare_trace(scene, d_rays, d_order, d_hits, n, flags, cfg, d_counters, stream, &L);
    if (rc) return rc;
    if (n == 0) return CWBVH_OK;
    const cwbvh_trace_config& c = L.cfg;
    rc = stream_counters(L.stream, &L.counters);
    if (rc) return rc;
    if (c.ray_order == 1 && d_order == nullptr && n >= kMortonMinRays) {
        // the order lives in the stream's workspace: keep it locked until
        // the traversal that reads it is queued behind the sort
        auto lock = lock_stream_scratch();
        for (int64_t b = 0; b < n; b += kMortonMaxRays) {
            TraceLaunch P = L;
            P.rays = d_rays + b;
            P.hits = d_hits + b;
            P.n = std::min(n - b, kMortonMaxRays);
            rc = morton_order_scratch(P.rays, P.n, scene->lo, scene->hi, L.stream, &P.order);
            if (rc == CWBVH_ERR_OUT_OF_MEMORY) {
                // no room for the sort's workspace: trace in the given order
                // (hip_fail has already cleared the runtime's last error)
                P.order = nullptr;
                set_error("");
            } else if (rc) {
                return rc;
            }
            rc = scene->kind == CWBVH_KIND_CWBVH ? launch_cwbvh(P) : launch_bvh2(P);
            if (rc) return rc;
        }
        return CWBVH_OK;
    }
    return scene->kind == CWBVH_KIND_CWBVH ? launch_cwbvh(L) : launch_bvh2(L);
}

int32_t cwbvh_trace(const cwbvh_scene* scene, const cwbvh_ray* d_rays, cwbvh_hit* d_hits, int64_t n, uint32_t flags,
                    const cwbvh_trace_config* cfg, uint64_t* d_counters, void* stream) {
    return trace_impl(scene, d_rays, nullptr, d_hits, n, flags, cfg, d_counters, stream);
}

int32_t cwbvh_rays_ao_from_samples_device(const cwbvh_float4* d_positions, const cwbvh_float4* d_normals,
                                          const float* d_radius, int64_t num_texels, int32_t rays_per_texel,
                                          float tmin, float tmax, uint64_t seed, cwbvh_ray* d_out, void* stream) {
    CW_CHECK_ARG(num_texels >= 0 && rays_per_texel >= 1, "bad ray count");
    CW_CHECK_ARG(num_texels * (int64_t)rays_per_texel <= 0x7FFFFFFF, "more than 2^31 - 1 rays");
    CW_CHECK_ARG(num_texels == 0 || (d_positions && d_normals && d_radius && d_out), "NULL argument");
    CW_CHECK_ARG(aligned16(d_positions) && aligned16(d_normals) && aligned16(d_out),
                 "surfel and ray arrays must be 16-byte aligned");
    return launch_surfel_rays(d_positions, d_normals, d_radius, num_texels, rays_per_texel, tmin, tmax, seed, d_out,
                              (hipStream_t)stream);
}

int32_t cwbvh_trace_surfels(const cwbvh_scene* scene, const cwbvh_float4* d_positions, const cwbvh_float4* d_normals,
                            const float* d_radius, int64_t num_texels, int32_t rays_per_texel, float tmin, float tmax,
                            uint64_t seed, uint32_t flags, cwbvh_hit* d_hits, uint32_t* d_texel_hits,
                            const cwbvh_trace_config* cfg, uint64_t* d_counters, void* stream) {
    CW_CHECK_ARG(scene != nullptr, "scene is NULL");
    CW_CHECK_ARG(scene->kind == CWBVH_KIND_CWBVH, "surfel traces run the CWBVH kernel");
    CW_CHECK_ARG(num_texels >= 0 && rays_per_texel >= 1, "bad ray count");
    const int64_t n = num_texels * (int64_t)rays_per_texel;
    CW_CHECK_ARG(n <= 0x7FFFFFFF, "more than 2^31 - 1 rays");
    CW_CHECK_ARG(n == 0 || (d_positions && d_normals && d_radius), "NULL surfel array");
    CW_CHECK_ARG(aligned16(d_positions) && aligned16(d_normals) && aligned16(d_hits),
                 "surfel and hit arrays must be 16-byte aligned");
    CW_CHECK_ARG((d_hits != nullptr) != (d_texel_hits != nullptr), "give exactly one of hits and per-surfel counts");
    CW_CHECK_ARG((flags & ~(uint32_t)(CWBVH_TRACE_ANY | CWBVH_TRACE_PRIM_ID | CWBVH_TRACE_STATS)) == 0, "unknown flags");
    CW_CHECK_ARG(!(flags & CWBVH_TRACE_PRIM_ID) || (d_hits && scene->prim_idx), "PRIM_ID needs hits and a prim index");
    CW_CHECK_ARG(!(flags & CWBVH_TRACE_STATS) || d_counters != nullptr, "STATS needs a counter buffer");
    cwbvh_trace_config c;
    if (cfg) c = *cfg;
    else {
        // the scene's default without node prefetch, which surfel traces do
        // not have (rays are made at refill; ray_order does not apply)
        cwbvh_trace_config_for_scene(scene, &c);
        c.prefetch = 0;
    }
    int32_t rc = validate_config(scene->kind, c);
    if (rc) return rc;
    if (n == 0) return CWBVH_OK;
    TraceLaunch L{};
    L.scene = scene;
    L.hits = d_hits;
    L.n = n;
    L.flags = flags;
    L.cfg = c;
    L.stats = d_counters;
    L.stream = (hipStream_t)stream;
    L.surf_pos = d_positions;
    L.surf_nrm = d_normals;
    L.surf_rad = d_radius;
    L.rays_per_texel = rays_per_texel;
    L.gen_tmin = tmin;
    L.gen_tmax = tmax;
    L.seed = seed;
    L.texel_hits = d_texel_hits;
    rc = stream_counters(L.stream, &L.counters);
    if (rc) return rc;
    return launch_cwbvh_surfels(L);
}

int32_t cwbvh_trace_ordered(const cwbvh_scene* scene, const cwbvh_ray* d_rays, const int32_t* d_order,
                            cwbvh_hit* d_hits, int64_t n, uint32_t flags, const cwbvh_trace_config* cfg,
                            uint64_t* d_counters, void* stream) {
    CW_CHECK_ARG(d_order != nullptr || n == 0, "order is NULL");
    return trace_impl(scene, d_rays, d_order, d_hits, n, flags, cfg, d_counters, stream);
}

int32_t cwbvh_loop_profile(void* stream, uint64_t* out6) {
    CW_CHECK_ARG(out6 != nullptr, "out is NULL");
    unsigned* c = nullptr;
    int32_t rc = stream_counters((hipStream_t)stream, &c);
    if (rc) return rc;
    CW_HIP(hipMemcpyAsync(out6, c + kWorkProfile, 6 * sizeof(uint64_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
    CW_HIP(hipMemsetAsync(c + kWorkProfile, 0, 6 * sizeof(uint64_t), (hipStream_t)stream));
    CW_HIP(hipStreamSynchronize((hipStream_t)stream));
    return CWBVH_OK;
}

int32_t cwbvh_stack_overflowed(void* stream, int32_t* flag) {
    CW_CHECK_ARG(flag != nullptr, "flag is NULL");
    unsigned* c = nullptr;
    int32_t rc = stream_counters((hipStream_t)stream, &c);
    if (rc) return rc;
    unsigned v = 0;
    CW_HIP(hipMemcpyAsync(&v, c + kWorkOverflow, sizeof(unsigned), hipMemcpyDeviceToHost, (hipStream_t)stream));
    CW_HIP(hipStreamSynchronize((hipStream_t)stream));
    *flag = v ? 1 : 0;
    if (v) CW_HIP(hipMemsetAsync(c + kWorkOverflow, 0, sizeof(unsigned), (hipStream_t)stream));
    return CWBVH_OK;
}

}  // extern "C"
