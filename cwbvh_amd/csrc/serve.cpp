// Serving loop: resident scenes, HBM-budget admission, continuous batching.
//
// No counterpart in the reference (a single-shot demo, src/main.cpp).  One
// scheduler thread owns the device side:
//   * a device arena of `memory_budget` bytes holds the rays and hits of every
//     admitted request (48 B per ray); requests wait in FIFO order until the
//     arena has room, so the budget bounds the HBM held by in-flight work;
//   * up to `max_inflight` batch slots, each with its own stream, job table and
//     fetch counters: whenever a slot is free, every pending request that fits
//     the arena (up to the batch limits) is copied in and traversed by ONE
//     persistent multi-scene launch, so requests that arrive while earlier
//     batches are on the GPU join the next batch and overlap with them;
//   * completion is detected by polling each slot's event; hits are copied
//     back on the slot's stream before the event, arena space is released and
//     waiters are woken.
// Every ray's result is independent of which batch, slot or wave traversed it
// (deterministic tie rule), so a request returns the same hits alone and in
// the mix.
//
// Three request kinds:
//   * host rays: rays are copied into the arena and hits copied back (48 B
//     per ray over PCIe);
//   * device rays: the caller's device ray and hit buffers are used in place,
//     nothing crosses PCIe and nothing is taken from the arena;
//   * surfels: positions, normals and radii (36 B per surfel) are copied in,
//     the fused generate-and-trace kernel makes rays_per_texel AO rays per
//     surfel, and a 4 B count of the rays that hit comes back per surfel.
//     A surfel request is one launch of its own on its slot (the fused
//     kernel traverses one scene); ray requests share multi-scene launches.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>

#include "internal.h"

using namespace cw;

namespace {

using Clock = std::chrono::steady_clock;

double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

constexpr int64_t kAlign = 256;

int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) with coalescing.
class Arena {
public:
    explicit Arena(int64_t size = 0) { reset(size); }
    void reset(int64_t size) {
        free_.clear();
        if (size > 0) free_[0] = size;
        used_ = 0;
    }
    int64_t alloc(int64_t bytes) {
        bytes = align_up(std::max<int64_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                int64_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest > 0) free_[off + bytes] = rest;
                used_ += bytes;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = align_up(std::max<int64_t>(bytes, 1));
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t used() const { return used_; }

private:
    std::map<int64_t, int64_t> free_;
    int64_t used_ = 0;
};

struct Request {
    uint64_t id = 0;
    int32_t kind = CWBVH_REQUEST_HOST_RAYS;
    int32_t scene = 0;
    const cwbvh_ray* rays = nullptr;   // host (HOST_RAYS) or device (DEVICE_RAYS)
    cwbvh_hit* hits = nullptr;
    int64_t n = 0;           // rays (surfels: texels x rays per texel)
    cwbvh_surfel_request surf{};
    int64_t off = -1;        // arena offset (rays, then hits; or the surfel arrays, then counts)
    double t_submit = 0, t_done = 0;
    bool done = false;
    int32_t status = CWBVH_OK;

    // arena bytes the request holds while in flight
    int64_t arena_bytes() const {
        switch (kind) {
            case CWBVH_REQUEST_DEVICE_RAYS: return 0;
            case CWBVH_REQUEST_SURFELS: {
                const int64_t t = surf.num_texels;
                return 2 * align_up(t * 16) + align_up(t * 4) + t * 4;
            }
            default: return align_up(n * (int64_t)sizeof(cwbvh_ray)) + n * (int64_t)sizeof(cwbvh_hit);
        }
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    BatchJob* d_jobs = nullptr;
    BatchJob* h_jobs = nullptr;   // pinned
    unsigned* counters = nullptr;
    std::vector<Request*> reqs;
    bool busy = false;
};

}  // namespace

struct cwbvh_server {
    // reserved to CWBVH_SERVER_MAX_SCENES at creation, so cwbvh_server_add_scene
    // appends (under mu) without moving the entries the worker reads
    std::vector<const cwbvh_scene*> scenes;
    cwbvh_server_config cfg{};
    char* d_arena = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> requests;
    uint64_t next_id = 1;
    int64_t outstanding = 0;
    bool stop = false;
    int32_t worker_status = CWBVH_OK;
    std::string worker_error;
    std::thread worker;
    // stats (guarded by mu)
    cwbvh_server_stats stats{};
    int inflight = 0;
    double busy_since = 0;

    int32_t num_scenes() {
        std::lock_guard<std::mutex> g(mu);
        return (int32_t)scenes.size();
    }
    int32_t launch(Slot& sl, const std::vector<Request*>& batch);
    void complete(Slot& sl);
    void fail_pending();
    void loop();
};

// lds_depth 0 in the server's trace configuration: each launch keeps the
// deepest stack need of its trees in LDS, so the LDS-only kernels serve it
// (10 entries when a tree's need is unknown)
static int32_t launch_lds_depth(int32_t configured, int32_t max_need) {
    if (configured > 0) return configured;
    return max_need > 0 && max_need <= CWBVH_STACK_CAPACITY ? max_need : 10;
}

int32_t cwbvh_server::launch(Slot& sl, const std::vector<Request*>& batch) {
    if (batch[0]->kind == CWBVH_REQUEST_SURFELS) {
        // one fused generate-and-trace launch; per-surfel counts come back
        const Request* r = batch[0];
        const cwbvh_surfel_request& q = r->surf;
        const int64_t t = q.num_texels;
        char* base = d_arena + r->off;
        char* pos = base;
        char* nrm = pos + align_up(t * 16);
        char* rad = nrm + align_up(t * 16);
        char* cnt = rad + align_up(t * 4);
        CW_HIP(hipMemcpyAsync(pos, q.positions, (size_t)t * 16, hipMemcpyHostToDevice, sl.stream));
        CW_HIP(hipMemcpyAsync(nrm, q.normals, (size_t)t * 16, hipMemcpyHostToDevice, sl.stream));
        CW_HIP(hipMemcpyAsync(rad, q.radius, (size_t)t * 4, hipMemcpyHostToDevice, sl.stream));
        CW_HIP(hipMemsetAsync(cnt, 0, (size_t)t * 4, sl.stream));
        TraceLaunch L{};
        L.scene = scenes[r->scene];
        L.n = r->n;
        L.flags = cfg.flags & CWBVH_TRACE_ANY;
        L.cfg = cfg.trace;
        L.cfg.lds_depth = launch_lds_depth(cfg.trace.lds_depth, L.scene->stack_need);
        L.cfg.prefetch = 0;   // the fused kernel has no node prefetch
        L.counters = sl.counters;
        L.stream = sl.stream;
        L.surf_pos = reinterpret_cast<const cwbvh_float4*>(pos);
        L.surf_nrm = reinterpret_cast<const cwbvh_float4*>(nrm);
        L.surf_rad = reinterpret_cast<const float*>(rad);
        L.rays_per_texel = q.rays_per_texel;
        L.gen_tmin = q.tmin;
        L.gen_tmax = q.tmax;
        L.seed = q.seed;
        L.texel_hits = reinterpret_cast<uint32_t*>(cnt);
        int32_t rc = launch_cwbvh_surfels(L);
        if (rc) return rc;
        CW_HIP(hipMemcpyAsync(q.texel_hits, cnt, (size_t)t * 4, hipMemcpyDeviceToHost, sl.stream));
        CW_HIP(hipEventRecord(sl.done, sl.stream));
        sl.reqs = batch;
        sl.busy = true;
        return CWBVH_OK;
    }
    const int32_t chunk = cfg.trace.fetch_chunk;
    uint32_t chunks = 0;
    int32_t max_need = 0;  // deepest stack need of the launch's trees (-1: unknown)
    for (size_t k = 0; k < batch.size(); ++k) {
        const Request* r = batch[k];
        const cwbvh_scene* sc = scenes[r->scene];
        BatchJob& J = sl.h_jobs[k];
        if (r->kind == CWBVH_REQUEST_DEVICE_RAYS) {
            J.rays = r->rays;
            J.hits = r->hits;
        } else {
            char* base = d_arena + r->off;
            CW_HIP(hipMemcpyAsync(base, r->rays, (size_t)r->n * sizeof(cwbvh_ray), hipMemcpyHostToDevice, sl.stream));
            J.rays = base;
            J.hits = base + align_up(r->n * (int64_t)sizeof(cwbvh_ray));
        }
        J.nodes = sc->nodes;
        J.woop = sc->woop;
        J.order = nullptr;
        J.prim_idx = sc->prim_idx;
        J.n = (int32_t)r->n;
        J.node_stride = sc->node_stride;
        J.chunk_begin = chunks;
        J.spec_limit = speculation_limit(sc->stack_need);
        chunks += (uint32_t)((r->n + chunk - 1) / chunk);
        if (sc->stack_need <= 0) max_need = -1;
        else if (max_need >= 0) max_need = std::max(max_need, sc->stack_need);
    }
    if (cfg.serial) {
        const Request* r = batch[0];
        TraceLaunch L{};
        L.scene = scenes[r->scene];
        L.rays = sl.h_jobs[0].rays;
        L.order = nullptr;
        L.hits = sl.h_jobs[0].hits;
        L.n = r->n;
        L.flags = cfg.flags;
        L.cfg = cfg.trace;
        L.cfg.lds_depth = launch_lds_depth(cfg.trace.lds_depth, max_need);
        L.counters = sl.counters;
        L.stream = sl.stream;
        int32_t rc = launch_cwbvh(L);
        if (rc) return rc;
    } else {
        CW_HIP(hipMemcpyAsync(sl.d_jobs, sl.h_jobs, batch.size() * sizeof(BatchJob), hipMemcpyHostToDevice, sl.stream));
        BatchLaunch L{};
        L.d_jobs = sl.d_jobs;
        L.num_jobs = (int32_t)batch.size();
        L.total_chunks = chunks;
        L.flags = cfg.flags;
        L.cfg = cfg.trace;
        L.cfg.lds_depth = launch_lds_depth(cfg.trace.lds_depth, max_need);
        L.counters = sl.counters;
        L.stream = sl.stream;
        L.lds_only = max_need > 0 && max_need <= L.cfg.lds_depth && L.cfg.lds_depth <= CWBVH_STACK_CAPACITY;
        int32_t rc = launch_cwbvh_batch(L);
        if (rc) return rc;
    }
    for (size_t k = 0; k < batch.size(); ++k) {
        const Request* r = batch[k];
        if (r->kind != CWBVH_REQUEST_HOST_RAYS) continue;
        CW_HIP(hipMemcpyAsync(r->hits, sl.h_jobs[k].hits, (size_t)r->n * sizeof(cwbvh_hit), hipMemcpyDeviceToHost,
                              sl.stream));
    }
    CW_HIP(hipEventRecord(sl.done, sl.stream));
    sl.reqs = batch;
    sl.busy = true;
    return CWBVH_OK;
}

void cwbvh_server::complete(Slot& sl) {
    const double t = now_s();
    std::lock_guard<std::mutex> g(mu);
    for (Request* r : sl.reqs) {
        r->t_done = t;
        r->done = true;
        if (r->off >= 0) arena.release(r->off, r->arena_bytes());
        stats.completed++;
        stats.rays += r->n;
        --outstanding;
    }
    sl.reqs.clear();
    sl.busy = false;
    if (--inflight == 0) stats.busy_s += t - busy_since;
    cv_done.notify_all();
}

// After a device failure no further batch is launched: every request still
// waiting for the arena ends with the failure status, so waiters, drain and
// destroy return.  Called with `mu` held.
void cwbvh_server::fail_pending() {
    const double t = now_s();
    for (Request* r : pending) {
        r->status = worker_status;
        r->done = true;
        r->t_done = t;
        --outstanding;
    }
    pending.clear();
}

void cwbvh_server::loop() {
    std::vector<Request*> batch;
    while (true) {
        bool progress = false;
        for (Slot& sl : slots) {
            if (!sl.busy) continue;
            hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) {
                complete(sl);
                progress = true;
            } else if (q != hipErrorNotReady) {
                const double t = now_s();
                std::lock_guard<std::mutex> g(mu);
                worker_status = hip_fail(q, "serving batch");
                worker_error = get_error();
                for (Request* r : sl.reqs) {
                    r->status = worker_status;
                    r->done = true;
                    r->t_done = t;
                    --outstanding;
                }
                sl.reqs.clear();
                sl.busy = false;
                if (--inflight == 0) stats.busy_s += t - busy_since;
                fail_pending();
                cv_done.notify_all();
            }
        }
        Slot* free_slot = nullptr;
        for (Slot& sl : slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            if (stop && pending.empty() && inflight == 0) break;
            if (free_slot && worker_status == CWBVH_OK) {
                const int32_t max_req = cfg.serial ? 1 : cfg.max_batch_requests;
                int64_t rays = 0;
                while (!pending.empty() && (int32_t)batch.size() < max_req) {
                    Request* r = pending.front();
                    const bool surfels = r->kind == CWBVH_REQUEST_SURFELS;
                    // a surfel request is a launch of its own
                    if (!batch.empty() && (surfels || rays + r->n > cfg.max_batch_rays)) break;
                    const int64_t bytes = r->arena_bytes();
                    if (bytes > 0) {
                        const int64_t off = arena.alloc(bytes);
                        if (off < 0) break;  // FIFO: wait for completions to free HBM
                        r->off = off;
                    }
                    batch.push_back(r);
                    pending.pop_front();
                    rays += r->n;
                    if (surfels) break;
                }
                if (!batch.empty()) {
                    stats.peak_bytes = std::max(stats.peak_bytes, arena.used());
                    if (inflight++ == 0) busy_since = now_s();
                    stats.max_inflight_seen = std::max(stats.max_inflight_seen, inflight);
                    stats.max_batch_requests_seen = std::max(stats.max_batch_requests_seen, (int32_t)batch.size());
                    stats.batches++;
                }
            }
            if (batch.empty() && !progress) {
                // nothing to launch: sleep until a submission or a short poll interval
                cv_work.wait_for(lk, std::chrono::microseconds(inflight ? 20 : 2000));
                continue;
            }
        }
        if (!batch.empty()) {
            int32_t rc = launch(*free_slot, batch);
            if (rc) {
                const std::string why = get_error();
                // copies queued before the failing call may still read the
                // callers' rays or write their hits: let them finish before
                // the requests are reported
                (void)hipStreamSynchronize(free_slot->stream);
                const double t = now_s();
                std::lock_guard<std::mutex> g(mu);
                worker_status = rc;
                worker_error = why;
                for (Request* r : batch) {
                    r->status = rc;
                    r->done = true;
                    r->t_done = t;
                    if (r->off >= 0) arena.release(r->off, r->arena_bytes());
                    --outstanding;
                }
                if (--inflight == 0) stats.busy_s += t - busy_since;
                fail_pending();
                cv_done.notify_all();
            }
        }
    }
}

extern "C" {

int32_t cwbvh_host_alloc(void** ptr, int64_t bytes) {
    CW_CHECK_ARG(ptr != nullptr && bytes >= 0, "bad arguments");
    *ptr = nullptr;
    if (bytes == 0) return CWBVH_OK;
    CW_HIP(hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault));
    return CWBVH_OK;
}

int32_t cwbvh_host_free(void* ptr) {
    if (ptr) CW_HIP(hipHostFree(ptr));
    return CWBVH_OK;
}

int32_t cwbvh_server_config_default(cwbvh_server_config* c) {
    CW_CHECK_ARG(c != nullptr, "cfg is NULL");
    std::memset(c, 0, sizeof(*c));
    c->memory_budget = (int64_t)8 << 30;
    c->max_inflight = 3;
    c->max_batch_requests = 64;
    c->max_batch_rays = (int64_t)16 << 20;
    c->serial = 0;
    c->flags = CWBVH_TRACE_CLOSEST;
    int32_t rc = cwbvh_trace_config_default(CWBVH_KIND_CWBVH, &c->trace);
    c->trace.lds_depth = 0;  // per launch: the deepest stack need of its trees
    return rc;
}

int32_t cwbvh_server_create(const cwbvh_scene* const* scenes, int32_t num_scenes, const cwbvh_server_config* cfg,
                            cwbvh_server** out) {
    CW_CHECK_ARG(out != nullptr && scenes != nullptr && num_scenes >= 1, "bad arguments");
    cwbvh_server_config c;
    if (cfg) c = *cfg;
    else cwbvh_server_config_default(&c);
    CW_CHECK_ARG(c.memory_budget >= 4096 && c.max_inflight >= 1 && c.max_inflight <= 16 &&
                     c.max_batch_requests >= 1 && c.max_batch_requests <= 4096 && c.max_batch_rays >= 1,
                 "invalid server configuration");
    CW_CHECK_ARG((c.flags & ~(uint32_t)(CWBVH_TRACE_ANY | CWBVH_TRACE_PRIM_ID)) == 0, "unsupported flags");
    CW_CHECK_ARG(c.trace.lds_depth >= 0, "lds_depth must be >= 0 (0 = each launch's deepest stack need)");
    cwbvh_trace_config vc = c.trace;
    if (vc.lds_depth == 0) vc.lds_depth = CWBVH_STACK_CAPACITY;  // the deepest a launch can pick
    int32_t rc = validate_config(CWBVH_KIND_CWBVH, vc);
    if (rc) return rc;
    for (int32_t i = 0; i < num_scenes; ++i)
        CW_CHECK_ARG(scenes[i] && scenes[i]->kind == CWBVH_KIND_CWBVH &&
                         (!(c.flags & CWBVH_TRACE_PRIM_ID) || scenes[i]->prim_idx),
                     "server scenes must be CWBVH scenes");
    if (c.trace.fetch_chunk == 0) c.trace.fetch_chunk = kBatchChunk;  // batches are planned in chunks
    CW_CHECK_ARG(num_scenes <= CWBVH_SERVER_MAX_SCENES, "too many scenes");
    auto* s = new cwbvh_server();
    s->scenes.reserve(CWBVH_SERVER_MAX_SCENES);
    s->scenes.assign(scenes, scenes + num_scenes);
    s->cfg = c;
    if (c.serial) s->cfg.max_inflight = 1;
    hipError_t e = hipMalloc(&s->d_arena, (size_t)c.memory_budget);
    s->arena.reset(c.memory_budget);
    s->slots.resize(s->cfg.max_inflight);
    const int32_t jobs = s->cfg.serial ? 1 : c.max_batch_requests;
    for (Slot& sl : s->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&sl.d_jobs, jobs * sizeof(BatchJob));
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_jobs, jobs * sizeof(BatchJob), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&sl.counters, kWorkWords * sizeof(unsigned));
        if (e == hipSuccess) e = hipMemset(sl.counters, 0, kWorkWords * sizeof(unsigned));
    }
    if (e != hipSuccess) {
        cwbvh_server_destroy(s);
        return hip_fail(e, "cwbvh_server_create");
    }
    s->worker = std::thread([s] { s->loop(); });
    *out = s;
    return CWBVH_OK;
}

}  // extern "C"

namespace {

int32_t enqueue(cwbvh_server* s, Request* r, uint64_t* request_id) {
    if (r->arena_bytes() > s->cfg.memory_budget) {
        set_error("request of " + std::to_string(r->n) + " rays exceeds the serving memory budget");
        delete r;
        return CWBVH_ERR_BUDGET;
    }
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->worker_status != CWBVH_OK) {
            delete r;
            set_error("server failed: " + s->worker_error);
            return s->worker_status;
        }
        r->id = s->next_id++;
        s->requests[r->id] = r;
        if (r->n == 0) {
            r->done = true;
            r->t_done = r->t_submit;
        } else {
            s->pending.push_back(r);
            ++s->outstanding;
        }
        *request_id = r->id;
    }
    s->cv_work.notify_one();
    return CWBVH_OK;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

extern "C" {

int32_t cwbvh_server_submit(cwbvh_server* s, int32_t scene, const cwbvh_ray* host_rays, int64_t n,
                            cwbvh_hit* host_hits, uint64_t* request_id) {
    CW_CHECK_ARG(s != nullptr && request_id != nullptr, "NULL argument");
    CW_CHECK_ARG(scene >= 0 && scene < s->num_scenes(), "scene index out of range");
    CW_CHECK_ARG(n >= 0 && n <= 0x7FFFFFFF && (n == 0 || (host_rays && host_hits)), "bad ray buffer");
    auto* r = new Request();
    r->kind = CWBVH_REQUEST_HOST_RAYS;
    r->scene = scene;
    r->rays = host_rays;
    r->hits = host_hits;
    r->n = n;
    return enqueue(s, r, request_id);
}

int32_t cwbvh_server_submit_device(cwbvh_server* s, int32_t scene, const cwbvh_ray* d_rays, int64_t n,
                                   cwbvh_hit* d_hits, uint64_t* request_id) {
    CW_CHECK_ARG(s != nullptr && request_id != nullptr, "NULL argument");
    CW_CHECK_ARG(scene >= 0 && scene < s->num_scenes(), "scene index out of range");
    CW_CHECK_ARG(n >= 0 && n <= 0x7FFFFFFF && (n == 0 || (d_rays && d_hits)), "bad ray buffer");
    CW_CHECK_ARG(aligned16(d_rays) && aligned16(d_hits), "ray and hit buffers must be 16-byte aligned");
    auto* r = new Request();
    r->kind = CWBVH_REQUEST_DEVICE_RAYS;
    r->scene = scene;
    r->rays = d_rays;
    r->hits = d_hits;
    r->n = n;
    return enqueue(s, r, request_id);
}

int32_t cwbvh_server_submit_surfels(cwbvh_server* s, int32_t scene, const cwbvh_surfel_request* q,
                                    uint64_t* request_id) {
    CW_CHECK_ARG(s != nullptr && request_id != nullptr && q != nullptr, "NULL argument");
    CW_CHECK_ARG(scene >= 0 && scene < s->num_scenes(), "scene index out of range");
    CW_CHECK_ARG(q->num_texels >= 0 && q->rays_per_texel >= 1 &&
                     q->num_texels * (int64_t)q->rays_per_texel <= 0x7FFFFFFF,
                 "bad surfel count");
    CW_CHECK_ARG(q->num_texels == 0 || (q->positions && q->normals && q->radius && q->texel_hits),
                 "NULL surfel array");
    auto* r = new Request();
    r->kind = CWBVH_REQUEST_SURFELS;
    r->scene = scene;
    r->surf = *q;
    r->n = q->num_texels * (int64_t)q->rays_per_texel;
    return enqueue(s, r, request_id);
}

int32_t cwbvh_server_add_scene(cwbvh_server* s, const cwbvh_scene* scene, int32_t* index) {
    CW_CHECK_ARG(s != nullptr && scene != nullptr && index != nullptr, "NULL argument");
    CW_CHECK_ARG(scene->kind == CWBVH_KIND_CWBVH && (!(s->cfg.flags & CWBVH_TRACE_PRIM_ID) || scene->prim_idx),
                 "server scenes must be CWBVH scenes");
    std::lock_guard<std::mutex> g(s->mu);
    if ((int32_t)s->scenes.size() >= CWBVH_SERVER_MAX_SCENES) {
        set_error("server holds CWBVH_SERVER_MAX_SCENES scenes already");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    *index = (int32_t)s->scenes.size();
    s->scenes.push_back(scene);
    return CWBVH_OK;
}

int32_t cwbvh_server_wait(cwbvh_server* s, uint64_t id, double timeout_s, double* latency_ms) {
    CW_CHECK_ARG(s != nullptr, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    CW_CHECK_ARG(it != s->requests.end(), "unknown request id");
    Request* r = it->second;
    auto pred = [r] { return r->done; };
    if (timeout_s < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
        set_error("timed out");
        return CWBVH_ERR_CANCELLED;
    }
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
    int32_t st = r->status;
    s->requests.erase(it);
    delete r;
    if (st) set_error(s->worker_error);
    return st;
}

int32_t cwbvh_server_drain(cwbvh_server* s) {
    CW_CHECK_ARG(s != nullptr, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [s] { return s->outstanding == 0; });
    return s->worker_status;
}

int32_t cwbvh_server_get_stats(const cwbvh_server* s, cwbvh_server_stats* out) {
    CW_CHECK_ARG(s != nullptr && out != nullptr, "NULL argument");
    auto* m = const_cast<cwbvh_server*>(s);
    std::lock_guard<std::mutex> g(m->mu);
    *out = m->stats;
    return CWBVH_OK;
}

void cwbvh_server_destroy(cwbvh_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto& kv : s->requests) delete kv.second;
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.d_jobs) (void)hipFree(sl.d_jobs);
        if (sl.h_jobs) (void)hipHostFree(sl.h_jobs);
        if (sl.counters) (void)hipFree(sl.counters);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->d_arena) (void)hipFree(s->d_arena);
    delete s;
}

}  // extern "C"

namespace {

// Poisson arrivals: request k arrives after an exponential gap, picks a scene
// uniformly and a ray count log-uniform in [min_rays, max_rays] (at most the
// scene's pool), and is handed to submit(k, scene, rays, pool offset, ring
// offset) -> status, where ring offsets wrap at ring_size.  Then every
// request is waited for and the latency percentiles are computed.
template <class Submit>
int32_t drive_poisson(cwbvh_server* s, const cwbvh_poisson_spec* spec, const int64_t* pool_sizes, int64_t ring_size,
                      int64_t ray_unit, double* latencies_ms, cwbvh_poisson_result* out, Submit submit) {
    const int32_t ns = s->num_scenes();
    std::mt19937_64 rng(spec->seed);
    std::exponential_distribution<double> gap(spec->rate);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    std::vector<uint64_t> ids(spec->num_requests);
    std::vector<int64_t> sizes(spec->num_requests);
    int64_t ring = 0;
    const double lmin = std::log((double)spec->min_rays), lmax = std::log((double)spec->max_rays);
    const double t0 = now_s();
    double t_next = t0;
    int32_t rc = CWBVH_OK;
    for (int32_t k = 0; k < spec->num_requests && rc == CWBVH_OK; ++k) {
        t_next += gap(rng);
        const int32_t sc = (int32_t)(u01(rng) * ns) % ns;
        // request size in pool units (rays, or surfels of ray_unit rays)
        int64_t n = (int64_t)std::exp(lmin + (lmax - lmin) * u01(rng)) / ray_unit;
        n = std::min(std::max<int64_t>(n, std::max<int64_t>(1, spec->min_rays / ray_unit)),
                     std::min<int64_t>(spec->max_rays / ray_unit, pool_sizes[sc]));
        n = std::max<int64_t>(n, 1);
        const int64_t off = pool_sizes[sc] > n ? (int64_t)(u01(rng) * (double)(pool_sizes[sc] - n)) : 0;
        if (ring + n > ring_size) ring = 0;
        while (now_s() < t_next) {
            // arrival clock: sleep when far ahead, spin for the last tens of microseconds
            const double ahead = t_next - now_s();
            if (ahead > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(ahead - 1e-4));
        }
        rc = submit(sc, n, off, ring, &ids[k]);
        ring += n;
        sizes[k] = n * ray_unit;
    }
    if (rc) return rc;
    std::vector<double> lat(spec->num_requests);
    for (int32_t k = 0; k < spec->num_requests; ++k) {
        rc = cwbvh_server_wait(s, ids[k], -1.0, &lat[k]);
        if (rc) return rc;
    }
    const double t_last = now_s();
    int64_t rays = 0;
    for (int64_t v : sizes) rays += v;
    if (latencies_ms) std::copy(lat.begin(), lat.end(), latencies_ms);
    std::vector<double> sorted = lat;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double p) {
        const size_t i = (size_t)std::min<double>((double)sorted.size() - 1, std::ceil(p * sorted.size()) - 1);
        return sorted[i];
    };
    double mean = 0;
    for (double v : lat) mean += v;
    out->wall_s = t_last - t0;
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->mean_ms = mean / lat.size();
    out->max_ms = sorted.back();
    out->rays = rays;
    out->mrays_per_s = rays / out->wall_s / 1e6;
    return CWBVH_OK;
}

int32_t check_spec(const cwbvh_poisson_spec* spec) {
    CW_CHECK_ARG(spec->rate > 0 && spec->num_requests >= 1 && spec->min_rays >= 1 && spec->max_rays >= spec->min_rays,
                 "bad Poisson spec");
    return CWBVH_OK;
}

}  // namespace

extern "C" {

int32_t cwbvh_server_run_poisson(cwbvh_server* s, const cwbvh_poisson_spec* spec, const cwbvh_ray* const* ray_pools,
                                 const int64_t* pool_sizes, cwbvh_hit* hit_pool, int64_t hit_pool_size,
                                 double* latencies_ms, cwbvh_poisson_result* out) {
    CW_CHECK_ARG(s && spec && ray_pools && pool_sizes && hit_pool && out, "NULL argument");
    int32_t rc = check_spec(spec);
    if (rc) return rc;
    for (int32_t k = 0; k < s->num_scenes(); ++k) CW_CHECK_ARG(ray_pools[k] && pool_sizes[k] >= 1, "empty ray pool");
    CW_CHECK_ARG(hit_pool_size >= spec->max_rays, "hit pool smaller than the largest request");
    return drive_poisson(s, spec, pool_sizes, hit_pool_size, 1, latencies_ms, out,
                         [&](int32_t sc, int64_t n, int64_t off, int64_t ring, uint64_t* id) {
                             return cwbvh_server_submit(s, sc, ray_pools[sc] + off, n, hit_pool + ring, id);
                         });
}

int32_t cwbvh_server_run_poisson_device(cwbvh_server* s, const cwbvh_poisson_spec* spec,
                                        const cwbvh_ray* const* d_ray_pools, const int64_t* pool_sizes,
                                        cwbvh_hit* d_hit_pool, int64_t hit_pool_size, double* latencies_ms,
                                        cwbvh_poisson_result* out) {
    CW_CHECK_ARG(s && spec && d_ray_pools && pool_sizes && d_hit_pool && out, "NULL argument");
    int32_t rc = check_spec(spec);
    if (rc) return rc;
    for (int32_t k = 0; k < s->num_scenes(); ++k) CW_CHECK_ARG(d_ray_pools[k] && pool_sizes[k] >= 1, "empty ray pool");
    CW_CHECK_ARG(hit_pool_size >= spec->max_rays, "hit pool smaller than the largest request");
    return drive_poisson(s, spec, pool_sizes, hit_pool_size, 1, latencies_ms, out,
                         [&](int32_t sc, int64_t n, int64_t off, int64_t ring, uint64_t* id) {
                             return cwbvh_server_submit_device(s, sc, d_ray_pools[sc] + off, n, d_hit_pool + ring, id);
                         });
}

int32_t cwbvh_server_run_poisson_surfels(cwbvh_server* s, const cwbvh_poisson_spec* spec,
                                         const cwbvh_surfel_pool* pools, int32_t rays_per_texel, float tmin,
                                         float tmax, uint32_t* count_pool, int64_t count_pool_size,
                                         double* latencies_ms, cwbvh_poisson_result* out) {
    CW_CHECK_ARG(s && spec && pools && count_pool && out, "NULL argument");
    int32_t rc = check_spec(spec);
    if (rc) return rc;
    CW_CHECK_ARG(rays_per_texel >= 1 && spec->max_rays >= rays_per_texel, "bad rays per texel");
    const size_t ns = (size_t)s->num_scenes();
    std::vector<int64_t> sizes(ns);
    for (size_t k = 0; k < ns; ++k) {
        CW_CHECK_ARG(pools[k].positions && pools[k].normals && pools[k].radius && pools[k].num_texels >= 1,
                     "empty surfel pool");
        sizes[k] = pools[k].num_texels;
    }
    CW_CHECK_ARG(count_pool_size >= spec->max_rays / rays_per_texel, "count pool smaller than the largest request");
    return drive_poisson(s, spec, sizes.data(), count_pool_size, rays_per_texel, latencies_ms, out,
                         [&](int32_t sc, int64_t t, int64_t off, int64_t ring, uint64_t* id) {
                             cwbvh_surfel_request q{};
                             q.positions = pools[sc].positions + off;
                             q.normals = pools[sc].normals + off;
                             q.radius = pools[sc].radius + off;
                             q.num_texels = t;
                             q.rays_per_texel = rays_per_texel;
                             q.tmin = tmin;
                             q.tmax = tmax;
                             q.seed = spec->seed + (uint64_t)off;
                             q.texel_hits = count_pool + ring;
                             return cwbvh_server_submit_surfels(s, sc, &q, id);
                         });
}

}  // extern "C"
