// Traversal kernels for gfx950 (wave64) and their launchers.
//
//  * cwbvh_trace: persistent CWBVH traversal with wave-level work fetching.
//    Follows the reference kernel src/TraversalKernelCWBVH.cu:24-375:
//      - ray fetch: terminated lanes of a wave are refilled together; one
//        lane claims `fetch_chunk` rays with a single atomic and the wave
//        hands them out by prefix rank (ballot + popcount), instead of one
//        atomic per refill as in the reference (:60-72);
//      - node group / triangle group state machine, highest-bit-first child
//        order with octant XOR, 8-child quantized slab test (:93-295);
//      - optional triangle postponing (:297-311) and the dynamic-fetch exit
//        when idle lanes accumulate (:365-371), both tunable for 64 lanes;
//      - traversal stack: first `lds_depth` entries in LDS (interleaved by
//        lane), the rest in per-lane scratch, capacity 64 entries with an
//        overflow flag (the reference has 8 shared + 32 local entries and
//        no check).
//    One template serves a single scene (uniform scene pointers) and the
//    multi-scene batch (per-lane scene pointers from a job table).
//  * bvh2_trace: one thread per ray, Aila-Laine while-while traversal with one
//    postponed leaf and a wave-wide "any lane wants nodes" exit, Woop leaves
//    closed by a 0x80000000 terminator (src/TraversalKernelBVH2.cu:26-178).
//
// Both kernels use the arithmetic contract of device_common.h, so results are
// bit-identical between kernel variants and to the scalar CPU restatement
// used by the tests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "aogen.h"
#include "device_common.h"
#include "internal.h"

using namespace cwd;

static_assert(kStackCapacity == CWBVH_STACK_CAPACITY, "device and public stack capacities agree");

namespace {

constexpr int kMaxBlock = 256;
constexpr int kWorkSlices = cw::kWorkSlices;
constexpr int kWorkStride = cw::kWorkStride;
constexpr int kWorkExited = cw::kWorkExited;
constexpr int kWorkOverflow = cw::kWorkOverflow;
constexpr int kWorkProfile = cw::kWorkProfile;
constexpr int kWorkDrained = cw::kWorkDrained;
constexpr int kWorkGroupExit = cw::kWorkGroupExit;

// Work distribution of the persistent single-scene grids (cwbvh_trace,
// bvh2_trace_persistent).  The rays are split into `slices` contiguous ranges
// with one claim counter each.  Blocks are dispatched round-robin over the
// XCDs, so block b takes slice b % slices (its home) first, then the others.
// (Batches claim chunk indices with a plain atomic per claim and count exits
// on one word: with the three changes below their kernel ran 6% slower on
// 32 scenes of 1 M rays each.)
//  * First claim without an atomic: wave i of the W_s waves whose home is
//    slice s owns [i * amount, (i + 1) * amount) of it outright, and the
//    counter hands out the slice from W_s * amount on.  Otherwise every wave
//    of the grid claims at once when the launch starts, and the claims queue
//    on eight counter lines.
//  * Drained mask: the first wave to find a slice dry sets its bit in one
//    word.  A wave that has run out of its home slice reads the word once and
//    skips the drained slices, instead of probing every counter with an
//    atomic at the end of the launch.
//  * Exit: blocks count their exits per home slice; the last block of a home
//    group counts on one word, and the last group resets everything, so no
//    line takes an atomic from every block.
//  * Splits: the host passes n = item_q * slices + item_r and the grid's
//    blocks = block_q * slices + block_r, so the claims divide nothing (a
//    64-bit division by a run-time value costs about 40 VALU instructions).
// Blocks whose home is slice s (blocks b with b mod slices = s).
__device__ __forceinline__ unsigned home_blocks(int s, unsigned block_q, unsigned block_r) {
    return block_q + ((unsigned)s < block_r ? 1u : 0u);
}

// First item of slice s: the first item_r slices hold item_q + 1 items, the
// others item_q.
__device__ __forceinline__ unsigned slice_begin(int s, unsigned item_q, unsigned item_r) {
    return (unsigned)s * item_q + min((unsigned)s, item_r);
}

// Slice-relative base of the wave's next claim of `amount` items from slice
// `sl` (>= the slice size when it has run dry).  `first`: the wave's first
// claim, from its home slice, with index `first_index` among that slice's
// waves.  The wave's first active lane does the memory work; the result is
// wave-uniform (scalar registers).
__device__ __forceinline__ unsigned slice_claim(unsigned* work, int sl, unsigned amount, bool first,
                                                unsigned first_index, unsigned block_q, unsigned block_r, int lane) {
    // 64-bit sums, saturated: a value past 2^32 - 1 reads as a dry slice
    if (first) return (unsigned)min((unsigned long long)first_index * amount, 0xFFFFFFFFull);
    unsigned base = 0;
    if (lane == __builtin_amdgcn_readfirstlane(lane)) base = atomicAdd(&work[sl * kWorkStride], amount);
    base = __builtin_amdgcn_readfirstlane(base);
    const unsigned long long static_items =
        (unsigned long long)home_blocks(sl, block_q, block_r) * (blockDim.x >> 6) * amount;
    return (unsigned)min(static_items + base, 0xFFFFFFFFull);
}

// Marks slice `sl` drained and returns the drained mask as now known: one
// load, and an atomic only when the slice is not marked yet.
__device__ __forceinline__ unsigned mark_drained(unsigned* work, int sl, int lane) {
    unsigned m = 0;
    if (lane == __builtin_amdgcn_readfirstlane(lane)) {
        m = __hip_atomic_load(&work[kWorkDrained], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!((m >> sl) & 1u)) m = atomicOr(&work[kWorkDrained], 1u << sl);
    }
    return __builtin_amdgcn_readfirstlane(m) | (1u << sl);
}

// Exit of a persistent block (all threads call it).  Every wave has consumed
// the result of its last claim before the barrier, so the claims are
// complete when the exit is counted; the last block out resets the counters
// for the next launch on the stream.
__device__ __forceinline__ void block_exit(unsigned* work, int total_blocks, int slices, unsigned block_q,
                                           unsigned block_r) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const int g = (int)(blockIdx.x % (unsigned)slices);
        const unsigned prev = atomicAdd(&work[kWorkGroupExit + g * kWorkStride], 1u);
        if (prev == home_blocks(g, block_q, block_r) - 1u) {
            const unsigned groups = (unsigned)min(slices, total_blocks);
            const unsigned gprev = atomicAdd(&work[kWorkExited], 1u);
            if (gprev == groups - 1u) {
                for (int k = 0; k < kWorkSlices; ++k) {
                    atomicExch(&work[k * kWorkStride], 0u);
                    atomicExch(&work[kWorkGroupExit + k * kWorkStride], 0u);
                }
                atomicExch(&work[kWorkDrained], 0u);
                atomicExch(&work[kWorkExited], 0u);
            }
        }
    }
}

struct CwbvhParams {
    const float4* __restrict__ nodes;
    const float4* __restrict__ tris;
    const int* __restrict__ prim_map;
    const Ray4* __restrict__ rays;
    const int* __restrict__ order;
    float4* __restrict__ hits;
    const cw::BatchJob* __restrict__ jobs;
    unsigned* work;                 // fetch counters, exited waves, overflow (kWork*)
    unsigned long long* stats;      // {nodes, triangles}
    int n;                          // rays (single) or total chunks (multi)
    int slices;                     // single scene: contiguous ray ranges, one per XCD
    int persistent;                 // 0: one ray per thread, no refill
    int num_jobs;
    int stride;                     // float4 per node: 5 or 8
    int lds_depth;
    int chunk;
    int idle_lanes;
    int idle_budget;
    int postpone_pct;
    int prim_id;
    int total_waves;
    int total_blocks;
    unsigned item_q, item_r;        // n = item_q * slices + item_r
    unsigned block_q, block_r;      // total_blocks = block_q * slices + block_r
    int spec_limit;                 // single scene: speculation_limit of its tree
    int lds_only;                   // host side: the tree's stack need fits lds_depth (LDS-only kernels)
    // SRC = kRaysFromSurfels: generate ray r from surfel r / rays_per_texel
    const float4* __restrict__ surf_pos;
    const float4* __restrict__ surf_nrm;
    const float* __restrict__ surf_rad;
    int rays_per_texel;
    float gen_tmin, gen_tmax;
    unsigned long long seed;
    unsigned* texel_hits;           // non-null: count hits per surfel instead of writing hits
};

constexpr int kRaysFromBuffer = 0;
constexpr int kRaysFromSurfels = 1;

__device__ __forceinline__ const float4* node_at(const float4* nodes, unsigned i, unsigned stride_sh) {
    return nodes + ((i << 2) + (i << stride_sh));
}

// Quantized slab test of 4 children (one meta dword, one byte per child).
__device__ __forceinline__ unsigned child_hits4(unsigned meta4, unsigned lox, unsigned loy, unsigned loz,
                                                unsigned hix, unsigned hiy, unsigned hiz, float adjx, float adjy,
                                                float adjz, float origx, float origy, float origz, float tmin,
                                                float tmax, unsigned octinv4) {
    const unsigned is_inner4 = (meta4 & (meta4 << 1)) & 0x10101010u;
    const unsigned ones = is_inner4 >> 4;                 // 0x01 per inner child byte
    const unsigned inner_mask4 = (ones << 8) - ones;      // 0xFF per inner child byte
    const unsigned bit_index4 = (meta4 ^ (octinv4 & inner_mask4)) & 0x1F1F1F1Fu;
    const unsigned child_bits4 = (meta4 >> 5) & 0x07070707u;
    unsigned hitmask = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int sh = 8 * j;
        const float tnx = __builtin_fmaf((float)((lox >> sh) & 0xFFu), adjx, origx);
        const float tny = __builtin_fmaf((float)((loy >> sh) & 0xFFu), adjy, origy);
        const float tnz = __builtin_fmaf((float)((loz >> sh) & 0xFFu), adjz, origz);
        const float tfx = __builtin_fmaf((float)((hix >> sh) & 0xFFu), adjx, origx);
        const float tfy = __builtin_fmaf((float)((hiy >> sh) & 0xFFu), adjy, origy);
        const float tfz = __builtin_fmaf((float)((hiz >> sh) & 0xFFu), adjz, origz);
        // int-ordered combine with tmin/tmax as well: equal to fmaxf/fminf for
        // tmin, tmax >= +0 and non-NaN slabs, without NaN canonicalization ops
        const float cmin = imax2(imax3(tnx, tny, tnz), tmin);
        const float cmax = imin2(imin3(tfx, tfy, tfz), tmax);
        const unsigned bits = (child_bits4 >> sh) & 0xFFu;
        const unsigned idx = (bit_index4 >> sh) & 0xFFu;
        hitmask |= (cmin <= cmax) ? (bits << idx) : 0u;
    }
    return hitmask;
}

// Inner-loop schedules (cwbvh_trace_config.loop_mode).
//  kLoopGroup: per iteration one node visit, then every triangle of the
//    group it produced (the reference's order, :93-340), with optional
//    postponing of the triangle group.
//  kLoopInterleave: per iteration a lane either visits a node (when it holds
//    no triangles) or tests ONE triangle; the per-ray sequence of visits and
//    tests is the same as kLoopGroup, but a lane with a long triangle group no
//    longer holds the other 63 lanes in a triangle loop.
//  kLoopSpeculate: as kLoopInterleave, and a lane that holds triangles still
//    visits a node in the same iteration; the new node's triangle group is
//    pushed on the stack and tested later, so node visits keep every lane
//    with inner children busy.  With postpone_pct > 0 the wave also holds
//    back its triangle step until that share of its lanes hold triangles and
//    have no node to visit (or no lane can visit a node), so the step runs
//    with more lanes on.
//    A lane speculates only while its stack holds fewer than spec_limit
//    entries (capacity minus the tree's stack need), so the parked groups
//    never push the stack past its capacity.
//  loop_mode 3 (a per-wave queue of triangle tests run 64 at a time) is
//    refused by validate_config: its first GPU run faulted the device
//    together with another untested path, and the cause was not shown, so the
//    schedule was taken out.
constexpr int kLoopGroup = 0;
constexpr int kLoopInterleave = 1;
constexpr int kLoopSpeculate = 2;
constexpr int kLoopQueue = 3;

// Register budget (cwbvh_trace_config.occupancy).  The plain interleaved
// kernel (one scene, no counters, no prefetch) takes 75 VGPRs unconstrained
// when it reads rays from a buffer (6 waves per SIMD), 91 when it makes them
// from surfels (5 waves).  Both are also built for 7 waves (72 VGPRs; 2 and
// 10 registers of cold state spilled), which traces launches of tens of
// millions of AO rays 3-5% faster and launches of a million rays 6-7% slower
// (more waves, each slower, lengthen the tail of a short launch).  The other
// instances stay unconstrained: the batch, counting and prefetching kernels
// spill 10 to 32 registers at 7 waves, and the grouped and speculative
// schedules serve scenes that stream from HBM, where spills and the extra
// wave cost more.
constexpr int kOccupancyAuto = 0;
constexpr int kOccupancyFree = 1;
constexpr int kOccupancy7 = 7;
constexpr long long kOccupancy7MinRays = 1ll << 22;  // automatic choice: launches this large

#ifndef CWBVH_TRACE_ATTR
#define CWBVH_TRACE_ATTR __attribute__((amdgpu_waves_per_eu(W7 ? 7 : 1)))
#endif

template <bool ANY, bool STATS, bool MULTI, bool PREFETCH, int LOOP, int SRC = kRaysFromBuffer, bool W7 = false,
          bool LDSO = false>
__global__ __launch_bounds__(kMaxBlock) CWBVH_TRACE_ATTR void cwbvh_trace(CwbvhParams P) {
    static_assert(!W7 || (LOOP == kLoopInterleave && !MULTI && !STATS && !PREFETCH),
                  "the 7-wave build is made for the plain interleaved kernels only");
    // LDSO: the launch's trees (a batch: every job's tree) need at most
    // lds_depth stack entries, so the stack lives in LDS only (interleaved
    // schedule: only node groups are pushed, one per level of the tree)
    static_assert(!LDSO || (LOOP == kLoopInterleave && !STATS), "LDS-only stacks serve the interleaved schedule");
    extern __shared__ unsigned long long lds_stack_raw[];
    CWD_LDS unsigned long long* lds_stack = lds_ptr(lds_stack_raw);
    unsigned long long spill[kStackCapacity];
    const StackIndex sx((int)threadIdx.x, P.lds_depth);
    const int lane = lane_id();

    // per-lane ray state
    int out_idx = -1;
    unsigned out_texel = 0;
    const float4* nodes = P.nodes;
    const float4* tris = P.tris;
    float4* hits = P.hits;
    const int* prim_map = P.prim_map;
    // node i starts at float4 (i << 2) + (i << stride_sh): stride 5 or 8
    // without a quarter-rate integer multiply
    unsigned stride_sh = P.stride == 8 ? 2u : 0u;
    int spec_limit = P.spec_limit;
    float3 o = make_float3(0.f, 0.f, 0.f), d = o;
    float tmin = 0.f, tmax = 0.f, idx = 0.f, idy = 0.f, idz = 0.f;
    unsigned octinv = 0, octinv4 = 0;
    uint2 ng = make_uint2(0u, 0u), tg = make_uint2(0u, 0u);
    int sp = 0;
    int hit_addr = -1;
    float hu = 0.f, hv = 0.f;
    unsigned long long st_nodes = 0, st_tris = 0;
    // STATS only: {slots, uses} of the loop iteration, the node code and the
    // triangle code; a slot is one lane of a wave instruction stream that ran
    // the code (64 per wave pass), a use is a lane that needed it
    unsigned long long pr_it = 0, pr_act = 0, pr_nd = 0, pr_ndu = 0, pr_tr = 0, pr_tru = 0;
    auto prof = [&](bool used, unsigned long long& slots, unsigned long long& uses) {
        if (STATS) {
            const unsigned long long act = __ballot(1);
            if (__any(used) && lane == __ffsll((long long)act) - 1) slots += 64;
            uses += used ? 1 : 0;
        }
    };
    bool have_pf = false;
    float4 pf0, pf1, pf2, pf3, pf4;
    // A lane is busy from loading a ray until writing its hit.  (Deriving
    // "terminated" from an empty stack and empty groups, as the reference
    // does at src/TraversalKernelCWBVH.cu:60, misreads a lane whose last
    // stack entry was a postponed triangle group just popped into ng.)
    bool busy = false;

    // wave-uniform work pool; single-scene launches split the rays into
    // `slices` contiguous ranges and a wave drains the range of its XCD
    // (workgroups are dispatched round-robin over the XCDs) before stealing
    unsigned pool_next = 0, pool_end = 0;
    int pool_job = 0;
    bool exhausted = false;
    const int home = (int)(blockIdx.x % (unsigned)P.slices);
    int slice_k = 0;
    const unsigned first_index =
        (blockIdx.x / (unsigned)P.slices) * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    while (true) {
        const bool terminated = !busy;
        unsigned long long need = __ballot(terminated);
        bool got = false;
        unsigned my = 0;
        int my_job = 0;
        while (need != 0ull) {
            if (pool_next >= pool_end) {
                if (exhausted) break;
                const int leader = __ffsll((long long)need) - 1;
                if (!MULTI && !P.persistent) {
                    // one ray per thread: the only "pool" is this wave's own 64 indices
                    if (slice_k > 0) {
                        exhausted = true;
                        break;
                    }
                    slice_k = 1;
                    const unsigned first = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u));
                    pool_next = min(first, (unsigned)P.n);
                    pool_end = min(first + 64u, (unsigned)P.n);
                    if (pool_next >= pool_end) {
                        exhausted = true;
                        break;
                    }
                } else if (MULTI) {
                    // claim one chunk index from the home slice of the chunk
                    // range, then from the others
                    bool found = false;
                    unsigned base = 0;
                    while (slice_k < P.slices) {
                        const int sl = home + slice_k - (home + slice_k >= P.slices ? P.slices : 0);
                        const unsigned lo = slice_begin(sl, P.item_q, P.item_r);
                        const unsigned hi = slice_begin(sl + 1, P.item_q, P.item_r);
                        unsigned claim = 0;
                        if (lane == leader) claim = atomicAdd(&P.work[sl * kWorkStride], 1u);
                        claim = __shfl(claim, leader);
                        if (claim < hi - lo) {
                            base = lo + claim;
                            found = true;
                            break;
                        }
                        ++slice_k;
                    }
                    if (!found) {
                        exhausted = true;
                        break;
                    }
                    // job j owns chunks [jobs[j].chunk_begin, jobs[j+1].chunk_begin)
                    int lo = 0, hi = P.num_jobs - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (P.jobs[mid].chunk_begin <= base) lo = mid;
                        else hi = mid - 1;
                    }
                    pool_job = lo;
                    const unsigned first = (base - P.jobs[lo].chunk_begin) * (unsigned)P.chunk;
                    pool_next = first;
                    pool_end = min(first + (unsigned)P.chunk, (unsigned)P.jobs[lo].n);
                } else {
                    bool found = false;
                    // the wave's first claim: nothing claimed yet; slices found dry in this refill
                    const bool first_claim = slice_k == 0 && pool_end == 0u;
                    unsigned drained = 0;
                    while (slice_k < P.slices) {
                        const int sl = home + slice_k - (home + slice_k >= P.slices ? P.slices : 0);
                        if (slice_k > 0 && ((drained >> sl) & 1u)) {
                            ++slice_k;
                            continue;
                        }
                        const unsigned lo = slice_begin(sl, P.item_q, P.item_r);
                        const unsigned hi = slice_begin(sl + 1, P.item_q, P.item_r);
                        const unsigned base = slice_claim(P.work, sl, (unsigned)P.chunk, first_claim && slice_k == 0, first_index,
                                                          P.block_q, P.block_r, lane);
                        if (base < hi - lo) {
                            pool_next = lo + base;
                            pool_end = lo + min(base + (unsigned)P.chunk, hi - lo);
                            found = true;
                            break;
                        }
                        drained = mark_drained(P.work, sl, lane);
                        ++slice_k;
                    }
                    if (!found) {
                        exhausted = true;
                        break;
                    }
                }
            }
            const unsigned avail = pool_end - pool_next;
            const unsigned rank = (unsigned)__popcll(need & lanemask_lt());
            if (((need >> lane) & 1ull) && rank < avail) {
                my = pool_next + rank;
                my_job = pool_job;
                got = true;
            }
            pool_next += min((unsigned)__popcll(need), avail);
            need = __ballot(terminated && !got);
        }
        if (terminated && !got) break;  // no work left for this lane

        if (got) {
            const Ray4* rays = P.rays;
            const int* order = P.order;
            unsigned ray_count = (unsigned)P.n;
            if (MULTI) {
                const cw::BatchJob& J = P.jobs[my_job];
                nodes = (const float4*)J.nodes;
                tris = (const float4*)J.woop;
                hits = (float4*)J.hits;
                prim_map = J.prim_idx;
                stride_sh = J.node_stride == 8 ? 2u : 0u;
                spec_limit = J.spec_limit;
                rays = (const Ray4*)J.rays;
                order = J.order;
                ray_count = (unsigned)J.n;
            }
            if constexpr (SRC == kRaysFromSurfels) {
                // the ray is made here from its surfel instead of read back
                const int ri = (int)my;
                out_idx = ri;
                out_texel = (unsigned)ri / (unsigned)P.rays_per_texel;
                const float4 sp = P.surf_pos[out_texel], sn = P.surf_nrm[out_texel];
                cwao::F3 go, gd;
                cwao::surfel_ray(cwao::F3{sp.x, sp.y, sp.z}, cwao::F3{sn.x, sn.y, sn.z}, P.surf_rad[out_texel],
                                 P.seed, (unsigned long long)ri, go, gd);
                o = make_float3(go.x, go.y, go.z);
                d = make_float3(gd.x, gd.y, gd.z);
                tmin = P.gen_tmin;
                tmax = P.gen_tmax;
            } else {
                int ri = order ? order[my] : (int)my;
                // an order entry outside the launch's rays traces ray `my`
                // itself, so a malformed order cannot read or write out of bounds
                if ((unsigned)ri >= ray_count) ri = (int)my;
                out_idx = ri;
                // rays are read once: keep them out of the caches the BVH lives in
                const float4 ro = nt_load4(&rays[ri].o);
                const float4 rd = nt_load4(&rays[ri].d);
                o = make_float3(ro.x, ro.y, ro.z);
                d = make_float3(rd.x, rd.y, rd.z);
                tmin = ro.w;
                tmax = rd.w;
            }
            safe_rcp3(d, idx, idy, idz);
            const unsigned oct = ((d.x < 0.f ? 1u : 0u) << 2) | ((d.y < 0.f ? 1u : 0u) << 1) | (d.z < 0.f ? 1u : 0u);
            octinv = 7u - oct;
            octinv4 = octinv * 0x01010101u;
            ng = make_uint2(0u, 0x80000000u);
            tg = make_uint2(0u, 0u);
            sp = 0;
            hit_addr = -1;
            hu = hv = 0.f;
            have_pf = false;
            busy = true;
        }

        // stack push of a node or triangle group (overflow is flagged, not written)
        auto push = [&](uint2 e) {
            if constexpr (LDSO) {
                if (sp < P.lds_depth) {
                    lds_stack[(sp << 6) + sx.col] = pack2(e);
                    ++sp;
                } else {
                    P.work[kWorkOverflow] = 1u;
                }
            } else {
                if (sp < kStackCapacity) {
                    stack_put(lds_stack, spill, sx, sp, pack2(e));
                    ++sp;
                } else {
                    P.work[kWorkOverflow] = 1u;
                }
            }
        };
        auto stack_at = [&](int i) -> unsigned long long {
            if constexpr (LDSO) return lds_stack[(i << 6) + sx.col];
            else return stack_get(lds_stack, spill, sx, i);
        };
        // Visit the highest-priority child of ng: the remainder of ng goes on
        // the stack, ng becomes the child's inner children, the child's leaf
        // triangles are returned.
        auto visit_node = [&]() -> uint2 {
            const unsigned imask = ng.y;
            const int bit = 31 - __clz((int)ng.y);
            const unsigned base = ng.x;
            ng.y &= ~(1u << bit);
            if (ng.y > 0x00FFFFFFu) push(ng);
            float4 n0, n1, n2, n3, n4;
            if (PREFETCH && have_pf) {
                n0 = pf0; n1 = pf1; n2 = pf2; n3 = pf3; n4 = pf4;
                have_pf = false;
            } else {
                const unsigned slot = (unsigned)(bit - 24) ^ octinv;
                const unsigned rel = (unsigned)__popc(imask & ~(0xFFFFFFFFu << slot));
                const float4* np = node_at(nodes, base + rel, stride_sh);
                n0 = np[0]; n1 = np[1]; n2 = np[2]; n3 = np[3]; n4 = np[4];
            }
            if (STATS) ++st_nodes;
            const unsigned eb = f2u(n0.w);
            const int ex = (int)(signed char)(eb & 0xFFu);
            const int ey = (int)(signed char)((eb >> 8) & 0xFFu);
            const int ez = (int)(signed char)((eb >> 16) & 0xFFu);
            const float adjx = u2f((unsigned)(ex + 127) << 23) * idx;
            const float adjy = u2f((unsigned)(ey + 127) << 23) * idy;
            const float adjz = u2f((unsigned)(ez + 127) << 23) * idz;
            const float origx = -(o.x - n0.x) * idx;
            const float origy = -(o.y - n0.y) * idy;
            const float origz = -(o.z - n0.z) * idz;
            const bool sgx = idx < 0.f, sgy = idy < 0.f, sgz = idz < 0.f;
            unsigned hitmask;
            {
                const unsigned qlx = f2u(n2.x), qly = f2u(n2.z), qlz = f2u(n3.x);
                const unsigned qhx = f2u(n3.z), qhy = f2u(n4.x), qhz = f2u(n4.z);
                hitmask = child_hits4(f2u(n1.z), sgx ? qhx : qlx, sgy ? qhy : qly, sgz ? qhz : qlz,
                                      sgx ? qlx : qhx, sgy ? qly : qhy, sgz ? qlz : qhz, adjx, adjy, adjz,
                                      origx, origy, origz, tmin, tmax, octinv4);
            }
            {
                const unsigned qlx = f2u(n2.y), qly = f2u(n2.w), qlz = f2u(n3.y);
                const unsigned qhx = f2u(n3.w), qhy = f2u(n4.y), qhz = f2u(n4.w);
                hitmask |= child_hits4(f2u(n1.w), sgx ? qhx : qlx, sgy ? qhy : qly, sgz ? qhz : qlz,
                                       sgx ? qlx : qhx, sgy ? qly : qhy, sgz ? qlz : qhz, adjx, adjy, adjz,
                                       origx, origy, origz, tmin, tmax, octinv4);
            }
            ng.x = f2u(n1.x);
            ng.y = (hitmask & 0xFF000000u) | (eb >> 24);
            if (PREFETCH && ng.y > 0x00FFFFFFu) {
                // the next node visit takes the highest hit child: start its
                // loads now so they overlap the triangle tests in between
                const int nb = 31 - __clz((int)ng.y);
                const unsigned nslot = (unsigned)(nb - 24) ^ octinv;
                const unsigned nrel = (unsigned)__popc(ng.y & ~(0xFFFFFFFFu << nslot));
                const float4* np = node_at(nodes, ng.x + nrel, stride_sh);
                pf0 = np[0]; pf1 = np[1]; pf2 = np[2]; pf3 = np[3]; pf4 = np[4];
                have_pf = true;
            }
            return make_uint2(f2u(n1.y), hitmask & 0x00FFFFFFu);
        };
        // Test the highest triangle of tg; true when an any-hit query is done.
        auto test_triangle = [&]() -> bool {
            const int ti = 31 - __clz((int)tg.y);
            const int addr = (int)(tg.x * 3u + (unsigned)ti * 3u);
            const float4 v0 = tris[addr], v1 = tris[addr + 1], v2 = tris[addr + 2];
            if (STATS) ++st_tris;
            tg.y &= ~(1u << ti);
            float t, u, v;
            if (woop_test<ANY>(v0, v1, v2, o, d, tmin, tmax, addr, hit_addr, t, u, v)) {
                tmax = t;
                hit_addr = addr;
                hu = u;
                hv = v;
                return ANY;
            }
            return false;
        };

        int lost = 0;
        while (true) {
            bool finished = false;
            prof(true, pr_it, pr_act);
            if constexpr (LOOP == kLoopGroup) {
                prof(ng.y > 0x00FFFFFFu, pr_nd, pr_ndu);
                if (ng.y > 0x00FFFFFFu) {
                    tg = visit_node();
                } else {
                    // a postponed triangle group came off the stack
                    tg = ng;
                    ng = make_uint2(0u, 0u);
                }
                const int total = P.postpone_pct > 0 ? __popcll(__ballot(1)) : 0;
                while (tg.y != 0u) {
                    if (P.postpone_pct > 0) {
                        const int active = __popcll(__ballot(1));
                        if (active * 100 < total * P.postpone_pct) {
                            push(tg);
                            tg.y = 0u;
                            break;
                        }
                    }
                    prof(true, pr_tr, pr_tru);
                    if (test_triangle()) {
                        finished = true;
                        break;
                    }
                }
                if (!finished && ng.y <= 0x00FFFFFFu) {
                    if (sp > 0) {
                        --sp;
                        ng = unpack2(stack_at(sp));
                    } else {
                        finished = true;
                    }
                }
            } else {
                // node step: lanes without triangles (and, speculating, every
                // lane with inner children left) visit one node
                const bool spec_ok = LOOP == kLoopSpeculate && sp < spec_limit;
                const bool node_step = (spec_ok || tg.y == 0u) && ng.y > 0x00FFFFFFu;
                prof(node_step, pr_nd, pr_ndu);
                if (node_step) {
                    const uint2 ntg = visit_node();
                    if (tg.y != 0u) {
                        if (ntg.y != 0u) push(ntg);
                    } else {
                        tg = ntg;
                    }
                }
                // triangle step: one test (speculating with postpone_pct, only
                // once enough lanes hold triangles or a lane holds nothing else)
                bool run_tri = true;
                if (LOOP == kLoopSpeculate && P.postpone_pct > 0) {
                    // lanes that hold triangles and cannot visit a node next
                    // iteration idle until the triangle step runs
                    const bool can_visit = ng.y > 0x00FFFFFFu && (tg.y == 0u || sp < spec_limit);
                    const int stuck = __popcll(__ballot(tg.y != 0u && !can_visit));
                    run_tri = !__any(can_visit) || stuck * 100 >= P.postpone_pct * __popcll(__ballot(1));
                }
                prof(run_tri && tg.y != 0u, pr_tr, pr_tru);
                if (run_tri && tg.y != 0u) finished = test_triangle();
                // out of inner children: take the next group off the stack
                if (!finished && ng.y <= 0x00FFFFFFu) {
                    if (tg.y == 0u) {
                        if (sp > 0) {
                            --sp;
                            const uint2 e = unpack2(stack_at(sp));
                            if (e.y > 0x00FFFFFFu) {
                                ng = e;
                            } else {
                                tg = e;
                                ng = make_uint2(0u, 0u);
                            }
                        } else {
                            finished = true;
                        }
                    } else if (LOOP == kLoopSpeculate && sp > 0) {
                        const uint2 e = unpack2(stack_at(sp - 1));
                        if (e.y > 0x00FFFFFFu) {
                            --sp;
                            ng = e;
                        }
                    }
                }
            }
            if (finished) {
                if (SRC == kRaysFromSurfels && P.texel_hits) {
                    if (hit_addr >= 0) atomicAdd(&P.texel_hits[out_texel], 1u);
                } else {
                    const int out_id = (hit_addr >= 0 && P.prim_id) ? prim_map[hit_addr] : hit_addr;
                    nt_store4(make_float4(tmax, i2f(out_id), hu, hv), &hits[out_idx]);
                }
                sp = 0;
                ng = make_uint2(0u, 0u);
                tg = make_uint2(0u, 0u);
                have_pf = false;
                busy = false;
                break;
            }
            lost += __popcll(~__ballot(1)) - P.idle_lanes;
            if (lost >= P.idle_budget) break;
        }
    }

    if (STATS) {
        atomicAdd(&P.stats[0], st_nodes);
        atomicAdd(&P.stats[1], st_tris);
        unsigned long long* pr = reinterpret_cast<unsigned long long*>(P.work + kWorkProfile);
        atomicAdd(&pr[0], pr_it);
        atomicAdd(&pr[1], pr_act);
        atomicAdd(&pr[2], pr_nd);
        atomicAdd(&pr[3], pr_ndu);
        atomicAdd(&pr[4], pr_tr);
        atomicAdd(&pr[5], pr_tru);
    }
    // one-ray-per-thread launches claim nothing and skip the exit count
    if (!MULTI && !P.persistent) return;
    if constexpr (MULTI) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned prev = atomicAdd(&P.work[kWorkExited], 1u);
            if (prev == (unsigned)P.total_blocks - 1u) {
                for (int k = 0; k < kWorkSlices; ++k) atomicExch(&P.work[k * kWorkStride], 0u);
                atomicExch(&P.work[kWorkExited], 0u);
            }
        }
    } else {
        block_exit(P.work, P.total_blocks, P.slices, P.block_q, P.block_r);
    }
}

struct Bvh2Params {
    const float4* __restrict__ nodes;
    const float4* __restrict__ tris;
    const int* __restrict__ prim_map;
    const Ray4* __restrict__ rays;
    const int* __restrict__ order;
    float4* __restrict__ hits;
    unsigned* work;
    unsigned long long* stats;
    int n;
    int lds_depth;
    int woop_last;   // last valid float4 index of tris (loads past a leaf terminator clamp here)
    int prim_id;
    // persistent launches (bvh2_trace_persistent) only
    int slices;
    int chunk;
    int idle_lanes;
    int idle_budget;
    int total_blocks;
    unsigned item_q, item_r;   // n = item_q * slices + item_r
    unsigned block_q, block_r; // total_blocks = block_q * slices + block_r
};

// Slab test of one BVH2 child box (reference src/TraversalKernelBVH2.cu:14-24).
__device__ __forceinline__ bool box2(float lox, float hix, float loy, float hiy, float loz, float hiz, float idx,
                                     float idy, float idz, float oodx, float oody, float oodz, float tmin, float tmax,
                                     float& dist) {
    const float lx = __builtin_fmaf(lox, idx, -oodx), hx = __builtin_fmaf(hix, idx, -oodx);
    const float ly = __builtin_fmaf(loy, idy, -oody), hy = __builtin_fmaf(hiy, idy, -oody);
    const float lz = __builtin_fmaf(loz, idz, -oodz), hz = __builtin_fmaf(hiz, idz, -oodz);
    const float smin = imax2(imax3(imin2(lx, hx), imin2(ly, hy), imin2(lz, hz)), tmin);
    const float smax = imin2(imin3(imax2(lx, hx), imax2(ly, hy), imax2(lz, hz)), tmax);
    dist = smin;
    return smin <= smax;
}

template <bool ANY, bool STATS>
__global__ __launch_bounds__(kMaxBlock) void bvh2_trace(Bvh2Params P) {
    constexpr int SENTINEL = 0x76543210;
    extern __shared__ int lds_stack_raw_i[];
    CWD_LDS int* lds_stack_i = lds_ptr(lds_stack_raw_i);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    int spill[kBvh2StackCapacity];
    const StackIndex sx((int)threadIdx.x, P.lds_depth);

    int ri = P.order ? P.order[i] : i;
    if ((unsigned)ri >= (unsigned)P.n) ri = i;  // malformed order: stay in bounds
    const float4 ro = nt_load4(&P.rays[ri].o);
    const float4 rd = nt_load4(&P.rays[ri].d);
    const float3 o = make_float3(ro.x, ro.y, ro.z), d = make_float3(rd.x, rd.y, rd.z);
    const float tmin = ro.w;
    float hitT = rd.w;
    float idx, idy, idz;
    safe_rcp3(d, idx, idy, idz);
    const float oodx = o.x * idx, oody = o.y * idy, oodz = o.z * idz;
    unsigned long long st_nodes = 0, st_tris = 0;

    int sp = 0;
    stack_put(lds_stack_i, spill, sx, 0, SENTINEL);
    int node = 0, leaf = 0, hit_addr = -1;
    float hu = 0.f, hv = 0.f;
    bool done = false;
    while (node != SENTINEL) {
        leaf = 0;
        while (node != SENTINEL && node >= 0) {
            const float4* np = P.nodes + node;
            const float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
            if (STATS) ++st_nodes;
            float d0, d1;
            const bool t0 = box2(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, idx, idy, idz, oodx, oody, oodz, tmin, hitT, d0);
            const bool t1 = box2(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, idx, idy, idz, oodx, oody, oodz, tmin, hitT, d1);
            const int c0 = f2i(n3.x), c1 = f2i(n3.y);
            if (!t0 && !t1) {
                node = stack_get(lds_stack_i, spill, sx, sp);
                --sp;
            } else {
                node = t0 ? c0 : c1;
                if (t0 && t1) {
                    int far = c1;
                    if (d1 < d0) {
                        far = node;
                        node = c1;
                    }
                    if (sp + 1 < kBvh2StackCapacity) {
                        ++sp;
                        stack_put(lds_stack_i, spill, sx, sp, far);
                    } else {
                        P.work[kWorkOverflow] = 1u;
                    }
                }
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = stack_get(lds_stack_i, spill, sx, sp);
                --sp;
            }
            if (!__any(leaf >= 0)) break;
        }
        while (leaf < 0) {
            for (int a = ~leaf;; a += 3) {
                const float4 v0 = P.tris[a];
                const float4 v1 = P.tris[min(a + 1, P.woop_last)];
                const float4 v2 = P.tris[min(a + 2, P.woop_last)];
                if (f2u(v0.x) == 0x80000000u) break;
                if (STATS) ++st_tris;
                float t, u, v;
                if (woop_test<ANY>(v0, v1, v2, o, d, tmin, hitT, a, hit_addr, t, u, v)) {
                    hitT = t;
                    hit_addr = a;
                    hu = u;
                    hv = v;
                    if (ANY) {
                        done = true;
                        break;
                    }
                }
            }
            if (done) break;
            leaf = node;
            if (node < 0) {
                node = stack_get(lds_stack_i, spill, sx, sp);
                --sp;
            }
        }
        if (done) break;
    }
    const int out_id = (hit_addr >= 0 && P.prim_id) ? P.prim_map[hit_addr] : hit_addr;
    nt_store4(make_float4(hitT, i2f(out_id), hu, hv), &P.hits[ri]);
    if (STATS) {
        atomicAdd(&P.stats[0], st_nodes);
        atomicAdd(&P.stats[1], st_tris);
    }
}

// Persistent BVH2 traversal.  One thread per ray leaves a lane idle from the
// moment its ray ends until the slowest ray of its wave ends (33 M AO rays on
// a 1 M-triangle scene of icospheres: 34% VALU lane use).  Here a grid sized to the device keeps its waves and
// hands rays to idle lanes: rays are claimed `chunk` at a time from `slices`
// contiguous ranges (a wave starts with the range of its XCD) and given to the
// idle lanes by ballot rank.  Each lane runs the while-while rounds of
// bvh2_trace on its ray unchanged (a round: node steps until every lane of the
// wave holds a leaf, then the leaves), so each ray's visits and tests, and its
// hit, are the same.  After each round a wave counts its idle lanes; once it
// has lost `idle_budget` lane-rounds beyond `idle_lanes` per round, it leaves
// the rounds to refill.  The last block to exit resets the claim counters.
template <bool ANY, bool STATS>
__global__ __launch_bounds__(kMaxBlock) void bvh2_trace_persistent(Bvh2Params P) {
    constexpr int SENTINEL = 0x76543210;
    extern __shared__ int lds_stack_raw_i[];
    CWD_LDS int* lds_stack_i = lds_ptr(lds_stack_raw_i);
    int spill[kBvh2StackCapacity];
    const StackIndex sx((int)threadIdx.x, P.lds_depth);
    const int lane = lane_id();

    int out_idx = 0;
    float3 o = make_float3(0.f, 0.f, 0.f), d = o;
    float tmin = 0.f, hitT = 0.f, idx = 0.f, idy = 0.f, idz = 0.f, oodx = 0.f, oody = 0.f, oodz = 0.f;
    int sp = 0, node = SENTINEL, leaf = 0, hit_addr = -1;
    float hu = 0.f, hv = 0.f;
    bool busy = false;
    unsigned long long st_nodes = 0, st_tris = 0;

    // wave-uniform claim state
    unsigned pool_next = 0, pool_end = 0;
    bool exhausted = false;
    const int home = (int)(blockIdx.x % (unsigned)P.slices);
    int slice_k = 0;
    const unsigned first_index =
        (blockIdx.x / (unsigned)P.slices) * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    while (true) {
        const bool idle = !busy;
        unsigned long long need = __ballot(idle);
        bool got = false;
        unsigned my = 0;
        while (need != 0ull) {
            if (pool_next >= pool_end) {
                if (exhausted) break;
                bool found = false;
                // the wave's first claim: nothing claimed yet; slices found dry in this refill
                const bool first_claim = slice_k == 0 && pool_end == 0u;
                unsigned drained = 0;
                while (slice_k < P.slices) {
                    const int sl = home + slice_k - (home + slice_k >= P.slices ? P.slices : 0);
                    if (slice_k > 0 && ((drained >> sl) & 1u)) {
                        ++slice_k;
                        continue;
                    }
                    const unsigned lo = slice_begin(sl, P.item_q, P.item_r);
                    const unsigned hi = slice_begin(sl + 1, P.item_q, P.item_r);
                    const unsigned base = slice_claim(P.work, sl, (unsigned)P.chunk, first_claim && slice_k == 0, first_index,
                                                      P.block_q, P.block_r, lane);
                    if (base < hi - lo) {
                        pool_next = lo + base;
                        pool_end = lo + min(base + (unsigned)P.chunk, hi - lo);
                        found = true;
                        break;
                    }
                    drained = mark_drained(P.work, sl, lane);
                    ++slice_k;
                }
                if (!found) {
                    exhausted = true;
                    break;
                }
            }
            const unsigned avail = pool_end - pool_next;
            const unsigned rank = (unsigned)__popcll(need & lanemask_lt());
            if (((need >> lane) & 1ull) && rank < avail) {
                my = pool_next + rank;
                got = true;
            }
            pool_next += min((unsigned)__popcll(need), avail);
            need = __ballot(idle && !got);
        }
        if (idle && !got) break;  // no rays left for this lane

        if (got) {
            int ri = P.order ? P.order[my] : (int)my;
            if ((unsigned)ri >= (unsigned)P.n) ri = (int)my;  // malformed order: stay in bounds
            out_idx = ri;
            const float4 ro = nt_load4(&P.rays[ri].o);
            const float4 rd = nt_load4(&P.rays[ri].d);
            o = make_float3(ro.x, ro.y, ro.z);
            d = make_float3(rd.x, rd.y, rd.z);
            tmin = ro.w;
            hitT = rd.w;
            safe_rcp3(d, idx, idy, idz);
            oodx = o.x * idx;
            oody = o.y * idy;
            oodz = o.z * idz;
            sp = 0;
            stack_put(lds_stack_i, spill, sx, 0, SENTINEL);
            node = 0;
            hit_addr = -1;
            hu = hv = 0.f;
            busy = true;
        }

        int lost = 0;
        while (true) {
            bool done = false;
            leaf = 0;
            while (node != SENTINEL && node >= 0) {
                const float4* np = P.nodes + node;
                const float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
                if (STATS) ++st_nodes;
                float d0, d1;
                const bool t0 = box2(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, idx, idy, idz, oodx, oody, oodz, tmin, hitT, d0);
                const bool t1 = box2(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, idx, idy, idz, oodx, oody, oodz, tmin, hitT, d1);
                const int c0 = f2i(n3.x), c1 = f2i(n3.y);
                if (!t0 && !t1) {
                    node = stack_get(lds_stack_i, spill, sx, sp);
                    --sp;
                } else {
                    node = t0 ? c0 : c1;
                    if (t0 && t1) {
                        int far = c1;
                        if (d1 < d0) {
                            far = node;
                            node = c1;
                        }
                        if (sp + 1 < kBvh2StackCapacity) {
                            ++sp;
                            stack_put(lds_stack_i, spill, sx, sp, far);
                        } else {
                            P.work[kWorkOverflow] = 1u;
                        }
                    }
                }
                if (node < 0 && leaf >= 0) {
                    leaf = node;
                    node = stack_get(lds_stack_i, spill, sx, sp);
                    --sp;
                }
                if (!__any(leaf >= 0)) break;
            }
            while (leaf < 0) {
                for (int a = ~leaf;; a += 3) {
                    const float4 v0 = P.tris[a];
                    const float4 v1 = P.tris[min(a + 1, P.woop_last)];
                    const float4 v2 = P.tris[min(a + 2, P.woop_last)];
                    if (f2u(v0.x) == 0x80000000u) break;
                    if (STATS) ++st_tris;
                    float t, u, v;
                    if (woop_test<ANY>(v0, v1, v2, o, d, tmin, hitT, a, hit_addr, t, u, v)) {
                        hitT = t;
                        hit_addr = a;
                        hu = u;
                        hv = v;
                        if (ANY) {
                            done = true;
                            break;
                        }
                    }
                }
                if (done) break;
                leaf = node;
                if (node < 0) {
                    node = stack_get(lds_stack_i, spill, sx, sp);
                    --sp;
                }
            }
            if (done || node == SENTINEL) {
                const int out_id = (hit_addr >= 0 && P.prim_id) ? P.prim_map[hit_addr] : hit_addr;
                nt_store4(make_float4(hitT, i2f(out_id), hu, hv), &P.hits[out_idx]);
                node = SENTINEL;
                busy = false;
                break;
            }
            lost += __popcll(~__ballot(1)) - P.idle_lanes;
            if (lost >= P.idle_budget) break;
        }
    }
    if (STATS) {
        atomicAdd(&P.stats[0], st_nodes);
        atomicAdd(&P.stats[1], st_tris);
    }
    block_exit(P.work, P.total_blocks, P.slices, P.block_q, P.block_r);
}

__global__ void copy_u4_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long step = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += step) dst[i] = src[i];
}

// VALU issue-rate probes: `iters` iterations of a fixed vector body on
// independent chains, no memory access but one store per lane at the end, so
// a full grid keeps every SIMD issuing vector instructions back to back.
// Timing one launch against its SQ_INSTS_VALU count gives the issue ceiling
// of that instruction mix:
//   kValuProbeFma   8 chains of FP32 fma;
//   kValuProbeInt   8 chains of 32-bit integer add / xor / min3 / max3;
//   kValuProbeSlab  two 4-child quantized slab tests per iteration (the
//                   node-visit body of cwbvh_trace: byte->float conversions,
//                   fma, integer-ordered min3/max3, mask logic) on operands
//                   that change every iteration.
constexpr int kValuProbeFma = 0;
constexpr int kValuProbeInt = 1;
constexpr int kValuProbeSlab = 2;

// Every binary32 pattern x, by magnitude: counts[0] += |x| in [2^-125, 2^125]
// and rcp_refined(x) differs from 1.0f / x; counts[1] += |x| in that range;
// counts[2] += |x| outside it and the two differ.
__global__ __launch_bounds__(256) void rcp_check_kernel(unsigned long long* __restrict__ counts) {
    const unsigned long long total = 1ull << 32;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    unsigned long long bad_in = 0, n_in = 0, bad_out = 0;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const float x = u2f((unsigned)i);
        const bool in_range = fabsf(x) >= 0x1p-125f && fabsf(x) <= 0x1p125f;
        const bool differ = f2u(rcp_refined(x)) != f2u(1.0f / x);
        bad_in += (in_range && differ) ? 1 : 0;
        n_in += in_range ? 1 : 0;
        bad_out += (!in_range && differ) ? 1 : 0;
    }
    atomicAdd(&counts[0], bad_in);
    atomicAdd(&counts[1], n_in);
    atomicAdd(&counts[2], bad_out);
}

template <int MODE>
__global__ __launch_bounds__(256) void valu_probe_kernel(float* __restrict__ out, long long iters, float a, float b) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    float f[8];
    unsigned u[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        f[k] = (float)(tid + k) * 1e-7f;
        u[k] = tid * 2654435761u + (unsigned)k * 40503u;
    }
    for (long long i = 0; i < iters; ++i) {
        if constexpr (MODE == kValuProbeFma) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int k = 0; k < 8; ++k) f[k] = __builtin_fmaf(f[k], a, b);
        } else if constexpr (MODE == kValuProbeInt) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int x = (int)(u[k] + (unsigned)r), y = (int)(u[(k + 1) & 7] ^ 0x5bd1e995u);
                    u[k] = (unsigned)max(max(x, y), (int)u[(k + 3) & 7]) ^ (unsigned)min(min(x, y), (int)u[(k + 5) & 7]);
                }
        } else {
            const unsigned s = (unsigned)i;
            const unsigned h0 = child_hits4(u[0] ^ s, u[1] + s, u[2] ^ s, u[3] + s, u[4] ^ s, u[5] + s, u[6] ^ s, f[0], f[1],
                                            f[2], f[3], f[4], f[5], a, b, u[7]);
            const unsigned h1 = child_hits4(u[7] ^ s, u[6] + s, u[5] ^ s, u[4] + s, u[3] ^ s, u[2] + s, u[1] ^ s, f[2], f[0],
                                            f[1], f[5], f[3], f[4], a, b, u[0]);
            u[0] ^= h0;
            u[7] += h1;
            f[0] = __builtin_fmaf(f[0], a, b);
            f[3] = __builtin_fmaf(f[3], a, b);
        }
    }
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += f[k] + (float)u[k];
    out[tid] = acc;
}

// ------------------------------------------------------------ occupancy

template <class K>
int resident_blocks(K kernel, int block, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t>, int> cache;
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_tuple((const void*)kernel, block, lds);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, lds) != hipSuccess) {
        (void)hipGetLastError();  // tolerated: one block per unit
        nb = 1;
    }
    if (nb < 1) nb = 1;
    cache[key] = nb;
    return nb;
}

// Items (rays or chunks) and grid blocks split over the work slices.
template <class Params>
void set_splits(Params& p) {
    const unsigned sl = (unsigned)std::max(1, p.slices);
    p.item_q = (unsigned)p.n / sl;
    p.item_r = (unsigned)p.n % sl;
    p.block_q = (unsigned)p.total_blocks / sl;
    p.block_r = (unsigned)p.total_blocks % sl;
}

template <bool ANY, bool STATS, bool MULTI, bool PREFETCH, int LOOP, int SRC = kRaysFromBuffer, bool W7 = false,
          bool LDSO = false>
int32_t launch_cwbvh_t(CwbvhParams& p, const cwbvh_trace_config& cfg, hipStream_t stream) {
    const int block = cfg.block_size;
    const size_t lds = (size_t)block * (size_t)cfg.lds_depth * sizeof(unsigned long long);
    auto kernel = cwbvh_trace<ANY, STATS, MULTI, PREFETCH, LOOP, SRC, W7, LDSO>;
    long long grid = cfg.grid_blocks;
    if (!MULTI && !p.persistent) {
        grid = ((long long)p.n + block - 1) / block;
        p.idle_budget = 0x7FFFFFFF;  // nothing to refill from
    } else if (grid <= 0) {
        grid = (long long)resident_blocks(kernel, block, lds) * cw::device_compute_units();
    }
    p.total_waves = (int)(grid * (block / 64));
    p.total_blocks = (int)grid;
    set_splits(p);
    if (p.chunk <= 0) {
        // automatic: about 32 claims per wave over the launch, a multiple of 64
        // rays, at most 512.  At least 128 once the launch averages 1024 rays
        // per wave; below that one ray per lane per claim balances the short
        // launch better (1-3 M AO rays: 2-5% faster than 128; from 8 M rays
        // on, 64 is slower, by 9% at 33 M)
        const long long per_wave = (long long)p.n / p.total_waves;
        long long c = (long long)p.n / ((long long)p.total_waves * 32);
        c = std::min<long long>(512, std::max<long long>(per_wave < 1024 ? 64 : 128, (c + 63) / 64 * 64));
        p.chunk = (int)c;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(block), lds, stream, p);
    CW_HIP(hipGetLastError());
    return CWBVH_OK;
}

template <bool ANY, bool STATS>
int32_t launch_bvh2_t(Bvh2Params& p, const cwbvh_trace_config& cfg, bool persistent, hipStream_t stream) {
    const int block = cfg.block_size;
    const size_t lds = (size_t)block * (size_t)cfg.lds_depth * sizeof(int);
    if (!persistent) {
        const long long grid = ((long long)p.n + block - 1) / block;
        hipLaunchKernelGGL((bvh2_trace<ANY, STATS>), dim3((unsigned)grid), dim3(block), lds, stream, p);
        CW_HIP(hipGetLastError());
        return CWBVH_OK;
    }
    auto kernel = bvh2_trace_persistent<ANY, STATS>;
    long long grid = cfg.grid_blocks;
    if (grid <= 0) grid = (long long)resident_blocks(kernel, block, lds) * cw::device_compute_units();
    p.total_blocks = (int)grid;
    p.slices = cfg.work_slices;
    p.idle_lanes = cfg.refill_idle;
    p.idle_budget = cfg.refill_budget;
    p.chunk = cfg.fetch_chunk;
    set_splits(p);
    if (p.chunk <= 0) {
        // automatic: about 32 claims per wave over the launch, 128..512 rays each
        const long long waves = grid * (block / 64);
        long long c = (long long)p.n / (waves * 32);
        c = std::min<long long>(512, std::max<long long>(128, (c + 63) / 64 * 64));
        p.chunk = (int)c;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(block), lds, stream, p);
    CW_HIP(hipGetLastError());
    return CWBVH_OK;
}

template <bool ANY, bool STATS, bool MULTI, bool PREFETCH>
int32_t dispatch_loop(CwbvhParams& p, const cwbvh_trace_config& cfg, hipStream_t st) {
    switch (cfg.loop_mode) {
        case kLoopInterleave:
            if constexpr (!STATS && !MULTI) {
                const bool w7 = !PREFETCH && (cfg.occupancy == kOccupancy7 ||
                                              (cfg.occupancy == kOccupancyAuto && (long long)p.n >= kOccupancy7MinRays));
                if (p.lds_only) {
                    if (w7) return launch_cwbvh_t<ANY, false, false, false, kLoopInterleave, kRaysFromBuffer, true, true>(p, cfg, st);
                    return launch_cwbvh_t<ANY, false, false, PREFETCH, kLoopInterleave, kRaysFromBuffer, false, true>(p, cfg, st);
                }
                if (w7) return launch_cwbvh_t<ANY, false, false, false, kLoopInterleave, kRaysFromBuffer, true>(p, cfg, st);
            }
            if constexpr (!STATS && MULTI) {
                if (p.lds_only)
                    return launch_cwbvh_t<ANY, false, true, PREFETCH, kLoopInterleave, kRaysFromBuffer, false, true>(p, cfg, st);
            }
            return launch_cwbvh_t<ANY, STATS, MULTI, PREFETCH, kLoopInterleave>(p, cfg, st);
        case kLoopSpeculate: return launch_cwbvh_t<ANY, STATS, MULTI, PREFETCH, kLoopSpeculate>(p, cfg, st);
        default: return launch_cwbvh_t<ANY, STATS, MULTI, PREFETCH, kLoopGroup>(p, cfg, st);
    }
}

template <bool ANY, bool STATS, bool MULTI>
int32_t dispatch_prefetch(CwbvhParams& p, const cwbvh_trace_config& cfg, hipStream_t st) {
    return cfg.prefetch ? dispatch_loop<ANY, STATS, MULTI, true>(p, cfg, st)
                        : dispatch_loop<ANY, STATS, MULTI, false>(p, cfg, st);
}

template <bool MULTI>
int32_t dispatch_cwbvh(CwbvhParams& p, uint32_t flags, const cwbvh_trace_config& cfg, hipStream_t st) {
    const bool any = (flags & CWBVH_TRACE_ANY) != 0, stats = (flags & CWBVH_TRACE_STATS) != 0;
    if (any) return stats ? dispatch_prefetch<true, true, MULTI>(p, cfg, st) : dispatch_prefetch<true, false, MULTI>(p, cfg, st);
    return stats ? dispatch_prefetch<false, true, MULTI>(p, cfg, st) : dispatch_prefetch<false, false, MULTI>(p, cfg, st);
}

template <bool ANY, bool STATS>
int32_t dispatch_surfels(CwbvhParams& p, const cwbvh_trace_config& cfg, hipStream_t st) {
    switch (cfg.loop_mode) {
        case kLoopInterleave:
            if constexpr (!STATS) {
                const bool w7 = cfg.occupancy == kOccupancy7 ||
                                (cfg.occupancy == kOccupancyAuto && (long long)p.n >= kOccupancy7MinRays);
                if (p.lds_only) {
                    if (w7) return launch_cwbvh_t<ANY, false, false, false, kLoopInterleave, kRaysFromSurfels, true, true>(p, cfg, st);
                    return launch_cwbvh_t<ANY, false, false, false, kLoopInterleave, kRaysFromSurfels, false, true>(p, cfg, st);
                }
                if (w7) return launch_cwbvh_t<ANY, false, false, false, kLoopInterleave, kRaysFromSurfels, true>(p, cfg, st);
            }
            return launch_cwbvh_t<ANY, STATS, false, false, kLoopInterleave, kRaysFromSurfels>(p, cfg, st);
        case kLoopSpeculate: return launch_cwbvh_t<ANY, STATS, false, false, kLoopSpeculate, kRaysFromSurfels>(p, cfg, st);
        default: return launch_cwbvh_t<ANY, STATS, false, false, kLoopGroup, kRaysFromSurfels>(p, cfg, st);
    }
}

// One thread per ray: the unfused generator (same arithmetic as the fused one).
__global__ void surfel_rays_kernel(const float4* __restrict__ pos, const float4* __restrict__ nrm,
                                   const float* __restrict__ rad, long long n_rays, int rays_per_texel, float tmin,
                                   float tmax, unsigned long long seed, Ray4* __restrict__ out) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    const long long t = r / rays_per_texel;
    const float4 p = pos[t], n = nrm[t];
    cwao::F3 o, d;
    cwao::surfel_ray(cwao::F3{p.x, p.y, p.z}, cwao::F3{n.x, n.y, n.z}, rad[t], seed, (unsigned long long)r, o, d);
    nt_store4(make_float4(o.x, o.y, o.z, tmin), &out[r].o);
    nt_store4(make_float4(d.x, d.y, d.z, tmax), &out[r].d);
}

}  // namespace

namespace cw {

int32_t validate_config(int32_t kind, const cwbvh_trace_config& c) {
    if (!(c.block_size == 64 || c.block_size == 128 || c.block_size == 256)) {
        set_error("block_size must be 64, 128 or 256");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    const int cap = kind == CWBVH_KIND_CWBVH ? kStackCapacity : kBvh2StackCapacity;
    if (c.lds_depth < 1 || c.lds_depth > cap) {
        set_error("lds_depth must be in [1, stack capacity]");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    const size_t lds = (size_t)c.block_size * c.lds_depth * (kind == CWBVH_KIND_CWBVH ? 8 : 4);
    if (lds > 160 * 1024) {
        set_error("LDS stack exceeds 160 KiB per block");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    if (c.ray_order < 0 || c.ray_order > 1) {
        set_error("ray_order must be 0 or 1");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    if (kind == CWBVH_KIND_CWBVH) {
        if (c.fetch_chunk < 0 || c.fetch_chunk > (1 << 20) || c.grid_blocks < 0 || c.refill_budget < 1 ||
            c.persistent < 0 || c.persistent > 2 ||
            c.refill_idle < 0 || c.refill_idle > 64 || c.postpone_pct < 0 || c.postpone_pct > 100 ||
            c.prefetch < 0 || c.prefetch > 1 || c.work_slices < 1 || c.work_slices > kWorkSlices ||
            c.loop_mode < kLoopGroup || c.loop_mode > kLoopQueue ||
            !(c.occupancy == kOccupancyAuto || c.occupancy == kOccupancyFree || c.occupancy == kOccupancy7)) {
            set_error("invalid CWBVH trace configuration");
            return CWBVH_ERR_INVALID_ARGUMENT;
        }
        if (c.loop_mode == kLoopQueue) {
            set_error("loop_mode 3 (triangle queue) is switched off: use 0, 1 or 2");
            return CWBVH_ERR_INVALID_ARGUMENT;
        }
        if (c.postpone_pct > 0 && c.loop_mode == kLoopInterleave) {
            set_error("postpone_pct applies to loop_mode 0 and 2");
            return CWBVH_ERR_INVALID_ARGUMENT;
        }
    } else {
        // BVH2: one ray per thread (persistent 0) or the persistent grid
        // (1, or 2 = automatic by ray count), whose claim and refill fields
        // then apply; the CWBVH-only fields must stay 0
        if (c.persistent < 0 || c.persistent > 2 || c.loop_mode != 0 || c.prefetch != 0 || c.postpone_pct != 0 ||
            c.occupancy != 0) {
            set_error("invalid BVH2 trace configuration: persistent is 0, 1 or 2; loop_mode, prefetch, "
                      "postpone_pct and occupancy are 0");
            return CWBVH_ERR_INVALID_ARGUMENT;
        }
        if (c.persistent != 0 && (c.fetch_chunk < 0 || c.fetch_chunk > (1 << 20) || c.grid_blocks < 0 ||
                                  c.refill_budget < 1 || c.refill_idle < 0 || c.refill_idle > 64 ||
                                  c.work_slices < 1 || c.work_slices > kWorkSlices)) {
            set_error("invalid BVH2 persistent configuration (refill_budget >= 1, refill_idle in [0, 64], "
                      "work_slices in [1, 8])");
            return CWBVH_ERR_INVALID_ARGUMENT;
        }
    }
    return CWBVH_OK;
}

// A tree whose stack need (known from its layout, or declared when caller
// arrays are wrapped) fits the LDS part of the stack is traversed by kernels
// without the scratch part.
static bool lds_only(const cwbvh_scene* s, const cwbvh_trace_config& c) {
    return s->stack_need > 0 && s->stack_need <= c.lds_depth && c.lds_depth <= kStackCapacity;
}

int32_t launch_cwbvh(const TraceLaunch& L) {
    const cwbvh_scene* s = L.scene;
    CwbvhParams p{};
    p.nodes = (const float4*)s->nodes;
    p.tris = (const float4*)s->woop;
    p.prim_map = s->prim_idx;
    p.rays = (const Ray4*)L.rays;
    p.order = L.order;
    p.hits = (float4*)L.hits;
    p.jobs = nullptr;
    p.work = L.counters;
    p.stats = (unsigned long long*)L.stats;
    p.n = (int)L.n;
    p.num_jobs = 0;
    p.stride = s->node_stride;
    p.lds_depth = L.cfg.lds_depth;
    p.chunk = L.cfg.fetch_chunk;
    p.idle_lanes = L.cfg.refill_idle;
    p.idle_budget = L.cfg.refill_budget;
    p.postpone_pct = L.cfg.postpone_pct;
    p.prim_id = (L.flags & CWBVH_TRACE_PRIM_ID) ? 1 : 0;
    p.slices = L.cfg.work_slices;
    // 2 = automatic: below kAutoPersistentRays the persistent grid's fixed
    // start-up and exit cost outweighs its load balancing (MI355X, AO rays on a
    // 65 k-triangle heightfield: 262 k rays trace 1.9x faster one per thread,
    // 1 M rays 15% slower)
    p.persistent = L.cfg.persistent == 2 ? (L.n >= cw::kAutoPersistentRays ? 1 : 0) : L.cfg.persistent;
    p.spec_limit = cw::speculation_limit(s->stack_need);
    p.lds_only = lds_only(s, L.cfg) ? 1 : 0;
    return dispatch_cwbvh<false>(p, L.flags, L.cfg, L.stream);
}

int32_t launch_cwbvh_surfels(const TraceLaunch& L) {
    if (L.cfg.prefetch != 0) {
        set_error("surfel traces do not use node prefetch");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    const cwbvh_scene* s = L.scene;
    CwbvhParams p{};
    p.nodes = (const float4*)s->nodes;
    p.tris = (const float4*)s->woop;
    p.prim_map = s->prim_idx;
    p.rays = nullptr;
    p.order = nullptr;
    p.hits = (float4*)L.hits;
    p.work = L.counters;
    p.stats = (unsigned long long*)L.stats;
    p.n = (int)L.n;
    p.stride = s->node_stride;
    p.lds_depth = L.cfg.lds_depth;
    p.chunk = L.cfg.fetch_chunk;
    p.idle_lanes = L.cfg.refill_idle;
    p.idle_budget = L.cfg.refill_budget;
    p.postpone_pct = L.cfg.postpone_pct;
    p.prim_id = (L.flags & CWBVH_TRACE_PRIM_ID) ? 1 : 0;
    p.slices = L.cfg.work_slices;
    p.persistent = L.cfg.persistent == 2 ? (L.n >= cw::kAutoPersistentRays ? 1 : 0) : L.cfg.persistent;
    p.spec_limit = cw::speculation_limit(s->stack_need);
    p.lds_only = lds_only(s, L.cfg) ? 1 : 0;
    p.surf_pos = (const float4*)L.surf_pos;
    p.surf_nrm = (const float4*)L.surf_nrm;
    p.surf_rad = L.surf_rad;
    p.rays_per_texel = L.rays_per_texel;
    p.gen_tmin = L.gen_tmin;
    p.gen_tmax = L.gen_tmax;
    p.seed = L.seed;
    p.texel_hits = L.texel_hits;
    const bool any = (L.flags & CWBVH_TRACE_ANY) != 0, stats = (L.flags & CWBVH_TRACE_STATS) != 0;
    if (any) return stats ? dispatch_surfels<true, true>(p, L.cfg, L.stream) : dispatch_surfels<true, false>(p, L.cfg, L.stream);
    return stats ? dispatch_surfels<false, true>(p, L.cfg, L.stream) : dispatch_surfels<false, false>(p, L.cfg, L.stream);
}

int32_t launch_surfel_rays(const cwbvh_float4* pos, const cwbvh_float4* nrm, const float* rad, int64_t num_texels,
                           int32_t rays_per_texel, float tmin, float tmax, uint64_t seed, cwbvh_ray* out,
                           hipStream_t stream) {
    const long long n = (long long)num_texels * rays_per_texel;
    if (n == 0) return CWBVH_OK;
    const int block = 256;
    hipLaunchKernelGGL(surfel_rays_kernel, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, stream,
                       (const float4*)pos, (const float4*)nrm, rad, n, rays_per_texel, tmin, tmax,
                       (unsigned long long)seed, (Ray4*)out);
    CW_HIP(hipGetLastError());
    return CWBVH_OK;
}

int32_t launch_cwbvh_batch(const BatchLaunch& L) {
    if (L.cfg.fetch_chunk <= 0) {
        set_error("batch launch needs the chunk size its job table was planned with");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    CwbvhParams p{};
    p.jobs = L.d_jobs;
    p.num_jobs = L.num_jobs;
    p.n = (int)L.total_chunks;
    p.work = L.counters;
    p.stats = (unsigned long long*)L.stats;
    p.lds_depth = L.cfg.lds_depth;
    p.chunk = L.cfg.fetch_chunk;
    p.idle_lanes = L.cfg.refill_idle;
    p.idle_budget = L.cfg.refill_budget;
    p.postpone_pct = L.cfg.postpone_pct;
    p.prim_id = (L.flags & CWBVH_TRACE_PRIM_ID) ? 1 : 0;
    p.slices = L.cfg.work_slices;
    p.persistent = 1;
    p.lds_only = L.lds_only;
    return dispatch_cwbvh<true>(p, L.flags, L.cfg, L.stream);
}

int32_t launch_bvh2(const TraceLaunch& L) {
    const cwbvh_scene* s = L.scene;
    Bvh2Params p{};
    p.nodes = (const float4*)s->nodes;
    p.tris = (const float4*)s->woop;
    p.prim_map = s->prim_idx;
    p.rays = (const Ray4*)L.rays;
    p.order = L.order;
    p.hits = (float4*)L.hits;
    p.work = L.counters;
    p.stats = (unsigned long long*)L.stats;
    p.n = (int)L.n;
    p.lds_depth = L.cfg.lds_depth;
    p.woop_last = s->woop_f4 > 0 ? (int)(s->woop_f4 - 1) : 0x7FFFFFFF;
    p.prim_id = (L.flags & CWBVH_TRACE_PRIM_ID) ? 1 : 0;
    // persistent 2 = automatic: the persistent grid from kAutoPersistentRays up
    const bool pers = L.cfg.persistent == 2 ? L.n >= cw::kAutoPersistentRays : L.cfg.persistent == 1;
    const bool any = (L.flags & CWBVH_TRACE_ANY) != 0, stats = (L.flags & CWBVH_TRACE_STATS) != 0;
    if (any) return stats ? launch_bvh2_t<true, true>(p, L.cfg, pers, L.stream) : launch_bvh2_t<true, false>(p, L.cfg, pers, L.stream);
    return stats ? launch_bvh2_t<false, true>(p, L.cfg, pers, L.stream) : launch_bvh2_t<false, false>(p, L.cfg, pers, L.stream);
}

}  // namespace cw

extern "C" int32_t cwbvh_copy_u4(void* dst, const void* src, int64_t n_u4, void* stream) {
    CW_CHECK_ARG(n_u4 >= 0 && (n_u4 == 0 || (dst && src)), "bad copy arguments");
    if (n_u4 == 0) return CWBVH_OK;
    const int block = 256;
    long long grid = (n_u4 + block - 1) / block;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(copy_u4_kernel, dim3((unsigned)grid), dim3(block), 0, (hipStream_t)stream, (uint4*)dst,
                       (const uint4*)src, (long long)n_u4);
    CW_HIP(hipGetLastError());
    return CWBVH_OK;
}

extern "C" int32_t cwbvh_rcp_check(uint64_t* d_counts, void* stream) {
    CW_CHECK_ARG(d_counts != nullptr, "d_counts is null");
    hipStream_t st = (hipStream_t)stream;
    CW_HIP(hipMemsetAsync(d_counts, 0, 3 * sizeof(uint64_t), st));
    hipLaunchKernelGGL(rcp_check_kernel, dim3(8192), dim3(256), 0, st, (unsigned long long*)d_counts);
    CW_HIP(hipGetLastError());
    return CWBVH_OK;
}

extern "C" int32_t cwbvh_valu_probe(int32_t mode, int64_t iters, int32_t blocks, float* d_out, void* stream) {
    CW_CHECK_ARG(mode >= kValuProbeFma && mode <= kValuProbeSlab, "mode must be 0, 1 or 2");
    CW_CHECK_ARG(iters >= 1 && blocks >= 1 && blocks <= (1 << 20) && d_out != nullptr, "bad probe arguments");
    const dim3 grid((unsigned)blocks), block(256);
    hipStream_t st = (hipStream_t)stream;
    if (mode == kValuProbeFma)
        hipLaunchKernelGGL(valu_probe_kernel<kValuProbeFma>, grid, block, 0, st, d_out, (long long)iters, 0.999f, 1e-7f);
    else if (mode == kValuProbeInt)
        hipLaunchKernelGGL(valu_probe_kernel<kValuProbeInt>, grid, block, 0, st, d_out, (long long)iters, 0.999f, 1e-7f);
    else
        hipLaunchKernelGGL(valu_probe_kernel<kValuProbeSlab>, grid, block, 0, st, d_out, (long long)iters, 0.999f, 1e-7f);
    CW_HIP(hipGetLastError());
    return CWBVH_OK;
}
