// Stable LSD radix sort of (key, value) pairs for gfx950: the sort behind the
// Morton ray pre-pass (cwbvh_morton_order, ray_order 1 launches, the batch
// pre-pass and the no-alloc trace).  No counterpart in the reference; written
// for wave64.
//
// 8-bit digits.  A histogram kernel reads the keys once and counts every
// pass's digits; a one-block kernel turns the counts into each pass's digit
// start offsets.  Then ONE kernel per pass ("onesweep", single-pass prefix
// with decoupled look-back):
//   1. a block takes the next tile of kTile keys from a counter (so every
//      tile with a lower index has started before it);
//   2. it ranks its keys among equal digits in input order: wave w owns
//      kItems x 64 consecutive keys (item i, lane l -> key w*kItems*64 +
//      i*64 + l, coalesced loads); per item the lanes with equal digits find
//      each other with eight ballots, the group's lowest lane adds the group
//      size to the wave's LDS count of that digit (the atomic's old value is
//      the group's start) and the lanes' ranks follow from their place in the
//      group; the per-wave counts are summed across waves in wave order;
//   3. it publishes its per-digit counts (flag AGGREGATE), then sums the
//      counts of the tiles before it, walking back until a tile that has
//      published its inclusive prefix (flag PREFIX), and publishes its own;
//   4. it writes its keys in digit order into LDS and from there to their
//      global positions, so that consecutive lanes write consecutive
//      addresses inside each digit's run.
// The order inside a tile is wave, item, lane = input order, so every pass is
// stable and the sort is an LSD sort.
//
// Status words are 64-bit: flag in bits 62..63, count in bits 0..61, one per
// (pass, tile, digit), cleared by the histogram kernel's launcher.  The
// look-back waits only on tiles that have started (step 1) and publish their
// aggregate without waiting on anyone, so it always ends; a spin limit bounds
// it regardless.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace {

constexpr int kBlock = 256;               // 4 waves
constexpr int kWaves = kBlock / 64;
constexpr int kItems = 16;                // keys per lane per tile
constexpr int kTile = kBlock * kItems;    // 4096 keys per tile
constexpr int kBins = 256;                // 8-bit digits
constexpr int kMaxPasses = 8;
static_assert(kBlock == kBins, "one thread per digit in the per-tile steps");

constexpr unsigned long long kFlagAggregate = 1ull << 62;
constexpr unsigned long long kFlagPrefix = 2ull << 62;
constexpr unsigned long long kCountMask = (1ull << 62) - 1;
constexpr long long kSpinLimit = 1ll << 26;

template <class K>
__device__ __forceinline__ unsigned digit_of(K key, int shift) {
    return (unsigned)(key >> shift) & (kBins - 1);
}

// Workspace: [hist: kMaxPasses x 256 u32][offsets: kMaxPasses x 256 u32]
// [tile counters: kMaxPasses u32, padded to 256 B][status: passes x tiles x
// 256 u64].
struct Layout {
    unsigned* hist;
    unsigned* offsets;
    unsigned* counters;
    unsigned long long* status;
};

__host__ __device__ inline size_t layout_head_bytes() { return (size_t)2 * kMaxPasses * kBins * 4 + 256; }

Layout layout_of(void* ws) {
    char* b = static_cast<char*>(ws);
    Layout L;
    L.hist = reinterpret_cast<unsigned*>(b);
    L.offsets = L.hist + kMaxPasses * kBins;
    L.counters = L.offsets + kMaxPasses * kBins;
    L.status = reinterpret_cast<unsigned long long*>(b + layout_head_bytes());
    return L;
}

__global__ void clear_u32(unsigned* __restrict__ p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long step = (long long)gridDim.x * blockDim.x;
    for (long long k = i; k < n; k += step) p[k] = 0u;
}

// Every pass's digit histogram from one read of the keys.
template <class K>
__global__ __launch_bounds__(kBlock) void radix_hist_all(const K* __restrict__ keys, long long n, int passes,
                                                          unsigned* __restrict__ hist) {
    __shared__ unsigned h[kMaxPasses][kBins];
    for (int p = 0; p < passes; ++p) h[p][threadIdx.x] = 0u;
    __syncthreads();
    const long long step = (long long)gridDim.x * kBlock;
    for (long long g = (long long)blockIdx.x * kBlock + threadIdx.x; g < n; g += step) {
        const K k = keys[g];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p][digit_of(k, 8 * p)], 1u);
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p)
        if (h[p][threadIdx.x]) atomicAdd(&hist[p * kBins + threadIdx.x], h[p][threadIdx.x]);
}

// Exclusive prefix of a 256-entry array held one entry per thread.
__device__ unsigned block_exclusive_256(unsigned v, unsigned* s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBins; off <<= 1) {
        const unsigned add = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    const unsigned incl = s[threadIdx.x];
    __syncthreads();
    return incl - v;
}

// Digit start offsets of every pass (one block).
__global__ __launch_bounds__(kBlock) void radix_offsets(const unsigned* __restrict__ hist, int passes,
                                                         unsigned* __restrict__ offsets) {
    __shared__ unsigned s[kBins];
    for (int p = 0; p < passes; ++p)
        offsets[p * kBins + threadIdx.x] = block_exclusive_256(hist[p * kBins + threadIdx.x], s);
}

__device__ __forceinline__ unsigned long long load_status(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One pass.  vals_in == nullptr: the value of key g is g.  keys_out ==
// nullptr: keys are not written (the last pass).
template <class K>
__global__ __launch_bounds__(kBlock) void radix_onesweep(const K* __restrict__ keys_in, const int* __restrict__ vals_in,
                                                          K* __restrict__ keys_out, int* __restrict__ vals_out,
                                                          long long n, int shift, unsigned* __restrict__ tile_counter,
                                                          const unsigned* __restrict__ digit_start,
                                                          unsigned long long* __restrict__ status) {
    __shared__ unsigned wcnt[kWaves][kBins];   // per-wave digit counts, then per-wave bases
    __shared__ unsigned local_start[kBins];    // start of each digit in the tile's digit order
    __shared__ unsigned global_base[kBins];    // digit start + keys of earlier tiles
    __shared__ unsigned scan_tmp[kBins];
    __shared__ unsigned tile_sh;
    __shared__ K stage_k[kTile];
    __shared__ int stage_v[kTile];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < kWaves * kBins; d += kBlock) (&wcnt[0][0])[d] = 0u;
    if (threadIdx.x == 0) tile_sh = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const unsigned tile = tile_sh;
    const long long tbase = (long long)tile * kTile;
    const long long wbase = tbase + (long long)w * (kItems * 64);
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));

    // step 2: ranks
    K key[kItems];
    unsigned grp[kItems];   // lane of the group's leader (bits 0..5) | rank in the group << 6
    unsigned old[kItems];   // leader lanes: the wave's count of the digit before this group
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const long long g = wbase + i * 64 + lane;
        const bool valid = g < n;
        key[i] = valid ? keys_in[g] : K(0);
        const unsigned d = digit_of(key[i], shift);
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long on = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? on : ~on;
        }
        const unsigned leader = valid ? (unsigned)(__ffsll((long long)m) - 1) : (unsigned)lane;
        grp[i] = leader | ((unsigned)__popcll(m & lt) << 6);
        old[i] = 0u;
        if (valid && leader == (unsigned)lane) old[i] = atomicAdd(&wcnt[w][d], (unsigned)__popcll(m));
    }
    unsigned rank[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) rank[i] = (unsigned)__shfl((int)old[i], (int)(grp[i] & 63u)) + (grp[i] >> 6);
    __syncthreads();

    // per-digit tile totals and per-wave bases (thread t = digit t)
    const unsigned d0 = threadIdx.x;
    unsigned total = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
        const unsigned c = wcnt[v][d0];
        wcnt[v][d0] = total;
        total += c;
    }
    // step 3: publish the aggregate, look back, publish the inclusive prefix
    unsigned long long* st = status + (size_t)tile * kBins + d0;
    unsigned long long excl = 0;
    if (tile == 0) {
        store_status(st, kFlagPrefix | (unsigned long long)total);
    } else {
        store_status(st, kFlagAggregate | (unsigned long long)total);
        long long p = (long long)tile - 1;
        long long spins = 0;
        while (p >= 0) {
            const unsigned long long v = load_status(status + (size_t)p * kBins + d0);
            if ((v & ~kCountMask) == 0ull) {
                if (++spins > kSpinLimit) break;   // cannot happen: tile p has started
                continue;
            }
            excl += v & kCountMask;
            if ((v & ~kCountMask) == kFlagPrefix) break;
            --p;
        }
        store_status(st, kFlagPrefix | (excl + total));
    }
    global_base[d0] = digit_start[d0] + (unsigned)excl;
    local_start[d0] = block_exclusive_256(total, scan_tmp);
    __syncthreads();

    // step 4: stage in digit order, then write out
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const long long g = wbase + i * 64 + lane;
        if (g >= n) continue;
        const unsigned d = digit_of(key[i], shift);
        const unsigned li = local_start[d] + wcnt[w][d] + rank[i];
        stage_k[li] = key[i];
        stage_v[li] = vals_in ? vals_in[g] : (int)g;
    }
    __syncthreads();
    const int count = (int)std::min<long long>(kTile, n - tbase);
    for (int idx = threadIdx.x; idx < count; idx += kBlock) {
        const K k = stage_k[idx];
        const unsigned d = digit_of(k, shift);
        const unsigned pos = global_base[d] + ((unsigned)idx - local_start[d]);
        if (keys_out) keys_out[pos] = k;
        vals_out[pos] = stage_v[idx];
    }
}

int passes_for(int key_bits) { return std::max(1, std::min(kMaxPasses, (key_bits + 7) / 8)); }

// Values move through va and vb (ping-pong) and land in dst: pass 0 reads
// vals_init (nullptr: value = input index) and writes va, later passes
// alternate vb, va, and the last pass writes dst.  So va must differ from
// vals_init, vb may be vals_init, and dst must be neither va nor vb.  Keys go
// k0 -> k1 -> k0 ... (k0 is overwritten); the last pass writes no keys.
template <class K>
int32_t sort_pairs(void* ws, size_t ws_bytes, K* k0, K* k1, const int* vals_init, int* va, int* vb, int* dst,
                   long long n, int key_bits, hipStream_t st) {
    if (key_bits > 8 * kMaxPasses || key_bits > 8 * (int)sizeof(K)) {
        cw::set_error("radix sort: too many key bits");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    if (ws_bytes < cw::radix_sort_workspace(n, key_bits)) {
        cw::set_error("radix sort: workspace too small");
        return CWBVH_ERR_INTERNAL;
    }
    const int passes = passes_for(key_bits);
    const long long tiles = (n + kTile - 1) / kTile;
    const Layout L = layout_of(ws);
    // histograms, tile counters and every pass's status words start at zero
    const long long words = (long long)(layout_head_bytes() / 4) + (long long)passes * tiles * kBins * 2;
    hipLaunchKernelGGL(clear_u32, dim3((unsigned)std::min<long long>(4096, (words + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, st, L.hist, words);
    const long long hist_blocks = std::min<long long>(2048, tiles);
    hipLaunchKernelGGL(radix_hist_all<K>, dim3((unsigned)hist_blocks), dim3(kBlock), 0, st, k0, n, passes, L.hist);
    hipLaunchKernelGGL(radix_offsets, dim3(1), dim3(kBlock), 0, st, L.hist, passes, L.offsets);
    CW_HIP(hipGetLastError());
    K* kin = k0;
    K* kout = k1;
    const int* vin = vals_init;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        int* vout = last ? dst : ((p & 1) ? vb : va);
        hipLaunchKernelGGL(radix_onesweep<K>, dim3((unsigned)tiles), dim3(kBlock), 0, st, kin, vin,
                           last ? (K*)nullptr : kout, vout, n, 8 * p, L.counters + p, L.offsets + p * kBins,
                           L.status + (size_t)p * tiles * kBins);
        CW_HIP(hipGetLastError());
        std::swap(kin, kout);
        vin = vout;
    }
    return CWBVH_OK;
}

}  // namespace

namespace cw {

size_t radix_sort_workspace(int64_t n, int key_bits) {
    const long long tiles = (n + kTile - 1) / kTile;
    return layout_head_bytes() + (size_t)passes_for(key_bits) * (size_t)tiles * kBins * 8;
}

int32_t radix_sort_u32(void* ws, size_t ws_bytes, uint32_t* k0, uint32_t* k1, const int32_t* vals_init, int32_t* va,
                       int32_t* vb, int32_t* dst, int64_t n, int key_bits, hipStream_t st) {
    if (n <= 0) return CWBVH_OK;
    return sort_pairs<unsigned>(ws, ws_bytes, k0, k1, vals_init, va, vb, dst, (long long)n, key_bits, st);
}

int32_t radix_sort_u64(void* ws, size_t ws_bytes, uint64_t* k0, uint64_t* k1, const int32_t* vals_init, int32_t* va,
                       int32_t* vb, int32_t* dst, int64_t n, int key_bits, hipStream_t st) {
    if (n <= 0) return CWBVH_OK;
    return sort_pairs<unsigned long long>(ws, ws_bytes, (unsigned long long*)k0, (unsigned long long*)k1, vals_init,
                                          va, vb, dst, (long long)n, key_bits, st);
}

}  // namespace cw
