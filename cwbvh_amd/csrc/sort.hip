// Stable LSD radix sort of (key, value) pairs for gfx950: the sort behind the
// Morton ray pre-pass (cwbvh_morton_order, ray_order 1 launches, the batch
// pre-pass).  No counterpart in the reference; written for wave64 and for a
// chip whose 8 XCDs do not share an L2.
//
// 8-bit digits, one pass per digit, each pass "reduce then scan" with no
// dependency between workgroups inside a kernel:
//   1. upsweep: one workgroup per tile of kTile keys counts the tile's digits
//      (per-wave LDS histograms) and writes one row of 256 counts per tile;
//   2. column scan, in three small kernels over that tiles x 256 table: the
//      digit sums of each chunk of kChunkTiles tiles, their exclusive scan
//      over the chunks per digit (with each digit's total), then each tile's
//      global start per digit = digit start + keys of that digit in earlier
//      tiles, written over the counts;
//   3. downsweep: the tile again.  Wave w owns kItems x 64 consecutive keys
//      (item i, lane l -> key w*kItems*64 + i*64 + l, coalesced loads); per
//      item the lanes with equal digits find each other with eight ballots,
//      the group's lowest lane adds the group size to the wave's LDS count of
//      that digit (the atomic's old value is the group's start) and the
//      lanes' ranks follow from their place in the group; per-wave counts are
//      summed across waves in wave order.  Keys and values are staged in LDS
//      in digit order and written from there, so that consecutive lanes write
//      consecutive addresses inside each digit's run.
// The order inside a tile is wave, item, lane = input order, and tiles take
// their runs in tile order, so every pass is stable and the sort is an LSD
// sort.
//
// An earlier single-pass version (decoupled look-back over per-tile status
// words) spent most of each pass waiting on look-back loads: status words
// written on one XCD reach another only through the memory side, so each
// step of a look-back costs a far load, and with ~1000 tiles in flight the
// walks were long (2^26 keys: 0.85 ms per pass).  The extra read of the keys
// here (4 B per key per pass) costs far less.
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
constexpr int kChunkTiles = 64;           // tiles per chunk of the column scan
static_assert(kBlock == kBins, "one thread per digit in the per-tile steps");

template <class K>
__device__ __forceinline__ unsigned digit_of(K key, int shift) {
    return (unsigned)(key >> shift) & (kBins - 1);
}

// Workspace: [counts / tile starts: tiles x 256 u32][chunk sums / chunk
// prefixes: chunks x 256 u32][digit totals: 256 u32], each part 256-byte
// aligned.  Every part is written before it is read in each pass, so the
// workspace needs no clearing.
struct Layout {
    unsigned* counts;
    unsigned* chunks;
    unsigned* totals;
};

inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

inline long long tiles_of(long long n) { return (n + kTile - 1) / kTile; }
inline long long chunks_of(long long tiles) { return (tiles + kChunkTiles - 1) / kChunkTiles; }

Layout layout_of(void* ws, long long tiles) {
    char* b = static_cast<char*>(ws);
    Layout L;
    L.counts = reinterpret_cast<unsigned*>(b);
    b += round256((size_t)tiles * kBins * 4);
    L.chunks = reinterpret_cast<unsigned*>(b);
    b += round256((size_t)chunks_of(tiles) * kBins * 4);
    L.totals = reinterpret_cast<unsigned*>(b);
    return L;
}

// Inclusive scan of a 256-entry array held one entry per thread; *total gets
// the sum.  Ends with a barrier, so s can be reused at once.
__device__ unsigned block_inclusive_256(unsigned v, unsigned* s, unsigned* total) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBins; off <<= 1) {
        const unsigned add = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    const unsigned incl = s[threadIdx.x];
    if (total) *total = s[kBins - 1];
    __syncthreads();
    return incl;
}

// 1. per-tile digit counts: counts[tile * 256 + digit]
template <class K>
__global__ __launch_bounds__(kBlock) void radix_upsweep(const K* __restrict__ keys, long long n, int shift,
                                                         unsigned* __restrict__ counts) {
    __shared__ unsigned h[kWaves][kBins];
    for (int d = threadIdx.x; d < kWaves * kBins; d += kBlock) (&h[0][0])[d] = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long wbase = (long long)blockIdx.x * kTile + (long long)w * (kItems * 64);
    unsigned dig[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const long long g = wbase + i * 64 + lane;
        dig[i] = g < n ? digit_of(keys[g], shift) : 0xFFFFFFFFu;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; ++i)
        if (dig[i] != 0xFFFFFFFFu) atomicAdd(&h[w][dig[i]], 1u);
    __syncthreads();
    const unsigned d = threadIdx.x;
    counts[(size_t)blockIdx.x * kBins + d] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// 2a. digit sums of each chunk of kChunkTiles tiles: chunks[c * 256 + digit]
__global__ __launch_bounds__(kBins) void radix_chunk_sums(const unsigned* __restrict__ counts, long long tiles,
                                                           unsigned* __restrict__ chunks) {
    const long long t0 = (long long)blockIdx.x * kChunkTiles;
    const int m = (int)std::min<long long>(kChunkTiles, tiles - t0);
    const unsigned* p = counts + (size_t)t0 * kBins + threadIdx.x;
    unsigned s = 0;
#pragma unroll 16
    for (int t = 0; t < m; ++t) s += p[(size_t)t * kBins];
    chunks[(size_t)blockIdx.x * kBins + threadIdx.x] = s;
}

// 2b. one workgroup per digit: exclusive scan of the chunk sums over the
// chunks (in place) and the digit's total
__global__ __launch_bounds__(kBins) void radix_chunk_scan(unsigned* __restrict__ chunks, long long nchunks,
                                                           unsigned* __restrict__ totals) {
    __shared__ unsigned s[kBins];
    const unsigned d = blockIdx.x;
    unsigned carry = 0;
    for (long long c0 = 0; c0 < nchunks; c0 += kBins) {
        const long long c = c0 + threadIdx.x;
        const unsigned v = c < nchunks ? chunks[(size_t)c * kBins + d] : 0u;
        unsigned sum = 0;
        const unsigned incl = block_inclusive_256(v, s, &sum);
        if (c < nchunks) chunks[(size_t)c * kBins + d] = carry + incl - v;
        carry += sum;
    }
    if (threadIdx.x == 0) totals[d] = carry;
}

// 2c. one workgroup per chunk: each tile's global start per digit, written
// over its counts
__global__ __launch_bounds__(kBins) void radix_tile_starts(unsigned* __restrict__ counts, long long tiles,
                                                            const unsigned* __restrict__ chunks,
                                                            const unsigned* __restrict__ totals) {
    __shared__ unsigned s[kBins];
    const unsigned d = threadIdx.x;
    const unsigned tot = totals[d];
    const unsigned start = block_inclusive_256(tot, s, nullptr) - tot;
    unsigned run = start + chunks[(size_t)blockIdx.x * kBins + d];
    const long long t0 = (long long)blockIdx.x * kChunkTiles;
    const int m = (int)std::min<long long>(kChunkTiles, tiles - t0);
    unsigned* p = counts + (size_t)t0 * kBins + d;
    for (int tb = 0; tb < m; tb += 16) {
        unsigned v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = tb + j < m ? p[(size_t)(tb + j) * kBins] : 0u;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (tb + j < m) p[(size_t)(tb + j) * kBins] = run;
            run += v[j];
        }
    }
}

// 3. one tile: rank, stage in digit order, write at the tile's digit starts.
// vals_in == nullptr: the value of key g is g.  keys_out == nullptr: keys
// are not written (the last pass).
template <class K>
__global__ __launch_bounds__(kBlock) void radix_downsweep(const K* __restrict__ keys_in, const int* __restrict__ vals_in,
                                                           K* __restrict__ keys_out, int* __restrict__ vals_out,
                                                           long long n, int shift,
                                                           const unsigned* __restrict__ tile_starts) {
    __shared__ unsigned wcnt[kWaves][kBins];   // per-wave digit counts, then per-wave bases
    __shared__ unsigned local_start[kBins];    // start of each digit in the tile's digit order
    __shared__ unsigned global_base[kBins];    // the tile's global start of each digit
    __shared__ unsigned scan_tmp[kBins];
    __shared__ K stage_k[kTile];
    __shared__ int stage_v[kTile];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < kWaves * kBins; d += kBlock) (&wcnt[0][0])[d] = 0u;
    const long long tbase = (long long)blockIdx.x * kTile;
    const long long wbase = tbase + (long long)w * (kItems * 64);
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    global_base[threadIdx.x] = tile_starts[(size_t)blockIdx.x * kBins + threadIdx.x];

    K key[kItems];
    int val[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const long long g = wbase + i * 64 + lane;
        const bool valid = g < n;
        key[i] = valid ? keys_in[g] : K(0);
        val[i] = valid ? (vals_in ? vals_in[g] : (int)g) : 0;
    }
    __syncthreads();

    // ranks among equal digits, per wave: the group's leader (its lowest
    // lane) reserves the group's places in the wave's count of the digit
    unsigned rank[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const long long g = wbase + i * 64 + lane;
        const bool valid = g < n;
        const unsigned d = digit_of(key[i], shift);
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long on = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? on : ~on;
        }
        const int leader = valid ? __ffsll((long long)m) - 1 : lane;
        unsigned old = 0u;
        if (valid && leader == lane) old = atomicAdd(&wcnt[w][d], (unsigned)__popcll(m));
        rank[i] = (unsigned)__shfl((int)old, leader) + (unsigned)__popcll(m & lt);
    }
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
    local_start[d0] = block_inclusive_256(total, scan_tmp, nullptr) - total;
    __syncthreads();   // every digit's start and every wave's base are in place

    // stage in digit order, then write out
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const long long g = wbase + i * 64 + lane;
        if (g >= n) continue;
        const unsigned d = digit_of(key[i], shift);
        const unsigned li = local_start[d] + wcnt[w][d] + rank[i];
        if (li < (unsigned)kTile) {   // always true; keeps a defect from writing outside the stage
            stage_k[li] = key[i];
            stage_v[li] = val[i];
        }
    }
    __syncthreads();
    const int count = (int)std::min<long long>(kTile, n - tbase);
    for (int idx = threadIdx.x; idx < count; idx += kBlock) {
        const K k = stage_k[idx];
        const unsigned d = digit_of(k, shift);
        const unsigned pos = global_base[d] + ((unsigned)idx - local_start[d]);
        if (pos >= (unsigned long long)n) continue;   // always false; keeps a defect inside the buffers
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
    if (n > 0x7FFFFFFFll) {
        cw::set_error("radix sort: more than 2^31 - 1 keys");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    if (ws_bytes < cw::radix_sort_workspace(n, key_bits)) {
        cw::set_error("radix sort: workspace too small");
        return CWBVH_ERR_INTERNAL;
    }
    const int passes = passes_for(key_bits);
    const long long tiles = tiles_of(n), nchunks = chunks_of(tiles);
    const Layout L = layout_of(ws, tiles);
    K* kin = k0;
    K* kout = k1;
    const int* vin = vals_init;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        int* vout = last ? dst : ((p & 1) ? vb : va);
        hipLaunchKernelGGL(radix_upsweep<K>, dim3((unsigned)tiles), dim3(kBlock), 0, st, kin, n, 8 * p, L.counts);
        hipLaunchKernelGGL(radix_chunk_sums, dim3((unsigned)nchunks), dim3(kBins), 0, st, L.counts, tiles, L.chunks);
        hipLaunchKernelGGL(radix_chunk_scan, dim3(kBins), dim3(kBins), 0, st, L.chunks, nchunks, L.totals);
        hipLaunchKernelGGL(radix_tile_starts, dim3((unsigned)nchunks), dim3(kBins), 0, st, L.counts, tiles, L.chunks,
                           L.totals);
        hipLaunchKernelGGL(radix_downsweep<K>, dim3((unsigned)tiles), dim3(kBlock), 0, st, kin, vin,
                           last ? (K*)nullptr : kout, vout, n, 8 * p, L.counts);
        CW_HIP(hipGetLastError());
        std::swap(kin, kout);
        vin = vout;
    }
    return CWBVH_OK;
}

}  // namespace

namespace cw {

size_t radix_sort_workspace(int64_t n, int key_bits) {
    (void)key_bits;
    const long long tiles = tiles_of(std::max<int64_t>(n, 1));
    return round256((size_t)tiles * kBins * 4) + round256((size_t)chunks_of(tiles) * kBins * 4) + kBins * 4;
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
