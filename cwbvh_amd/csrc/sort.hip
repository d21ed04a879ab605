// Stable LSD radix sort of (key, value) pairs for gfx950: the sort behind the
// Morton ray pre-pass (cwbvh_morton_order, ray_order 1 launches and the batch
// pre-pass).  No counterpart in the reference; written for wave64.
//
// One pass per 8-bit digit, three steps each:
//   1. histogram: every tile of kTile keys counts its digits in LDS and stores
//      the 256 counts digit-major: counts[digit * tiles + tile];
//   2. scan: an exclusive prefix sum over that array gives, for every (digit,
//      tile), the output position of the tile's first key with that digit
//      (reduce per chunk, scan of the chunk sums in one block, add back);
//   3. scatter: every tile ranks its keys among equal digits in input order
//      and writes each pair to its position.
// Ranking in step 3: wave w of a block owns kItems x 64 consecutive keys of
// the tile (item i, lane l -> key w*kItems*64 + i*64 + l, so loads stay
// coalesced).  For each item the lanes with equal digits find each other with
// eight ballots (one per digit bit); a lane's rank is the number of equal
// lanes below it plus the wave's running count of that digit in LDS, which
// the group's lowest lane then advances.  After the wave's last item the
// per-wave counts are summed across the block's waves, in wave order.  The
// order inside a tile is wave, item, lane = input order, so every pass is
// stable and the sort is an LSD sort.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace {

constexpr int kBlock = 256;               // 4 waves
constexpr int kWaves = kBlock / 64;
constexpr int kItems = 16;                // keys per lane per tile
constexpr int kTile = kBlock * kItems;    // 4096 keys per tile
constexpr int kBins = 256;                // 8-bit digits
constexpr int kScanChunk = 4096;          // counts per block in the scan steps
static_assert(kBlock == kBins, "one thread per digit in the per-tile steps");

template <class K>
__device__ __forceinline__ unsigned digit_of(K key, int shift) {
    return (unsigned)(key >> shift) & (kBins - 1);
}

// Step 1.
template <class K>
__global__ __launch_bounds__(kBlock) void radix_hist(const K* __restrict__ keys, long long n, int shift, int tiles,
                                                      unsigned* __restrict__ counts) {
    __shared__ unsigned h[kBins];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kTile;
#pragma unroll 4
    for (int i = 0; i < kItems; ++i) {
        const long long g = base + (long long)i * kBlock + threadIdx.x;
        if (g < n) atomicAdd(&h[digit_of(keys[g], shift)], 1u);
    }
    __syncthreads();
    counts[(long long)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Step 2a: sum of every kScanChunk counts.
__global__ __launch_bounds__(kBlock) void scan_reduce(const unsigned* __restrict__ c, long long len,
                                                       unsigned* __restrict__ partial) {
    __shared__ unsigned s[kBlock];
    const long long base = (long long)blockIdx.x * kScanChunk;
    unsigned acc = 0;
    for (int k = threadIdx.x; k < kScanChunk; k += kBlock) {
        const long long g = base + k;
        if (g < len) acc += c[g];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Exclusive scan of v[0..count) in place by one block (count <= kBlock * 64).
__device__ void block_exclusive_scan(unsigned* v, int count, unsigned* s) {
    const int per = (count + kBlock - 1) / kBlock;
    const int lo = threadIdx.x * per;
    unsigned sum = 0;
    for (int k = 0; k < per; ++k)
        if (lo + k < count) sum += v[lo + k];
    s[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums
    for (int off = 1; off < kBlock; off <<= 1) {
        const unsigned add = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    unsigned run = s[threadIdx.x] - sum;
    for (int k = 0; k < per; ++k)
        if (lo + k < count) {
            const unsigned x = v[lo + k];
            v[lo + k] = run;
            run += x;
        }
}

// Step 2b: exclusive scan of the chunk sums (one block).
__global__ __launch_bounds__(kBlock) void scan_partials(unsigned* __restrict__ partial, int count) {
    __shared__ unsigned s[kBlock];
    block_exclusive_scan(partial, count, s);
}

// Step 2c: exclusive scan of each chunk in place, offset by its chunk sum.
__global__ __launch_bounds__(kBlock) void scan_add(unsigned* __restrict__ c, long long len,
                                                    const unsigned* __restrict__ partial) {
    __shared__ unsigned v[kScanChunk];
    __shared__ unsigned s[kBlock];
    const long long base = (long long)blockIdx.x * kScanChunk;
    const int count = (int)std::min<long long>(kScanChunk, len - base);
    for (int k = threadIdx.x; k < count; k += kBlock) v[k] = c[base + k];
    __syncthreads();
    block_exclusive_scan(v, count, s);
    __syncthreads();
    const unsigned off = partial[blockIdx.x];
    for (int k = threadIdx.x; k < count; k += kBlock) c[base + k] = v[k] + off;
}

// Step 3.  vals_in == nullptr: the value of key g is g (the first pass of an
// index sort).  keys_out == nullptr: keys are not written (the last pass,
// when only the values are wanted).
template <class K>
__global__ __launch_bounds__(kBlock) void radix_scatter(const K* __restrict__ keys_in, const int* __restrict__ vals_in,
                                                         K* __restrict__ keys_out, int* __restrict__ vals_out,
                                                         long long n, int shift, int tiles,
                                                         const unsigned* __restrict__ offsets) {
    __shared__ unsigned wcnt[kWaves][kBins];   // running digit counts of each wave
    __shared__ unsigned tile_off[kBins];       // output position of the tile's first key per digit
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < kWaves * kBins; d += kBlock) (&wcnt[0][0])[d] = 0u;
    tile_off[threadIdx.x] = offsets[(long long)threadIdx.x * tiles + blockIdx.x];
    __syncthreads();
    const long long wbase = (long long)blockIdx.x * kTile + (long long)w * (kItems * 64);
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    K key[kItems];
    unsigned rank[kItems];
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
        unsigned r = 0;
        if (valid) {
            const unsigned prior = wcnt[w][d];
            r = prior + (unsigned)__popcll(m & lt);
            // the group's lowest lane advances the wave's count of this digit
            if ((m & lt) == 0ull) wcnt[w][d] = prior + (unsigned)__popcll(m);
        }
        rank[i] = r;
    }
    __syncthreads();
    // keys of earlier waves of the tile come first
    for (int d = threadIdx.x; d < kBins; d += kBlock) {
        unsigned run = 0;
#pragma unroll
        for (int v = 0; v < kWaves; ++v) {
            const unsigned c = wcnt[v][d];
            wcnt[v][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const long long g = wbase + i * 64 + lane;
        if (g >= n) continue;
        const unsigned d = digit_of(key[i], shift);
        const unsigned pos = tile_off[d] + wcnt[w][d] + rank[i];
        if (keys_out) keys_out[pos] = key[i];
        vals_out[pos] = vals_in ? vals_in[g] : (int)g;
    }
}

// Values move through va and vb (ping-pong) and land in dst: pass 0 reads
// vals_init (nullptr: value = input index) and writes va, later passes
// alternate vb, va, and the last pass writes dst.  So va must differ from
// vals_init, vb may be vals_init, and dst must be neither va nor vb.  Keys go
// k0 -> k1 -> k0 ... (k0 is overwritten); the last pass writes no keys.
template <class K>
int32_t sort_pairs(void* ws, size_t ws_bytes, K* k0, K* k1, const int* vals_init, int* va, int* vb, int* dst,
                   long long n, int key_bits, hipStream_t st) {
    const int tiles = (int)((n + kTile - 1) / kTile);
    const long long len = (long long)tiles * kBins;
    const int chunks = (int)((len + kScanChunk - 1) / kScanChunk);
    if (chunks > kBlock * 64) {
        cw::set_error("radix sort: too many keys for the scan");
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    if (ws_bytes < cw::radix_sort_workspace(n)) {
        cw::set_error("radix sort: workspace too small");
        return CWBVH_ERR_INTERNAL;
    }
    unsigned* counts = static_cast<unsigned*>(ws);
    unsigned* partial = counts + ((len + 63) / 64 * 64);
    const int passes = std::max(1, (key_bits + 7) / 8);
    K* kin = k0;
    K* kout = k1;
    const int* vin = vals_init;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const bool last = p == passes - 1;
        int* vout = last ? dst : ((p & 1) ? vb : va);
        hipLaunchKernelGGL(radix_hist<K>, dim3((unsigned)tiles), dim3(kBlock), 0, st, kin, n, shift, tiles, counts);
        hipLaunchKernelGGL(scan_reduce, dim3((unsigned)chunks), dim3(kBlock), 0, st, counts, len, partial);
        hipLaunchKernelGGL(scan_partials, dim3(1), dim3(kBlock), 0, st, partial, chunks);
        hipLaunchKernelGGL(scan_add, dim3((unsigned)chunks), dim3(kBlock), 0, st, counts, len, partial);
        hipLaunchKernelGGL(radix_scatter<K>, dim3((unsigned)tiles), dim3(kBlock), 0, st, kin, vin,
                           last ? (K*)nullptr : kout, vout, n, shift, tiles, counts);
        CW_HIP(hipGetLastError());
        std::swap(kin, kout);
        vin = vout;
    }
    return CWBVH_OK;
}

}  // namespace

namespace cw {

size_t radix_sort_workspace(int64_t n) {
    const long long tiles = (n + kTile - 1) / kTile;
    const long long len = tiles * kBins;
    const long long chunks = (len + kScanChunk - 1) / kScanChunk;
    return (size_t)(((len + 63) / 64 * 64) + chunks + 64) * sizeof(unsigned);
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
