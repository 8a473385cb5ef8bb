// sort.hpp -- LDS-staged stable LSD radix sort of (key, value) uint2 pairs and an exclusive
// uint32 scan, for gfx950 (wave64, 256-thread blocks = 4 waves).
//
// Radix sort (replaces the SDK RadixSort the reference links, dempacking.cu:553-555): 8-bit
// digits, one pass per 8 key bits actually used.  Per pass:
//   hist    : each 2048-key tile counts its digits in LDS -> counts[digit * ntiles + tile]
//   scan    : exclusive scan of counts (digit-major) -> global offset of (digit, tile)
//   scatter : each tile ranks its keys stably (wave ballots give the rank among equal digits
//             inside a wave; LDS per-wave counts extend it across waves and rounds), stages
//             keys in LDS in digit order, then writes each digit run to its global offset,
//             so consecutive lanes write consecutive addresses.
// Stability: keys are ranked in input order (round-major, lane-minor), so equal keys keep
// their input order -- the result equals std::stable_sort by key.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dem {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;   // 2048 keys per tile
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Inclusive wave-level scan of a uint32 across 64 lanes.
__device__ __forceinline__ unsigned wave_incl_scan(unsigned v)
{
    unsigned lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned t = __shfl_up(v, d, 64);
        if (lane >= (unsigned)d) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the exclusive
// prefix and writes the block total to *total.  `wsum` is 4 words of LDS.
__device__ __forceinline__ unsigned block_excl_scan(unsigned v, unsigned *wsum, unsigned *total)
{
    unsigned lane = lane_id();
    unsigned w = threadIdx.x >> 6;
    unsigned inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    unsigned base = 0, tot = 0;
#pragma unroll
    for (unsigned k = 0; k < (kScanThreads >> 6); k++) {
        unsigned s = wsum[k];
        if (k < w) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

// ---------------------------------------------------------------------------------------
// Exclusive scan, 3 kernels.  out[i] = sum(in[0..i)), out[n] = total (out has n+1 slots),
// and *total_out (device) = total when non-null.  `active` == nullptr or *active != 0 runs.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const unsigned *in, unsigned n,
                                                             unsigned *tile_sums,
                                                             const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned wsum[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    unsigned s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        if (i < n) s += in[i];
    }
    unsigned tot;
    block_excl_scan(s, wsum, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of tile sums in place; total to *total_out.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(unsigned *tile_sums, unsigned ntiles,
                                                           unsigned *total_out,
                                                           const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned wsum[kScanThreads / 64];
    unsigned carry = 0;
    for (unsigned base = 0; base < ntiles; base += kScanThreads) {
        unsigned i = base + threadIdx.x;
        unsigned v = i < ntiles ? tile_sums[i] : 0u;
        unsigned tot;
        unsigned ex = block_excl_scan(v, wsum, &tot);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const unsigned *in, unsigned n,
                                                             const unsigned *tile_offs,
                                                             unsigned *out,
                                                             const unsigned *total,
                                                             const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned wsum[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    unsigned v[kScanItems];
    unsigned s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    unsigned tot;
    unsigned ex = block_excl_scan(s, wsum, &tot) + tile_offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && total) out[n] = *total;
}

// ---------------------------------------------------------------------------------------
// The same exclusive scan of two arrays of n words at once (the own-pair and transposed
// counts of a rebuild): out_k[i] = sum(in_k[0..i)), out_k[n] = total_k, total_out[k] =
// total_k.  Tile sums: 2 per tile (ts[2 * tile + k]).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kScanThreads) void k_scan2_tiles(const unsigned *in0,
                                                              const unsigned *in1, unsigned n,
                                                              unsigned *ts, const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned wsum[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    unsigned s0 = 0, s1 = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        if (i < n) {
            s0 += in0[i];
            s1 += in1[i];
        }
    }
    unsigned t0, t1;
    block_excl_scan(s0, wsum, &t0);
    block_excl_scan(s1, wsum, &t1);
    if (threadIdx.x == 0) {
        ts[2 * blockIdx.x] = t0;
        ts[2 * blockIdx.x + 1] = t1;
    }
}

// Single block: exclusive scans of both tile-sum sequences in place; totals to total_out[0..1].
__global__ __launch_bounds__(kScanThreads) void k_scan2_top(unsigned *ts, unsigned ntiles,
                                                            unsigned *total_out,
                                                            const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned wsum[kScanThreads / 64];
    unsigned c0 = 0, c1 = 0;
    for (unsigned base = 0; base < ntiles; base += kScanThreads) {
        unsigned i = base + threadIdx.x;
        unsigned v0 = i < ntiles ? ts[2 * i] : 0u, v1 = i < ntiles ? ts[2 * i + 1] : 0u;
        unsigned t0, t1;
        unsigned e0 = block_excl_scan(v0, wsum, &t0);
        unsigned e1 = block_excl_scan(v1, wsum, &t1);
        if (i < ntiles) {
            ts[2 * i] = c0 + e0;
            ts[2 * i + 1] = c1 + e1;
        }
        c0 += t0;
        c1 += t1;
    }
    if (threadIdx.x == 0) {
        total_out[0] = c0;
        total_out[1] = c1;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan2_apply(const unsigned *in0,
                                                              const unsigned *in1, unsigned n,
                                                              const unsigned *ts, unsigned *out0,
                                                              unsigned *out1,
                                                              const unsigned *total,
                                                              const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned wsum[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    unsigned v0[kScanItems], v1[kScanItems];
    unsigned s0 = 0, s1 = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        v0[k] = i < n ? in0[i] : 0u;
        v1[k] = i < n ? in1[i] : 0u;
        s0 += v0[k];
        s1 += v1[k];
    }
    unsigned t0, t1;
    unsigned e0 = block_excl_scan(s0, wsum, &t0) + ts[2 * blockIdx.x];
    unsigned e1 = block_excl_scan(s1, wsum, &t1) + ts[2 * blockIdx.x + 1];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        if (i < n) {
            out0[i] = e0;
            out1[i] = e1;
        }
        e0 += v0[k];
        e1 += v1[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out0[n] = total[0];
        out1[n] = total[1];
    }
}

// ---------------------------------------------------------------------------------------
// Radix sort kernels on uint2 (key, value).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint2 *kv, unsigned n,
                                                             int shift, unsigned ntiles,
                                                             unsigned *counts,
                                                             const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned hist[256];
    hist[threadIdx.x] = 0u;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        size_t i = base + (size_t)r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&hist[(kv[i].x >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint2 *in, unsigned n,
                                                                int shift, unsigned ntiles,
                                                                const unsigned *offs,
                                                                uint2 *out,
                                                                const unsigned *active)
{
    if (active && *active == 0u) return;
    __shared__ unsigned s_key[kSortTile];
    __shared__ unsigned s_val[kSortTile];
    __shared__ unsigned s_hist[256];      // tile histogram, then exclusive tile prefix
    __shared__ unsigned s_run[256];       // keys of each digit already ranked (earlier rounds)
    __shared__ unsigned s_wcnt[4][256];   // per-wave digit counts of the current round
    __shared__ unsigned s_wsum[4];

    const unsigned t = threadIdx.x;
    const unsigned lane = lane_id();
    const unsigned w = t >> 6;
    const size_t base = (size_t)blockIdx.x * kSortTile;

    uint2 item[kSortItems];
    s_hist[t] = 0u;
    s_run[t] = 0u;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        size_t i = base + (size_t)r * kSortThreads + t;
        if (i < n) {
            item[r] = in[i];
            atomicAdd(&s_hist[(item[r].x >> shift) & 0xffu], 1u);
        } else {
            item[r] = make_uint2(0u, 0u);
        }
    }
    __syncthreads();
    unsigned tile_total;
    unsigned my_prefix = block_excl_scan(s_hist[t], s_wsum, &tile_total);
    __syncthreads();
    s_hist[t] = my_prefix;   // exclusive prefix of digit t inside this tile
    __syncthreads();

    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        size_t i = base + (size_t)r * kSortThreads + t;
        bool valid = i < n;
        unsigned d = (item[r].x >> shift) & 0xffu;
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        unsigned rank_w = __popcll(peers & lt_mask);
        s_wcnt[0][t] = 0u;
        s_wcnt[1][t] = 0u;
        s_wcnt[2][t] = 0u;
        s_wcnt[3][t] = 0u;
        __syncthreads();
        // the highest lane of each peer group publishes the group's size
        if (valid && (peers >> lane) == 1ull) s_wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned before = s_run[d];
            for (unsigned k = 0; k < w; k++) before += s_wcnt[k][d];
            unsigned pos = s_hist[d] + before + rank_w;
            s_key[pos] = item[r].x;
            s_val[pos] = item[r].y;
        }
        __syncthreads();
        s_run[t] += s_wcnt[0][t] + s_wcnt[1][t] + s_wcnt[2][t] + s_wcnt[3][t];
        __syncthreads();
    }
    // keys are in LDS in (digit, input order); write each digit run to its global offset
    for (unsigned k = t; k < tile_total; k += kSortThreads) {
        unsigned key = s_key[k];
        unsigned d = (key >> shift) & 0xffu;
        unsigned g = offs[(size_t)d * ntiles + blockIdx.x] + (k - s_hist[d]);
        out[g] = make_uint2(key, s_val[k]);
    }
}

}  // namespace dem
