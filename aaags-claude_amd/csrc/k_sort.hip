// k_sort.hip -- tile binning and the stable LSD radix sort, written for wave64.
//
// scan:       reduce-then-scan over 4096-element blocks (three launches); the
//             wave-level prefix uses lane shuffles, the block level LDS.
// duplicate:  one thread per Gaussian writes its tile|depth keys at the
//             scanned offset (row-major over its tile rectangle).
// radix sort: per pass an upsweep (per-block digit histogram), a scan of each
//             digit's row of the digit-major histogram table (one wave per
//             digit), and a downsweep that adds the digit totals' scan, ranks keys
//             stably with wave ballots (match-any on the digit bits), stages
//             the block's keys in LDS in digit order and writes them out in
//             runs, so global writes are contiguous within each digit bucket.
//             1024 or 4096 keys per 256-thread block (by sort size).  Sorts
//             of 4M keys and more instead count every pass's digits in one
//             read up front and run each pass as one kernel (decoupled
//             look-back in place of the upsweep and the scan).
#include <algorithm>

#include "gs_device.h"
#include "gs_internal.h"
#include "gs_wave.h"

namespace gsk {
using namespace gsw;

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_ITEMS;  // 4096

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
// Keys per block = 256 threads x ROUNDS: 4096 for large sorts (long output
// runs per digit), 1024 below 4M keys so that a sort of ~1M keys still spreads
// over ~1000 blocks instead of ~250.
constexpr uint32_t RS_SMALL_LIMIT = 1u << 22;

// n_dev (optional): the item count lives on the device (graph capture); the
// launch is sized for n (the capacity) and the kernels use min(*n_dev, n).
__device__ __forceinline__ uint32_t live_count(uint32_t n, const uint32_t* n_dev) {
    if (n_dev) {
        uint32_t d = *n_dev;
        return d < n ? d : n;
    }
    return n;
}

// ---------------------------------------------------------------- scan ----
// A thread's SCAN_ITEMS consecutive 32-bit items: four 16-byte loads where the
// run is whole and the array 16-byte aligned (four loads touch 64 lines per
// wave instead of sixteen), else one item at a time.
template <typename TIn>
__device__ __forceinline__ void load_items(const TIn* __restrict__ in, uint32_t base, uint32_t n,
                                           uint32_t (&v)[SCAN_ITEMS]) {
    static_assert(sizeof(TIn) == 4 && SCAN_ITEMS % 4 == 0, "32-bit items, whole 16-byte runs");
    if (base + SCAN_ITEMS <= n && ((uintptr_t)in & 15u) == 0) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS / 4; q++) {
            const uint4 x = p[q];
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) v[k] = (base + k < n) ? (uint32_t)in[base + k] : 0u;
    }
}

// Phase 1: per-block sums of `in` (int32 or uint32 values treated as uint64).
template <typename TIn>
__global__ __launch_bounds__(SCAN_THREADS) void scan_reduce_kernel(const TIn* __restrict__ in, uint32_t n,
                                                                   unsigned long long* __restrict__ partial) {
    __shared__ unsigned long long s_w[SCAN_THREADS / 64];
    uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    load_items(in, base, n, v);
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s += (unsigned long long)v[k];
    s = wave_sum(s);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < SCAN_THREADS / 64; k++) t += s_w[k];
        partial[blockIdx.x] = t;
    }
}

// Phase 2: one block turns the partial sums into exclusive block offsets and
// writes the grand total (uint32 + overflow flag).
__global__ __launch_bounds__(1024) void scan_partials_kernel(unsigned long long* __restrict__ partial, uint32_t nb,
                                                             uint32_t* __restrict__ total32,
                                                             unsigned long long* __restrict__ total64) {
    __shared__ unsigned long long s_w[16];
    uint32_t per = (nb + 1023) / 1024;
    uint32_t b0 = threadIdx.x * per;
    unsigned long long s = 0;
    for (uint32_t k = 0; k < per; k++)
        if (b0 + k < nb) s += partial[b0 + k];
    int w = threadIdx.x >> 6, l = lane_id();
    unsigned long long inc = wave_inclusive_scan(s);
    if (l == 63) s_w[w] = inc;
    __syncthreads();
    unsigned long long wb = 0, tot = 0;
    for (int k = 0; k < 16; k++) {
        if (k < w) wb += s_w[k];
        tot += s_w[k];
    }
    unsigned long long run = wb + inc - s;
    for (uint32_t k = 0; k < per; k++)
        if (b0 + k < nb) {
            unsigned long long v = partial[b0 + k];
            partial[b0 + k] = run;
            run += v;
        }
    if (threadIdx.x == 0) {
        if (total32) *total32 = (uint32_t)tot;
        if (total64) *total64 = tot;
    }
}

// Phase 3: block-local scan plus the block offset.  INCLUSIVE selects the
// inclusive (binning offsets) or exclusive (radix histogram) variant.
template <typename TIn, bool INCLUSIVE>
__global__ __launch_bounds__(SCAN_THREADS) void scan_downsweep_kernel(const TIn* __restrict__ in, uint32_t n,
                                                                      const unsigned long long* __restrict__ partial,
                                                                      uint32_t* __restrict__ out) {
    __shared__ unsigned long long s_w[SCAN_THREADS / 64];
    uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    load_items(in, base, n, v);
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s += v[k];
    unsigned long long tot;
    unsigned long long run = block_exclusive_scan<unsigned long long>(s, s_w, &tot) + partial[blockIdx.x];
    uint32_t o[SCAN_ITEMS];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (INCLUSIVE) run += v[k];
        o[k] = (uint32_t)run;
        if (!INCLUSIVE) run += v[k];
    }
    if (base + SCAN_ITEMS <= n && ((uintptr_t)out & 15u) == 0) {
        uint4* q = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS / 4; j++) q[j] = make_uint4(o[4 * j], o[4 * j + 1], o[4 * j + 2], o[4 * j + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++)
            if (base + k < n) out[base + k] = o[k];
    }
}

// ----------------------------------------------------------- duplicate ----
// tile_binner_cpu.cpp:59-99: keys tile_id << 32 | float bits of the depth.
__global__ __launch_bounds__(256) void duplicate_kernel(int n, const float2* __restrict__ means2D,
                                                        const float* __restrict__ depths, const int* __restrict__ radii,
                                                        const uint32_t* __restrict__ offsets_incl, int gx, int gy,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ values,
                                                        uint32_t capacity, uint32_t* __restrict__ overflow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = radii[i];
    if (r <= 0) return;
    float2 m = means2D[i];
    int mn[2], mx[2];
    gsd::get_rect(m.x, m.y, r, gx, gy, mn, mx);
    uint32_t cnt = (uint32_t)((mx[1] - mn[1]) * (mx[0] - mn[0]));
    uint32_t off = offsets_incl[i] - cnt;
    uint64_t db = (uint64_t)__float_as_uint(depths[i]);
    for (int y = mn[1]; y < mx[1]; y++)
        for (int x = mn[0]; x < mx[0]; x++) {
            if (off < capacity) {
                keys[off] = ((uint64_t)(uint32_t)(y * gx + x) << 32) | db;
                values[off] = (uint32_t)i;
            } else if (overflow) {
                *overflow = 1u;
            }
            off++;
        }
}

// ------------------------------------------------- depth-ordered binning ----
// Stable-sorting pairs by (tile << 32 | depth bits) orders each tile's pairs
// by (depth bits, Gaussian index).  The same order results from sorting the
// visible Gaussians by depth bits (stable, so ties keep index order), emitting
// each Gaussian's pairs in that order, and stable-sorting the pairs by tile
// alone.  These kernels implement the first two steps; the pairs are 64-bit
// words tile << 32 | index, so the tile sort moves 8 bytes per pair.

// Depth keys: culled Gaussians get the largest key and emit no pairs.
__global__ __launch_bounds__(256) void depth_keys_kernel(int n, const float* __restrict__ depths,
                                                         const int* __restrict__ radii, uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = radii[i] > 0 ? (uint64_t)__float_as_uint(depths[i]) : 0xFFFFFFFFull;
    idx[i] = (uint32_t)i;
}

// The same keys with the index in the high word (index << 32 | key): sorted
// on bits [0, 32) as self-contained words, the index travels with its key
// (no separate value array through the passes).
// rects (optional, views of at most 255 x 255 tiles): each Gaussian's tile
// rectangle packed as x0 | x1 << 8 | y0 << 16 | y1 << 24 (0 when culled), in
// index order, so that the binning gathers one word per Gaussian in depth
// order instead of a mean and a radius.
__global__ __launch_bounds__(256) void depth_words_kernel(int n, const float* __restrict__ depths,
                                                          const int* __restrict__ radii, uint64_t* __restrict__ words,
                                                          const float2* __restrict__ means2D, int gx, int gy,
                                                          uint32_t* __restrict__ rects) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int rad = radii[i];
    const uint32_t key = rad > 0 ? __float_as_uint(depths[i]) : 0xFFFFFFFFu;
    words[i] = ((uint64_t)(uint32_t)i << 32) | key;
    if (rects) {
        uint32_t rec = 0u;
        if (rad > 0) {
            const float2 m = means2D[i];
            int mn[2], mx[2];
            gsd::get_rect(m.x, m.y, rad, gx, gy, mn, mx);
            rec = (uint32_t)mn[0] | ((uint32_t)mx[0] << 8) | ((uint32_t)mn[1] << 16) | ((uint32_t)mx[1] << 24);
        }
        rects[i] = rec;
    }
}

__global__ __launch_bounds__(256) void gather_tiles_kernel(int n, const uint32_t* __restrict__ order,
                                                           const int* __restrict__ tiles, int* __restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    out[r] = tiles[order[r]];
}

// Pairs of the Gaussian of depth rank r at [offsets_incl[r] - count, offsets_incl[r]).
__global__ __launch_bounds__(256) void duplicate_ranked_kernel(int n, const uint32_t* __restrict__ order,
                                                               const float2* __restrict__ means2D,
                                                               const int* __restrict__ radii,
                                                               const uint32_t* __restrict__ offsets_incl, int gx,
                                                               int gy, uint64_t* __restrict__ words,
                                                               uint32_t capacity, uint32_t* __restrict__ overflow) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t i = order[r];
    int rad = radii[i];
    if (rad <= 0) return;
    float2 m = means2D[i];
    int mn[2], mx[2];
    gsd::get_rect(m.x, m.y, rad, gx, gy, mn, mx);
    uint32_t cnt = (uint32_t)((mx[1] - mn[1]) * (mx[0] - mn[0]));
    uint32_t off = offsets_incl[r] - cnt;
    for (int y = mn[1]; y < mx[1]; y++)
        for (int x = mn[0]; x < mx[0]; x++) {
            if (off < capacity) words[off] = ((uint64_t)(uint32_t)(y * gx + x) << 32) | i;
            else if (overflow) *overflow = 1u;
            off++;
        }
}

// Multi-view: the depth sort runs over keys view << 32 | depth bits, so the
// sorted order is (view, depth, index).  Pair words carry the view too:
// view << 48 | local tile << 32 | index; the tile sort keys on the local tile
// bits only, which leaves each (view, tile) run contiguous and in depth order.
__global__ __launch_bounds__(256) void gather_tiles_multi_kernel(uint32_t total, int N,
                                                                 const uint64_t* __restrict__ skeys,
                                                                 const uint32_t* __restrict__ sidx,
                                                                 const int* __restrict__ tiles, int* __restrict__ out) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total) return;
    uint32_t v = (uint32_t)(skeys[r] >> 32);
    out[r] = tiles[(size_t)v * N + sidx[r]];
}

__global__ __launch_bounds__(256) void duplicate_multi_kernel(uint32_t total, int N, const uint64_t* __restrict__ skeys,
                                                              const uint32_t* __restrict__ sidx,
                                                              const float2* __restrict__ means2D,
                                                              const int* __restrict__ radii,
                                                              const uint32_t* __restrict__ offsets_incl,
                                                              const int4* __restrict__ views,
                                                              uint64_t* __restrict__ words, uint32_t capacity,
                                                              uint32_t* __restrict__ overflow) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total) return;
    uint32_t v = (uint32_t)(skeys[r] >> 32);
    uint32_t i = sidx[r];
    size_t o = (size_t)v * N + i;
    int rad = radii[o];
    if (rad <= 0) return;
    int4 vi = views[v];
    int gx = vi.z, gy = (vi.y + 15) / 16;
    float2 m = means2D[o];
    int mn[2], mx[2];
    gsd::get_rect(m.x, m.y, rad, gx, gy, mn, mx);
    uint32_t cnt = (uint32_t)((mx[1] - mn[1]) * (mx[0] - mn[0]));
    uint32_t off = offsets_incl[r] - cnt;
    const uint64_t hi = (uint64_t)v << 48;
    for (int y = mn[1]; y < mx[1]; y++)
        for (int x = mn[0]; x < mx[0]; x++) {
            if (off < capacity) words[off] = hi | ((uint64_t)(uint32_t)(y * gx + x) << 32) | i;
            else if (overflow) *overflow = 1u;
            off++;
        }
}

// Ranges per global tile (first global tile of the view + local tile).
__global__ __launch_bounds__(256) void tile_ranges_multi_kernel(const uint64_t* __restrict__ words, uint32_t n_cap,
                                                                const uint32_t* __restrict__ n_dev,
                                                                const int4* __restrict__ views,
                                                                uint32_t* __restrict__ ranges) {
    const uint32_t n = live_count(n_cap, n_dev);
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint64_t w = words[j];
    uint32_t g = (uint32_t)views[w >> 48].w + (uint32_t)((w >> 32) & 0xFFFFu);
    if (j == 0) ranges[g * 2] = 0;
    else {
        uint64_t wp = words[j - 1];
        uint32_t gp = (uint32_t)views[wp >> 48].w + (uint32_t)((wp >> 32) & 0xFFFFu);
        if (g != gp) {
            ranges[gp * 2 + 1] = j;
            ranges[g * 2] = j;
        }
    }
    if (j == n - 1) ranges[g * 2 + 1] = n;
}

// ---------------------------------------------------------- radix sort ----
template <int BITS, int ROUNDS>
__global__ __launch_bounds__(RS_THREADS) void rs_upsweep_kernel(const uint64_t* __restrict__ keys, uint32_t n_cap,
                                                                const uint32_t* __restrict__ n_dev, int shift,
                                                                uint32_t hs, uint32_t* __restrict__ hist) {
    constexpr int RADIX = 1 << BITS;
    constexpr int RS_TILE = RS_THREADS * ROUNDS, RS_WAVE_SPAN = ROUNDS * 64;
    const uint32_t n = live_count(n_cap, n_dev);
    __shared__ uint32_t s_h[RS_WAVES][RADIX];
    const int w = threadIdx.x >> 6, l = lane_id();
    for (int i = threadIdx.x; i < RS_WAVES * RADIX; i += RS_THREADS) (&s_h[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t wbase = blockIdx.x * RS_TILE + w * RS_WAVE_SPAN;
    // every round's key is loaded before any is counted: one memory latency
    // per wave instead of one per round
    uint64_t k_r[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        k_r[r] = i < n ? keys[i] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        uint32_t i = wbase + r * 64 + l;
        bool valid = i < n;
        uint64_t active = __ballot(valid);
        if (active == 0) break;
        uint32_t d = valid ? digit_of<BITS>(k_r[r], shift) : 0u;
        uint64_t peers = match_digit<BITS>(d, active);
        if (valid && (peers & lanemask_lt()) == 0) s_h[w][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += RS_THREADS) {
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < RS_WAVES; k++) s += s_h[k][d];
        hist[(size_t)d * hs + blockIdx.x] = s;
    }
}

// One wave per digit: exclusive scan of the digit's row of the digit-major
// count table (its count in every block, in block order; rows hs >= nb
// entries apart, hs a multiple of 4) and the digit's total.  Each lane takes
// four consecutive counts as one 16-byte load, so one wave-wide scan covers
// 256 blocks.  The downsweep adds the exclusive scan of the digit totals, so
// a pass needs no separate three-launch scan of the whole table.
__global__ __launch_bounds__(64) void rs_digit_scan_kernel(const uint32_t* __restrict__ hist, uint32_t nb,
                                                           uint32_t hs, uint32_t* __restrict__ goff,
                                                           uint32_t* __restrict__ dtot) {
    const uint4* h = reinterpret_cast<const uint4*>(hist + (size_t)blockIdx.x * hs);
    uint4* o = reinterpret_cast<uint4*>(goff + (size_t)blockIdx.x * hs);
    const uint32_t nq = hs / 4u;
    const int l = lane_id();
    uint32_t carry = 0u;
    // DS_LOADS steps of 64 quads are loaded before they are scanned, so the
    // loads of a row (1024 blocks per group of steps) overlap
    constexpr int DS_LOADS = 4;
    for (uint32_t q00 = 0; q00 < nq; q00 += 64u * DS_LOADS) {
        uint4 xs[DS_LOADS];
#pragma unroll
        for (int j = 0; j < DS_LOADS; j++) {
            const uint32_t q = q00 + 64u * j + (uint32_t)l;
            xs[j] = q < nq ? h[q] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int j = 0; j < DS_LOADS; j++) {
            const uint32_t q = q00 + 64u * j + (uint32_t)l;
            uint4 x = xs[j];
            const uint32_t i0 = 4u * q;  // counts past the last block are padding
            if (i0 >= nb) x.x = 0u;
            if (i0 + 1u >= nb) x.y = 0u;
            if (i0 + 2u >= nb) x.z = 0u;
            if (i0 + 3u >= nb) x.w = 0u;
            const uint32_t tot = x.x + x.y + x.z + x.w;
            const uint32_t inc = wave_inclusive_scan(tot);
            const uint32_t ex = carry + inc - tot;
            if (q < nq) o[q] = make_uint4(ex, ex + x.x, ex + x.x + x.y, ex + x.x + x.y + x.z);
            carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        }
    }
    if (l == 0) dtot[blockIdx.x] = carry;
}

// VALS = false sorts self-contained 64-bit words (no separate value array).
template <int BITS, bool VALS, int ROUNDS>
__global__ __launch_bounds__(RS_THREADS) void rs_downsweep_kernel(const uint64_t* __restrict__ keys_in,
                                                                  const uint32_t* __restrict__ vals_in, uint32_t n_cap,
                                                                  const uint32_t* __restrict__ n_dev, int shift,
                                                                  uint32_t hs, const uint32_t* __restrict__ goff,
                                                                  const uint32_t* __restrict__ dtot,
                                                                  uint64_t* __restrict__ keys_out,
                                                                  uint32_t* __restrict__ vals_out) {
    constexpr int RADIX = 1 << BITS;
    constexpr int RS_TILE = RS_THREADS * ROUNDS, RS_WAVE_SPAN = ROUNDS * 64;
    const uint32_t n = live_count(n_cap, n_dev);
    if (blockIdx.x * (uint32_t)RS_TILE >= n) return;  // whole block past the live count
    __shared__ uint64_t s_k[RS_TILE];
    __shared__ uint32_t s_v[VALS ? RS_TILE : 1];
    __shared__ uint32_t s_wh[RS_WAVES][RADIX];
    __shared__ uint32_t s_dstart[RADIX];
    __shared__ uint32_t s_goff[RADIX];
    __shared__ uint32_t s_scan[RS_WAVES];

    const int w = threadIdx.x >> 6, l = lane_id();
    for (int i = threadIdx.x; i < RS_WAVES * RADIX; i += RS_THREADS) (&s_wh[0][0])[i] = 0u;
    {
        // the digit's start in the output: all smaller digits, then this
        // digit's keys in the blocks before this one
        static_assert(RADIX <= RS_THREADS, "one digit per thread");
        const uint32_t t = threadIdx.x < RADIX ? dtot[threadIdx.x] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<uint32_t>(t, s_scan, &total);
        if (threadIdx.x < RADIX) s_goff[threadIdx.x] = ex + goff[(size_t)threadIdx.x * hs + blockIdx.x];
    }
    __syncthreads();

    const uint32_t bbase = blockIdx.x * RS_TILE;
    const uint32_t wbase = bbase + w * RS_WAVE_SPAN;
    uint64_t k_r[ROUNDS];
    uint32_t v_r[VALS ? ROUNDS : 1];
    uint32_t rank_r[ROUNDS];
    uint32_t d_r[ROUNDS];
    // every round's key (and value) is loaded before the ranking starts
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        k_r[r] = i < n ? keys_in[i] : 0ull;
        if constexpr (VALS) v_r[r] = i < n ? vals_in[i] : 0u;
    }
    // 1. per-wave stable ranks: rounds in key order, lanes in key order.
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        uint32_t i = wbase + r * 64 + l;
        bool valid = i < n;
        uint64_t active = __ballot(valid);
        uint32_t d = digit_of<BITS>(k_r[r], shift);
        d_r[r] = d;
        rank_r[r] = 0u;
        if (active != 0) {
            uint64_t peers = match_digit<BITS>(d, active);
            uint32_t before = 0u;
            if (valid) before = s_wh[w][d];
            // every lane has read the counter before the leader bumps it
            __builtin_amdgcn_wave_barrier();
            if (valid && (peers & lanemask_lt()) == 0) s_wh[w][d] = before + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            rank_r[r] = before + (uint32_t)__popcll(peers & lanemask_lt());
        }
    }
    __syncthreads();
    // 2. digit totals, per-wave offsets inside each digit, digit starts.
    for (int d = threadIdx.x; d < RADIX; d += RS_THREADS) {
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < RS_WAVES; k++) {
            uint32_t c = s_wh[k][d];
            s_wh[k][d] = run;
            run += c;
        }
        s_dstart[d] = run;  // digit total for now
    }
    __syncthreads();
    {
        uint32_t tot = (threadIdx.x < RADIX) ? s_dstart[threadIdx.x] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan<uint32_t>(tot, s_scan, &total);
        if (threadIdx.x < RADIX) s_dstart[threadIdx.x] = ex;
    }
    __syncthreads();
    // 3. stage the block's pairs in LDS in (digit, input order) order.
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        uint32_t i = wbase + r * 64 + l;
        if (i < n) {
            uint32_t p = s_dstart[d_r[r]] + s_wh[w][d_r[r]] + rank_r[r];
            if (p < (uint32_t)RS_TILE) {
                s_k[p] = k_r[r];
                if constexpr (VALS) s_v[p] = v_r[r];
            }
        }
    }
    __syncthreads();
    // 4. write out: consecutive threads write consecutive slots of a bucket.
    uint32_t cnt = (n - bbase < (uint32_t)RS_TILE) ? (n - bbase) : (uint32_t)RS_TILE;
    for (uint32_t p = threadIdx.x; p < cnt; p += RS_THREADS) {
        uint64_t k = s_k[p];
        uint32_t d = digit_of<BITS>(k, shift);
        uint32_t g = s_goff[d] + (p - s_dstart[d]);
        if (g < n) {
            keys_out[g] = k;
            if constexpr (VALS) vals_out[g] = s_v[p];
        }
    }
}

// ------------------------------------------------- one-sweep passes ----
// Sorts of at least RS_SMALL_LIMIT keys: one kernel counts the digits of every
// pass in a single read of the keys (rs_hist_all_kernel); then each pass is
// one kernel (rs_onesweep_kernel).  Its blocks take consecutive tickets, rank
// their keys as the downsweep does, publish their per-digit counts and obtain
// the counts of all earlier blocks by decoupled look-back over per-(block,
// digit) status words: bits 31..30 the state (0 none yet, 1 this block's count,
// 2 the inclusive count of all blocks up to it), bits 29..0 the count.  A
// block reaches back until it meets a state-2 word.  Tickets are handed out in
// the order blocks start, so every block looked at has started and publishes
// its own count before it waits on anything: the waits end.  A wait is
// nevertheless bounded (OS_SPIN_LIMIT sleeps, about a tenth of a second);
// past it the kernel sets *err and goes on, so a fault elsewhere cannot hang
// the device.  Each pass clears the status words of the next (its block b
// clears row b), so a sort needs one clearing launch.
constexpr int OS_MAX_PASSES = 8;
constexpr uint32_t OS_AGG = 1u << 30, OS_INC = 2u << 30, OS_CNT = (1u << 30) - 1u;
constexpr uint32_t OS_SPIN_LIMIT = 1u << 22;
constexpr uint32_t OS_HIST_BLOCKS = 512;
// keys per one-sweep block = OS_THREADS x OS_ROUNDS
constexpr int OS_THREADS = 256, OS_ROUNDS = 16;
// status words of earlier blocks read at once per look-back step (8, 16 and
// 32 measured slower at 47M keys)
constexpr int OS_WINDOW = 4;
// large sorts: decoupled look-back (true), or a counting kernel and a scan per
// pass in front of the same scatter kernel (false; 16% slower at 47M keys)
constexpr bool OS_LOOKBACK = true;
// the smallest sort that takes the one-sweep passes (below it, e.g. the ~1M-key
// depth sort, the three-launch passes on 2048-key blocks measured faster)
constexpr uint32_t OS_MIN_KEYS = RS_SMALL_LIMIT;

struct PassShifts {
    int s[OS_MAX_PASSES];
};

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Digit counts of every pass, summed into hist[pass * RADIX + digit] (zeroed
// beforehand).  LDS atomics into OS_HIST_COPIES copies of the counters, the
// copy chosen by lane % OS_HIST_COPIES, rows padded by one word so that the
// copies of one counter sit in different banks: a digit that most keys share
// (the exponent byte of a depth, the high tile bits of one Gaussian's run of
// pairs) costs a few conflicting lanes per instruction instead of a wave.
constexpr int OS_HIST_COPIES = 8;
constexpr int OS_HIST_THREADS = 1024;

// copies that fit 64 KB of LDS (eight, or four for eight 8-bit passes)
static int hist_copies(int radix_bits, int passes) {
    int c = OS_HIST_COPIES;
    while (c > 1 && (size_t)c * (((size_t)passes << radix_bits) + 1) * 4 > 65536) c /= 2;
    return c;
}
static size_t hist_lds_bytes(int radix_bits, int passes) {
    return (size_t)hist_copies(radix_bits, passes) * (((size_t)passes << radix_bits) + 1) * 4;
}

template <int BITS>
__global__ __launch_bounds__(OS_HIST_THREADS) void rs_hist_all_kernel(const uint64_t* __restrict__ keys,
                                                                      uint32_t n_cap,
                                                                      const uint32_t* __restrict__ n_dev,
                                                                      PassShifts sh, int passes, int copies,
                                                                      uint32_t* __restrict__ hist) {
    constexpr int RADIX = 1 << BITS;
    extern __shared__ uint32_t s_hc[];  // `copies` rows of passes * RADIX + 1 words
    const int row = passes * RADIX + 1;
    const uint32_t n = live_count(n_cap, n_dev);
    for (int i = threadIdx.x; i < copies * row; i += OS_HIST_THREADS) s_hc[i] = 0u;
    __syncthreads();
    uint32_t* mine = s_hc + (threadIdx.x & (copies - 1)) * row;
    const uint32_t stride = gridDim.x * (uint32_t)OS_HIST_THREADS;
    for (uint32_t i0 = blockIdx.x * (uint32_t)OS_HIST_THREADS + threadIdx.x; i0 < n; i0 += 4u * stride) {
        uint64_t k[4];
        bool ok[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t i = (uint64_t)i0 + (uint64_t)q * stride;
            ok[q] = i < n;
            k[q] = ok[q] ? keys[i] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!ok[q]) continue;
#pragma unroll
            for (int p = 0; p < OS_MAX_PASSES; p++)
                if (p < passes) atomicAdd(&mine[p * RADIX + digit_of<BITS>(k[q], sh.s[p])], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * RADIX; i += OS_HIST_THREADS) {
        uint32_t c = 0u;
        for (int k = 0; k < copies; k++) c += s_hc[k * row + i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// One pass: keys (and values) in, scattered stably by the digit at `shift`.
// LOOKBACK: the block's place among the blocks comes from its ticket, its
// digits' starts from the decoupled look-back (hist: this pass's digit totals;
// status / status_next: this pass's and the next pass's status words
// ([block][digit]); ticket: zeroed counter).  Otherwise from the scanned
// per-block counts of a counting kernel (goff: scanned digit-major table with
// row stride hs; dtot: digit totals), block = blockIdx.x.
template <int BITS, bool VALS, int ROUNDS, int THREADS, bool LOOKBACK>
__global__ __launch_bounds__(THREADS) void rs_scatter_kernel(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t n_cap,
    const uint32_t* __restrict__ n_dev, int shift, const uint32_t* __restrict__ hist, uint32_t* __restrict__ status,
    uint32_t* __restrict__ status_next, uint32_t* __restrict__ ticket, uint32_t* __restrict__ err, uint32_t hs,
    const uint32_t* __restrict__ goff, uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
    constexpr int RADIX = 1 << BITS;
    constexpr int WAVES = THREADS / 64;
    constexpr int RS_TILE = THREADS * ROUNDS, RS_WAVE_SPAN = ROUNDS * 64;
    static_assert(RADIX <= THREADS, "one digit per thread");
    const uint32_t n = live_count(n_cap, n_dev);
    // keys in digit order, then (digit << 32 | value) in the same slots
    __shared__ uint64_t s_k[RS_TILE];
    __shared__ uint32_t s_wh[WAVES][RADIX];
    __shared__ uint32_t s_dstart[RADIX];
    __shared__ uint32_t s_goff[RADIX];
    __shared__ uint32_t s_scan[WAVES];
    __shared__ uint32_t s_b;

    if (LOOKBACK && threadIdx.x == 0) s_b = atomicAdd(ticket, 1u);
    for (int i = threadIdx.x; i < WAVES * RADIX; i += THREADS) (&s_wh[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t b = LOOKBACK ? s_b : blockIdx.x;
    if (LOOKBACK && status_next && threadIdx.x < RADIX) status_next[(size_t)b * RADIX + threadIdx.x] = 0u;
    const uint32_t bbase = b * (uint32_t)RS_TILE;
    if (bbase >= n) return;  // past the live count: no block before it reads its words
    {
        // the digits' starts in the output (the look-back adds the earlier
        // blocks' counts; without it the scanned table holds them)
        const uint32_t t = threadIdx.x < RADIX ? hist[threadIdx.x] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<uint32_t>(t, s_scan, &total);
        if (threadIdx.x < RADIX) s_goff[threadIdx.x] = ex + (LOOKBACK ? 0u : goff[(size_t)threadIdx.x * hs + b]);
    }

    const int w = threadIdx.x >> 6, l = lane_id();
    const uint32_t wbase = bbase + w * RS_WAVE_SPAN;
    uint64_t k_r[ROUNDS];
    uint32_t v_r[VALS ? ROUNDS : 1];
    uint32_t rank_r[ROUNDS];  // digits are recomputed from the keys (fewer registers)
    // 1. per-wave stable ranks: rounds in key order, lanes in key order
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        const bool valid = i < n;
        k_r[r] = valid ? keys_in[i] : 0ull;
        if constexpr (VALS) v_r[r] = valid ? vals_in[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        const bool valid = i < n;
        const uint64_t active = __ballot(valid);
        const uint32_t d = digit_of<BITS>(k_r[r], shift);
        rank_r[r] = 0u;
        if (active != 0) {
            const uint64_t peers = match_digit<BITS>(d, active);
            uint32_t before = 0u;
            if (valid) before = s_wh[w][d];
            __builtin_amdgcn_wave_barrier();
            if (valid && (peers & lanemask_lt()) == 0) s_wh[w][d] = before + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            rank_r[r] = before + (uint32_t)__popcll(peers & lanemask_lt());
        }
    }
    __syncthreads();
    // 2. the block's count of each digit (published at once) and per-wave offsets
    uint32_t cnt = 0u;
    if (threadIdx.x < RADIX) {
        const int d = threadIdx.x;
#pragma unroll
        for (int k = 0; k < WAVES; k++) {
            const uint32_t c = s_wh[k][d];
            s_wh[k][d] = cnt;
            cnt += c;
        }
        if (LOOKBACK) st_agent(&status[(size_t)b * RADIX + d], (b == 0 ? OS_INC : OS_AGG) | cnt);
    }
    {
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<uint32_t>(cnt, s_scan, &total);
        if (threadIdx.x < RADIX) s_dstart[threadIdx.x] = ex;
    }
    // 3. look-back: the counts of this digit in all earlier blocks, OS_WINDOW words at a time
    if (LOOKBACK && threadIdx.x < RADIX && b > 0) {
        const int d = threadIdx.x;
        const uint32_t* col = status + d;
        uint32_t acc = 0u;
        int j = (int)b - 1;
        bool done = false;
        while (!done) {
            uint32_t s[OS_WINDOW];
#pragma unroll
            for (int q = 0; q < OS_WINDOW; q++)
                s[q] = (j - q >= 0) ? ld_agent(&col[(size_t)(j - q) * RADIX]) : OS_INC;
#pragma unroll
            for (int q = 0; q < OS_WINDOW; q++) {
                if (done) break;
                uint32_t v = s[q];
                uint32_t spins = 0u;
                while ((v & ~OS_CNT) == 0u) {
                    if (++spins > OS_SPIN_LIMIT) {
                        *err = 1u;
                        v = OS_INC;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    v = ld_agent(&col[(size_t)(j - q) * RADIX]);
                }
                acc += v & OS_CNT;
                if (v & OS_INC) done = true;
            }
            j -= OS_WINDOW;
        }
        st_agent(&status[(size_t)b * RADIX + d], OS_INC | (acc + cnt));
        s_goff[d] += acc;
    }
    __syncthreads();
    // from here on s_goff[d] is the output index of the block's slot 0 of digit d's run
    // minus that slot: output index = s_goff[d] + slot
    if (threadIdx.x < RADIX) s_goff[threadIdx.x] -= s_dstart[threadIdx.x];
    // 4. stage the keys in LDS in (digit, input order) order and write them out in runs
    uint32_t p_r[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        const uint32_t d = digit_of<BITS>(k_r[r], shift);
        p_r[r] = s_dstart[d] + s_wh[w][d] + rank_r[r];
        if (i < n && p_r[r] < (uint32_t)RS_TILE) s_k[p_r[r]] = k_r[r];
    }
    __syncthreads();
    const uint32_t count = (n - bbase < (uint32_t)RS_TILE) ? (n - bbase) : (uint32_t)RS_TILE;
    for (uint32_t p = threadIdx.x; p < count; p += THREADS) {
        const uint64_t k = s_k[p];
        const uint32_t g = s_goff[digit_of<BITS>(k, shift)] + p;
        if (g < n) keys_out[g] = k;
    }
    if constexpr (VALS) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            const uint32_t i = wbase + r * 64 + l;
            if (i < n && p_r[r] < (uint32_t)RS_TILE) s_k[p_r[r]] = ((uint64_t)digit_of<BITS>(k_r[r], shift) << 32) | v_r[r];
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < count; p += THREADS) {
            const uint64_t e = s_k[p];
            const uint32_t g = s_goff[(uint32_t)(e >> 32)] + p;
            if (g < n) vals_out[g] = (uint32_t)e;
        }
    }
}

// Per-block digit counts for the scatter kernel without look-back:
// hist[digit * hs + block] (digit-major, as rs_digit_scan_kernel reads it).
// LDS atomics into eight padded copies of the counters (see rs_hist_all_kernel).
template <int BITS, int ROUNDS, int THREADS>
__global__ __launch_bounds__(THREADS) void rs_count_kernel(const uint64_t* __restrict__ keys, uint32_t n_cap,
                                                           const uint32_t* __restrict__ n_dev, int shift, uint32_t hs,
                                                           uint32_t* __restrict__ hist) {
    constexpr int RADIX = 1 << BITS, COPIES = 8, ROW = RADIX + 1;
    constexpr int TILE = THREADS * ROUNDS;
    __shared__ uint32_t s_h[COPIES * ROW];
    const uint32_t n = live_count(n_cap, n_dev);
    for (int i = threadIdx.x; i < COPIES * ROW; i += THREADS) s_h[i] = 0u;
    __syncthreads();
    uint32_t* mine = s_h + (threadIdx.x & (COPIES - 1)) * ROW;
    const uint32_t base = blockIdx.x * (uint32_t)TILE + threadIdx.x;
    uint64_t k[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = base + r * THREADS;
        k[r] = i < n ? keys[i] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; r++)
        if (base + r * THREADS < n) atomicAdd(&mine[digit_of<BITS>(k[r], shift)], 1u);
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += THREADS) {
        uint32_t c = 0u;
#pragma unroll
        for (int q = 0; q < COPIES; q++) c += s_h[q * ROW + d];
        hist[(size_t)d * hs + blockIdx.x] = c;
    }
}

// identify_tile_ranges (radix_sort.cl:14-46), ranges pre-zeroed.
// Tile ids at or beyond num_tiles are ignored (they cannot come from binning).
__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint64_t* __restrict__ keys, uint32_t n_cap,
                                                          const uint32_t* __restrict__ n_dev, uint32_t num_tiles,
                                                          uint32_t* __restrict__ ranges) {
    const uint32_t n = live_count(n_cap, n_dev);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t cur = (uint32_t)(keys[i] >> 32);
    if (i == 0) {
        if (cur < num_tiles) ranges[cur * 2] = 0;
    } else {
        uint32_t prev = (uint32_t)(keys[i - 1] >> 32);
        if (cur != prev) {
            if (prev < num_tiles) ranges[prev * 2 + 1] = i;
            if (cur < num_tiles) ranges[cur * 2] = i;
        }
    }
    if (i == n - 1 && cur < num_tiles) ranges[cur * 2 + 1] = n;
}

}  // namespace gsk

using namespace gsk;

uint32_t scan_partials_needed(uint32_t n) { return (n + SCAN_BLOCK - 1) / SCAN_BLOCK; }

// Inclusive scan of int32 tiles_touched into uint32 offsets.
hipError_t launch_scan_tiles(const int* in, uint32_t n, uint32_t* out, unsigned long long* partial,
                             uint32_t* total32, unsigned long long* total64, hipStream_t s) {
    uint32_t nb = scan_partials_needed(n);
    if (nb == 0) nb = 1;
    scan_reduce_kernel<int><<<nb, SCAN_THREADS, 0, s>>>(in, n, partial);
    scan_partials_kernel<<<1, 1024, 0, s>>>(partial, nb, total32, total64);
    scan_downsweep_kernel<int, true><<<nb, SCAN_THREADS, 0, s>>>(in, n, partial, out);
    return hipGetLastError();
}

// Exclusive scan of uint32 counts (in != out).
hipError_t launch_scan_u32_exclusive(const uint32_t* in, uint32_t n, uint32_t* out,
                                            unsigned long long* partial, hipStream_t s, uint32_t* total32,
                                            unsigned long long* total64) {
    uint32_t nb = scan_partials_needed(n);
    scan_reduce_kernel<uint32_t><<<nb, SCAN_THREADS, 0, s>>>(in, n, partial);
    scan_partials_kernel<<<1, 1024, 0, s>>>(partial, nb, total32, total64);
    scan_downsweep_kernel<uint32_t, false><<<nb, SCAN_THREADS, 0, s>>>(in, n, partial, out);
    return hipGetLastError();
}

hipError_t launch_duplicate(int n, const float* means2D, const float* depths, const int* radii,
                            const uint32_t* offsets, int gx, int gy, uint64_t* keys, uint32_t* values,
                            uint32_t capacity, uint32_t* overflow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    duplicate_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, reinterpret_cast<const float2*>(means2D), depths, radii,
                                                     offsets, gx, gy, keys, values, capacity, overflow);
    return hipGetLastError();
}

static int radix_rounds(uint32_t n) { return n >= RS_SMALL_LIMIT ? 16 : 8; }
uint32_t radix_blocks(uint32_t n) {
    uint32_t tile = (uint32_t)(RS_THREADS * radix_rounds(n));
    return (n + tile - 1) / tile;
}

// Workspace layout for one sort of n items with the given radix:
//   [ping keys n*8][ping vals n*4][small: OS_MAX_PASSES*R + 64 words][table A R*nb*4][table B R*nb*4]
// Three-launch passes: table A = per-block digit counts, table B = their
// scans, small = digit totals.  One-sweep passes: small = every pass's digit
// totals, the per-pass tickets and the error word; tables A and B = the status
// words of alternate passes.  Row stride of the digit-major count table: the
// block count rounded up to a multiple of 4 (16-byte rows for the digit scan).
static uint32_t radix_row_stride(uint32_t nb) { return (nb + 3u) & ~3u; }

static size_t small_words(int radix_bits) { return ((size_t)OS_MAX_PASSES << radix_bits) + 64; }

static uint32_t os_blocks(uint32_t n) { return (n + OS_THREADS * OS_ROUNDS - 1) / (OS_THREADS * OS_ROUNDS); }

// rows of the two tables: the three-launch passes' blocks or the one-sweep blocks
static size_t table_rows(uint32_t n) {
    return std::max<size_t>(radix_row_stride(radix_blocks(n)), radix_row_stride(os_blocks(n)));
}

size_t radix_workspace_bytes(uint32_t n, int radix_bits) {
    size_t nb = table_rows(n);
    size_t R = (size_t)1 << radix_bits;
    size_t tab = R * nb;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return al((size_t)n * 8) + al((size_t)n * 4) + al(small_words(radix_bits) * 4) + 2 * al(tab * 4);
}

// the one-sweep path: sorts of at least OS_MIN_KEYS whose counts fit the status words' 30 bits
static bool use_onesweep(uint32_t n, int passes) {
    return n >= OS_MIN_KEYS && n <= OS_CNT && passes <= OS_MAX_PASSES;
}

template <int BITS>
static hipError_t onesweep_sort(const uint64_t* ki, const uint32_t* vi, uint64_t* keys_out, uint32_t* vals_out,
                                uint64_t* pk, uint32_t* pv, uint32_t n, const uint32_t* n_dev, const int* shifts,
                                int passes, uint32_t* small, uint32_t* tab_a, uint32_t* tab_b, hipStream_t s) {
    constexpr int RADIX = 1 << BITS;
    constexpr int ROUNDS = OS_ROUNDS, THREADS = OS_THREADS;
    const uint32_t nb = os_blocks(n), hs = radix_row_stride(nb);
    uint32_t* hist = small;                                       // [passes][RADIX]
    uint32_t* tickets = small + ((size_t)OS_MAX_PASSES << BITS);  // [OS_MAX_PASSES]
    uint32_t* err = tickets + OS_MAX_PASSES;
    const bool vals = vi != nullptr;
    if (OS_LOOKBACK) {
        // the small words and the first pass's status words
        hipError_t e = launch_zero(small, small_words(BITS) * 4, s);
        if (e == hipSuccess) e = launch_zero(tab_a, (size_t)nb * RADIX * 4, s);
        if (e != hipSuccess) return e;
        PassShifts sh{};
        for (int p = 0; p < passes; p++) sh.s[p] = shifts[p];
        const uint32_t hb = std::min<uint32_t>(OS_HIST_BLOCKS, (n + 16383) / 16384);
        rs_hist_all_kernel<BITS><<<hb, OS_HIST_THREADS, hist_lds_bytes(BITS, passes), s>>>(
            ki, n, n_dev, sh, passes, hist_copies(BITS, passes), hist);
    }
    for (int p = 0; p < passes; p++) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint64_t* ko = to_out ? keys_out : pk;
        uint32_t* vo = vals ? (to_out ? vals_out : pv) : nullptr;
        if (OS_LOOKBACK) {
            uint32_t* st = (p % 2 == 0) ? tab_a : tab_b;
            uint32_t* st_next = (p + 1 < passes) ? ((p % 2 == 0) ? tab_b : tab_a) : nullptr;
            if (vals)
                rs_scatter_kernel<BITS, true, ROUNDS, THREADS, true><<<nb, THREADS, 0, s>>>(
                    ki, vi, n, n_dev, shifts[p], hist + (size_t)p * RADIX, st, st_next, tickets + p, err, 0u, nullptr,
                    ko, vo);
            else
                rs_scatter_kernel<BITS, false, ROUNDS, THREADS, true><<<nb, THREADS, 0, s>>>(
                    ki, nullptr, n, n_dev, shifts[p], hist + (size_t)p * RADIX, st, st_next, tickets + p, err, 0u,
                    nullptr, ko, nullptr);
        } else {
            // per-block counts (table A), their scan per digit (table B) and the digit totals
            rs_count_kernel<BITS, ROUNDS, THREADS><<<nb, THREADS, 0, s>>>(ki, n, n_dev, shifts[p], hs, tab_a);
            rs_digit_scan_kernel<<<RADIX, 64, 0, s>>>(tab_a, nb, hs, tab_b, hist);
            if (vals)
                rs_scatter_kernel<BITS, true, ROUNDS, THREADS, false><<<nb, THREADS, 0, s>>>(
                    ki, vi, n, n_dev, shifts[p], hist, nullptr, nullptr, nullptr, nullptr, hs, tab_b, ko, vo);
            else
                rs_scatter_kernel<BITS, false, ROUNDS, THREADS, false><<<nb, THREADS, 0, s>>>(
                    ki, nullptr, n, n_dev, shifts[p], hist, nullptr, nullptr, nullptr, nullptr, hs, tab_b, ko,
                    nullptr);
        }
        ki = ko;
        vi = vo;
    }
    return hipGetLastError();
}

template <int BITS, int ROUNDS>
static hipError_t sort_pass_r(const uint64_t* ki, const uint32_t* vi, uint64_t* ko, uint32_t* vo, uint32_t n,
                              const uint32_t* n_dev, int shift, uint32_t* hist, uint32_t* goff,
                              uint32_t* dtot, hipStream_t s) {
    const uint32_t nb = radix_blocks(n), hs = radix_row_stride(nb);
    rs_upsweep_kernel<BITS, ROUNDS><<<nb, RS_THREADS, 0, s>>>(ki, n, n_dev, shift, hs, hist);
    rs_digit_scan_kernel<<<1u << BITS, 64, 0, s>>>(hist, nb, hs, goff, dtot);
    if (vi)
        rs_downsweep_kernel<BITS, true, ROUNDS>
            <<<nb, RS_THREADS, 0, s>>>(ki, vi, n, n_dev, shift, hs, goff, dtot, ko, vo);
    else
        rs_downsweep_kernel<BITS, false, ROUNDS>
            <<<nb, RS_THREADS, 0, s>>>(ki, nullptr, n, n_dev, shift, hs, goff, dtot, ko, nullptr);
    return hipGetLastError();
}

template <int BITS>
static hipError_t sort_pass(const uint64_t* ki, const uint32_t* vi, uint64_t* ko, uint32_t* vo, uint32_t n,
                            const uint32_t* n_dev, int shift, uint32_t* hist, uint32_t* goff, uint32_t* dtot,
                            hipStream_t s) {
    if (radix_rounds(n) == 16) return sort_pass_r<BITS, 16>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s);
    if (radix_rounds(n) == 8) return sort_pass_r<BITS, 8>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s);
    return sort_pass_r<BITS, 4>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s);
}

// Stable LSD sort of n keys (with optional uint32 values) on key bits
// [begin_bit, end_bit).  The last pass lands in keys_out / vals_out.
hipError_t launch_radix_sort_bits(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                                  uint32_t* vals_out, uint32_t n, int begin_bit, int end_bit, int radix_bits,
                                  void* workspace, hipStream_t s, const uint32_t* n_dev) {
    if (n == 0) return hipSuccess;
    size_t nb = table_rows(n);
    size_t tab = ((size_t)1 << radix_bits) * nb;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char* w = static_cast<char*>(workspace);
    uint64_t* pk = reinterpret_cast<uint64_t*>(w);
    w += al((size_t)n * 8);
    uint32_t* pv = reinterpret_cast<uint32_t*>(w);
    w += al((size_t)n * 4);
    uint32_t* small = reinterpret_cast<uint32_t*>(w);
    w += al(small_words(radix_bits) * 4);
    uint32_t* tab_a = reinterpret_cast<uint32_t*>(w);
    w += al(tab * 4);
    uint32_t* tab_b = reinterpret_cast<uint32_t*>(w);

    const bool vals = vals_in != nullptr;
    int bits = end_bit - begin_bit;
    int passes = bits > 0 ? (bits + radix_bits - 1) / radix_bits : 0;
    if (passes <= 0) {
        hipError_t e = hipMemcpyAsync(keys_out, keys_in, (size_t)n * 8, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess || !vals) return e;
        return hipMemcpyAsync(vals_out, vals_in, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
    }
    // A last digit that would reach past end_bit is moved down to end at it:
    // the bits above end_bit may be set (the depth-sort words carry the index
    // there).  Re-sorting bits an earlier pass already sorted keeps a stable
    // LSD sort correct.
    int shifts[64];
    for (int p = 0; p < passes; p++) {
        int shift = begin_bit + p * radix_bits;
        if (shift + radix_bits > end_bit && end_bit - radix_bits >= begin_bit) shift = end_bit - radix_bits;
        shifts[p] = shift;
    }
    if (use_onesweep(n, passes)) {
        switch (radix_bits) {
            case 4: return onesweep_sort<4>(keys_in, vals_in, keys_out, vals_out, pk, pv, n, n_dev, shifts, passes,
                                            small, tab_a, tab_b, s);
            case 5: return onesweep_sort<5>(keys_in, vals_in, keys_out, vals_out, pk, pv, n, n_dev, shifts, passes,
                                            small, tab_a, tab_b, s);
            case 6: return onesweep_sort<6>(keys_in, vals_in, keys_out, vals_out, pk, pv, n, n_dev, shifts, passes,
                                            small, tab_a, tab_b, s);
            case 7: return onesweep_sort<7>(keys_in, vals_in, keys_out, vals_out, pk, pv, n, n_dev, shifts, passes,
                                            small, tab_a, tab_b, s);
            case 8: return onesweep_sort<8>(keys_in, vals_in, keys_out, vals_out, pk, pv, n, n_dev, shifts, passes,
                                            small, tab_a, tab_b, s);
            default: return hipErrorInvalidValue;
        }
    }
    uint32_t* hist = tab_a;
    uint32_t* goff = tab_b;
    uint32_t* dtot = small;
    const uint64_t* ki = keys_in;
    const uint32_t* vi = vals_in;
    for (int p = 0; p < passes; p++) {
        // the last pass lands in the output; earlier passes alternate so that it does
        bool to_out = ((passes - 1 - p) % 2) == 0;
        uint64_t* ko = to_out ? keys_out : pk;
        uint32_t* vo = vals ? (to_out ? vals_out : pv) : nullptr;
        hipError_t e;
        const int shift = shifts[p];
        switch (radix_bits) {
            case 4: e = sort_pass<4>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s); break;
            case 5: e = sort_pass<5>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s); break;
            case 6: e = sort_pass<6>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s); break;
            case 7: e = sort_pass<7>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s); break;
            case 8: e = sort_pass<8>(ki, vi, ko, vo, n, n_dev, shift, hist, goff, dtot, s); break;
            default: return hipErrorInvalidValue;
        }
        if (e != hipSuccess) return e;
        ki = ko;
        vi = vo;
    }
    return hipSuccess;
}

hipError_t launch_radix_sort(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out, uint32_t* vals_out,
                             uint32_t n, int key_bits, int radix_bits, void* workspace, hipStream_t s) {
    return launch_radix_sort_bits(keys_in, vals_in, keys_out, vals_out, n, 0, key_bits, radix_bits, workspace, s,
                                  nullptr);
}

hipError_t launch_depth_keys(int n, const float* depths, const int* radii, uint64_t* keys, uint32_t* idx,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    depth_keys_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, depths, radii, keys, idx);
    return hipGetLastError();
}

hipError_t launch_depth_words(int n, const float* depths, const int* radii, uint64_t* words, hipStream_t s,
                              const float* means2D, int gx, int gy, uint32_t* rects) {
    if (n <= 0) return hipSuccess;
    depth_words_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, depths, radii, words,
                                                       reinterpret_cast<const float2*>(means2D), gx, gy, rects);
    return hipGetLastError();
}

hipError_t launch_gather_tiles(int n, const uint32_t* order, const int* tiles, int* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gather_tiles_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, order, tiles, out);
    return hipGetLastError();
}

hipError_t launch_duplicate_ranked(int n, const uint32_t* order, const float* means2D, const int* radii,
                                   const uint32_t* offsets, int gx, int gy, uint64_t* words, uint32_t capacity,
                                   uint32_t* overflow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    duplicate_ranked_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, order, reinterpret_cast<const float2*>(means2D), radii,
                                                            offsets, gx, gy, words, capacity, overflow);
    return hipGetLastError();
}

hipError_t launch_gather_tiles_multi(uint32_t total, int N, const uint64_t* skeys, const uint32_t* sidx,
                                     const int* tiles, int* out, hipStream_t s) {
    if (total == 0) return hipSuccess;
    gather_tiles_multi_kernel<<<(total + 255) / 256, 256, 0, s>>>(total, N, skeys, sidx, tiles, out);
    return hipGetLastError();
}

hipError_t launch_duplicate_multi(uint32_t total, int N, const uint64_t* skeys, const uint32_t* sidx,
                                  const float* means2D, const int* radii, const uint32_t* offsets, const int4* views,
                                  uint64_t* words, uint32_t capacity, uint32_t* overflow, hipStream_t s) {
    if (total == 0) return hipSuccess;
    duplicate_multi_kernel<<<(total + 255) / 256, 256, 0, s>>>(total, N, skeys, sidx,
                                                               reinterpret_cast<const float2*>(means2D), radii,
                                                               offsets, views, words, capacity, overflow);
    return hipGetLastError();
}

hipError_t launch_tile_ranges_multi(const uint64_t* words, uint32_t n_cap, const uint32_t* n_dev, const int4* views,
                                    uint32_t* ranges, hipStream_t s) {
    if (n_cap == 0) return hipSuccess;
    tile_ranges_multi_kernel<<<(n_cap + 255) / 256, 256, 0, s>>>(words, n_cap, n_dev, views, ranges);
    return hipGetLastError();
}

hipError_t launch_tile_ranges(const uint64_t* keys_sorted, uint32_t n, uint32_t num_tiles, uint32_t* ranges,
                              hipStream_t s, const uint32_t* n_dev) {
    if (n == 0) return hipSuccess;
    tile_ranges_kernel<<<(n + 255) / 256, 256, 0, s>>>(keys_sorted, n, n_dev, num_tiles, ranges);
    return hipGetLastError();
}
