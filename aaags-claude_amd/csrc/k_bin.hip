// k_bin.hip -- tile binning by direct placement (the render/train fast path).
//
// Replaces duplicate-with-keys + the pair sort (tile_binner_cpu.cpp:59-99,
// sorter_cpu.cpp:6-44; scatter.cl:73-154 + radix_sort_kv_cl.h) with the same
// result: per tile, the indices of the Gaussians whose tile rectangle covers
// it, in (depth bits, index) order, plus [start, end) ranges (empty tiles
// (0, 0)).
//
// Input is each view's Gaussians in depth order: a stable sort by the float
// bits of depth, so equal depths keep index order, with culled Gaussians last.
// That list is cut into groups (S groups of G1 Gaussians, then groups of G),
// and each view's tiles into bins of 8 x (8 * TPL) tiles.  One single-wave
// block handles one (group, bin), and
// lane l owns the TPL tiles (l % 8, l / 8 + 8 k) of the bin: a lane's counters
// and cursors live in its registers, so nothing is shared and no atomics are
// needed, and the order in which the wave visits Gaussians is the order each
// tile sees them.
//   rects:  the depth-ordered Gaussians as packed tile rectangles (8 bytes).
//   count:  each block walks its group in depth order; every lane counts the
//           Gaussians whose rectangle covers each of its tiles.  The counts go
//           to a table laid out tile-major, (view, tile, group), plus one
//           trailing entry.
//   scan:   one exclusive scan of the table gives, for every (tile, group), the
//           output slot of the group's first pair in that tile (the tile-major
//           layout makes the tile starts part of the same scan), and the total.
//   place:  the same walk again with each lane's cursors set from the scanned
//           table; each Gaussian index goes to slot cursor++ -- its stable rank
//           in the tile.  A lane collects its tile's slots in a 16-entry
//           (64-byte) LDS line buffer; when lines fill, the wave writes them
//           out one by one, 16 lanes per line, so HBM sees whole-line writes
//           except at the two ends of each (tile, group) run.
//   ranges: (start, end) of a tile are the scanned entries of its first group
//           and of the next tile's first group.
// A ballot over 64 rectangles at a time skips the Gaussians that miss the bin,
// so a block's cost is its group's length / 64 plus the Gaussians that touch
// its bin.  HBM traffic: 4 B per pair written once, the 8-byte records (read
// once per bin, mostly from cache), and the table (4 B per tile and group).
#include "gs_device.h"

namespace gsb {

constexpr int LINE = 16;  // pair slots per 64-byte output line

// Depth-ordered rectangle records, one per (view, rank): (x0 | x1 << 8 |
// y0 << 16 | y1 << 24, Gaussian index); culled Gaussians get an empty
// rectangle.  Views are at most 255 x 255 tiles.
__global__ __launch_bounds__(256) void bin_rects_kernel(BinArgs a) {
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (size_t)a.V * a.N) return;
    const int v = (int)(r / (size_t)a.N);
    const int4 vi = a.views ? a.views[v] : a.view0;
    const size_t seg = (size_t)v * a.N;
    const uint32_t id = a.sidx[r * (size_t)a.sidx_stride];
    if (a.rect_idx) {  // packed beforehand in index order: one word gathered
        a.rects[r] = make_uint2(a.rect_idx[seg + id], id);
        return;
    }
    const int rad = a.radii[seg + id];
    uint2 rec = make_uint2(0u, id);
    if (rad > 0) {
        const float2 m = a.means2D[seg + id];
        int mn[2], mx[2];
        gsd::get_rect(m.x, m.y, rad, vi.x, vi.y, mn, mx);
        rec.x = (uint32_t)mn[0] | ((uint32_t)mx[0] << 8) | ((uint32_t)mn[1] << 16) | ((uint32_t)mx[1] << 24);
    }
    a.rects[r] = rec;
}

// Ranks [r0, r1) of group c in a view's depth order: S groups of G1 ranks,
// then groups of G.
__device__ __forceinline__ void group_ranks(const BinArgs& a, int c, uint32_t& r0, uint32_t& r1) {
    uint32_t len;
    if (c < a.S) {
        r0 = (uint32_t)c * (uint32_t)a.G1;
        len = (uint32_t)a.G1;
    } else {
        r0 = (uint32_t)a.S * (uint32_t)a.G1 + (uint32_t)(c - a.S) * (uint32_t)a.G;
        len = (uint32_t)a.G;
    }
    r1 = (r0 + len < (uint32_t)a.N) ? r0 + len : (uint32_t)a.N;
}

// The (view, group, bin) of this block.  Blocks of one view are numbered
// group-major, so the blocks of one XCD (after the remap) share a few groups
// and their records stay in that XCD's L2.
template <int TPL>
struct BinBlock {
    int v, c, bin, bx0, by0;
    int4 vi;  // (tiles per row, tile rows, table offset, first global tile)
};

template <int TPL>
__device__ __forceinline__ BinBlock<TPL> bin_block(const BinArgs& a, int block0 = 0) {
    const int q = block0 + gsd::xcd_remap(blockIdx.x, gridDim.x);
    BinBlock<TPL> b;
    b.v = 0;
    if (a.vblock0)
        while (b.v + 1 < a.V && a.vblock0[b.v + 1] <= q) b.v++;
    const int local = q - (a.vblock0 ? a.vblock0[b.v] : 0);
    b.vi = a.views ? a.views[b.v] : a.view0;
    const int nbx = (b.vi.x + 7) / 8, nby = (b.vi.y + 8 * TPL - 1) / (8 * TPL);
    b.c = local / (nbx * nby);
    const int bin = local - b.c * (nbx * nby);
    b.bin = bin;
    b.bx0 = (bin % nbx) * 8;
    b.by0 = (bin / nbx) * (8 * TPL);
    return b;
}

// Walks the block's group in depth order and, for every Gaussian that meets the
// bin, calls f(in, gid) on all lanes with in[k] set where the lane's k-th tile
// is covered by the Gaussian's rectangle.  Rectangles are read WALK_DEPTH x 64
// at a time, one per lane per 64, and the next round is prefetched, so that
// several loads are in flight per wave; the ones that meet the bin are visited
// lowest lane first, round by round, i.e. in depth order.
constexpr int WALK_DEPTH = 4;

template <int TPL, typename F>
__device__ __forceinline__ void walk_bin(const BinArgs& a, const BinBlock<TPL>& b, int tx, int ty, int lane, F&& f) {
    const uint2* rects = a.rects + (size_t)b.v * a.N;
    uint32_t r0, r1;
    group_ranks(a, b.c, r0, r1);
    const int bx1 = b.bx0 + 8, by1 = b.by0 + 8 * TPL;
    constexpr uint32_t STEP = 64u * WALK_DEPTH;
    uint2 nxt[WALK_DEPTH];
#pragma unroll
    for (int q = 0; q < WALK_DEPTH; q++) {
        const uint32_t r = r0 + 64u * q + (uint32_t)lane;
        nxt[q] = r < r1 ? rects[r] : make_uint2(0u, 0u);
    }
    for (uint32_t base = r0; base < r1; base += STEP) {
        uint2 cur[WALK_DEPTH];
#pragma unroll
        for (int q = 0; q < WALK_DEPTH; q++) {
            cur[q] = nxt[q];
            const uint32_t r = base + STEP + 64u * q + (uint32_t)lane;
            if (r < r1) nxt[q] = rects[r];
        }
#pragma unroll
        for (int q = 0; q < WALK_DEPTH; q++) {
            const uint2 rec = cur[q];
            const int x0 = (int)(rec.x & 0xFFu), x1 = (int)((rec.x >> 8) & 0xFFu);
            const int y0 = (int)((rec.x >> 16) & 0xFFu), y1 = (int)(rec.x >> 24);
            const bool live = base + 64u * q + (uint32_t)lane < r1;
            uint64_t todo =
                __ballot(live && x0 < x1 && y0 < y1 && x0 < bx1 && x1 > b.bx0 && y0 < by1 && y1 > b.by0);
            while (todo) {
                const int j = __builtin_ctzll(todo);
                todo &= todo - 1ull;
                const uint32_t R = (uint32_t)__builtin_amdgcn_readlane((int)rec.x, j);
                const uint32_t gid = (uint32_t)__builtin_amdgcn_readlane((int)rec.y, j);
                const int jx0 = (int)(R & 0xFFu), jx1 = (int)((R >> 8) & 0xFFu);
                const int jy0 = (int)((R >> 16) & 0xFFu), jy1 = (int)(R >> 24);
                const bool inx = tx >= jx0 && tx < jx1;
                bool in[TPL];
#pragma unroll
                for (int k = 0; k < TPL; k++) {
                    const int y = ty + 8 * k;
                    in[k] = inx && y >= jy0 && y < jy1;
                }
                f(in, gid);
            }
        }
    }
}

template <int TPL>
__global__ __launch_bounds__(64) void bin_count_kernel(BinArgs a) {
    const BinBlock<TPL> b = bin_block<TPL>(a);
    const int lane = threadIdx.x;
    if (blockIdx.x == 0 && lane == 0) a.table[a.table_len - 1] = 0u;
    const int tx = b.bx0 + (lane & 7), ty = b.by0 + (lane >> 3);
    uint32_t cnt[TPL];
#pragma unroll
    for (int k = 0; k < TPL; k++) cnt[k] = 0u;
    walk_bin<TPL>(a, b, tx, ty, lane, [&](const bool (&in)[TPL], uint32_t) {
#pragma unroll
        for (int k = 0; k < TPL; k++) cnt[k] += in[k] ? 1u : 0u;
    });
    const int gx = b.vi.x, gy = b.vi.y;
    const size_t tab = (size_t)(uint32_t)b.vi.z;
#pragma unroll
    for (int k = 0; k < TPL; k++) {
        const int y = ty + 8 * k;
        if (tx < gx && y < gy) a.table[tab + (size_t)(y * gx + tx) * a.C + b.c] = cnt[k];
    }
}

// The same table from a two-dimensional difference array: one workgroup per
// (view, group).  A rectangle [x0, x1) x [y0, y1) adds +1 at (x0, y0) and
// (x1, y1) and -1 at (x1, y0) and (x0, y1) of a (tiles per row + 1) x (tile
// rows + 1) array in LDS; the array's prefix sum along rows, then along
// columns, is each tile's count.  Four LDS atomics per Gaussian whatever its
// size, instead of one walk over the group per bin.  Used when the array of
// the largest view fits in DIFF_MAX entries.
constexpr int DIFF_MAX = 16384;  // 64 KB of LDS

__global__ __launch_bounds__(256) void bin_count_diff_kernel(BinArgs a) {
    extern __shared__ int dif[];
    const int q = blockIdx.x;
    const int v = q / a.C, c = q - v * a.C;
    const int4 vi = a.views ? a.views[v] : a.view0;
    const int gx = vi.x, gy = vi.y, W1 = gx + 1;
    const int len = W1 * (gy + 1);
    const int tid = threadIdx.x;
    if (q == 0 && tid == 0) a.table[a.table_len - 1] = 0u;
    for (int i = tid; i < len; i += 256) dif[i] = 0;
    __syncthreads();
    const uint2* rects = a.rects + (size_t)v * a.N;
    uint32_t r0, r1;
    group_ranks(a, c, r0, r1);
    for (uint32_t r = r0 + (uint32_t)tid; r < r1; r += 256u) {
        const uint32_t R = rects[r].x;
        const int x0 = (int)(R & 0xFFu), x1 = (int)((R >> 8) & 0xFFu);
        const int y0 = (int)((R >> 16) & 0xFFu), y1 = (int)(R >> 24);
        if (x0 < x1 && y0 < y1) {
            atomicAdd(&dif[y0 * W1 + x0], 1);
            atomicAdd(&dif[y0 * W1 + x1], -1);
            atomicAdd(&dif[y1 * W1 + x0], -1);
            atomicAdd(&dif[y1 * W1 + x1], 1);
        }
    }
    __syncthreads();
    // rows: one per thread (a row's stride W1 is odd at the sizes that matter,
    // so a wave's 64 rows fall in distinct banks)
    for (int y = tid; y < gy; y += 256) {
        int* row = dif + y * W1;
        int s = 0;
#pragma unroll 8
        for (int x = 0; x < gx; x++) {
            s += row[x];
            row[x] = s;
        }
    }
    __syncthreads();
    // columns: one per thread; the column sums are the counts
    const size_t tab = (size_t)(uint32_t)vi.z;
    for (int x = tid; x < gx; x += 256) {
        int s = 0;
#pragma unroll 4
        for (int y = 0; y < gy; y++) {
            s += dif[y * W1 + x];
            a.table[tab + (size_t)(y * gx + x) * a.C + c] = (uint32_t)s;
        }
    }
}

// Writes slots [lo, hi] (one line, hi & 15 the line's last slot or its
// current end) of lane j's k-th line buffer with the whole wave: lane l < 16
// moves slot l, so the line leaves as one coalesced 64-byte store (fewer lanes
// for the partial lines at the ends of a run).  Slots past the capacity are
// dropped and raise the overflow flag.
__device__ __forceinline__ void flush_line_wave(const BinArgs& a, const uint32_t* line, uint32_t lo, uint32_t hi,
                                                int lane) {
    const uint32_t idx = (hi & ~(uint32_t)(LINE - 1)) + (uint32_t)lane;
    if (lane < LINE && idx >= lo && idx <= hi) {
        if (idx < a.capacity) a.out[idx] = line[lane];
        else *a.overflow = 1u;
    }
}

// block0 / gate: the launch covers logical blocks block0 .. (a later range of
// groups), and a block none of whose tiles has a nonzero gate word (per tile,
// one view) does nothing.
template <int TPL>
__global__ __launch_bounds__(64) void bin_place_kernel(BinArgs a, int block0, const uint32_t* __restrict__ gate) {
    // per (tile slot, lane): one line of LINE slots plus one spare slot that
    // takes the writes of lanes whose tile the Gaussian does not cover, so
    // the visit has no per-lane branch
    constexpr int LSTRIDE = LINE + 1;
    __shared__ uint32_t lines[TPL * 64 * LSTRIDE];
    const BinBlock<TPL> b = bin_block<TPL>(a, block0);
    const int lane = threadIdx.x;
    const int tx = b.bx0 + (lane & 7), ty = b.by0 + (lane >> 3);
    const int gx = b.vi.x, gy = b.vi.y;
    if (gate) {
        bool open = false;
#pragma unroll
        for (int k = 0; k < TPL; k++) {
            const int y = ty + 8 * k;
            open = open || (tx < gx && y < gy && gate[y * gx + tx] != 0u);
        }
        if (__ballot(open) == 0ull) return;  // one wave, no barrier follows
    }
    const size_t tab = (size_t)(uint32_t)b.vi.z;
    uint32_t cur[TPL], beg[TPL];
#pragma unroll
    for (int k = 0; k < TPL; k++) {
        const int y = ty + 8 * k;
        cur[k] = (tx < gx && y < gy) ? a.table[tab + (size_t)(y * gx + tx) * a.C + b.c] : 0u;
        beg[k] = cur[k];
    }
    walk_bin<TPL>(a, b, tx, ty, lane, [&](const bool (&in)[TPL], uint32_t gid) {
#pragma unroll
        for (int k = 0; k < TPL; k++) {
            const uint32_t p = cur[k];
            const uint32_t slot = in[k] ? (p & (LINE - 1)) : (uint32_t)LINE;
            lines[(k * 64 + lane) * LSTRIDE + slot] = gid;
            cur[k] = p + (in[k] ? 1u : 0u);
            // a line fills when slot LINE - 1 was written; such lines leave one
            // at a time, each written by the whole wave
            uint64_t fm = __ballot(slot == LINE - 1);
            while (fm) {
                const int j = __builtin_ctzll(fm);
                fm &= fm - 1ull;
                const uint32_t pj = (uint32_t)__builtin_amdgcn_readlane((int)p, j);
                const uint32_t bj = (uint32_t)__builtin_amdgcn_readlane((int)beg[k], j);
                const uint32_t lo = pj - (LINE - 1);
                flush_line_wave(a, lines + (k * 64 + j) * LSTRIDE, lo > bj ? lo : bj, pj, lane);
            }
        }
    });
    // the last, partly filled line of each tile
#pragma unroll
    for (int k = 0; k < TPL; k++) {
        const uint32_t end = cur[k], lo = end & ~(uint32_t)(LINE - 1);
        const uint32_t from = lo > beg[k] ? lo : beg[k];
        uint64_t fm = __ballot(from < end);
        while (fm) {
            const int j = __builtin_ctzll(fm);
            fm &= fm - 1ull;
            const uint32_t fj = (uint32_t)__builtin_amdgcn_readlane((int)from, j);
            const uint32_t ej = (uint32_t)__builtin_amdgcn_readlane((int)end, j);
            flush_line_wave(a, lines + (k * 64 + j) * LSTRIDE, fj, ej - 1u, lane);
        }
    }
}

// One thread per global tile.
// Ranges are clamped to the pair capacity, so that after an overflow (pairs
// dropped, flag raised) the rasterizers still read inside the buffer.
__global__ __launch_bounds__(256) void bin_ranges_kernel(const uint32_t* __restrict__ scanned,
                                                         const int4* __restrict__ views, int4 view0, int V, int C,
                                                         uint32_t total_tiles, uint32_t capacity,
                                                         uint32_t* __restrict__ ranges) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total_tiles) return;
    int v = 0;
    if (views)
        while (v + 1 < V && (uint32_t)views[v + 1].w <= g) v++;
    const int4 vi = views ? views[v] : view0;
    const size_t t = g - (uint32_t)vi.w;
    const size_t tab = (size_t)(uint32_t)vi.z;
    uint32_t s = scanned[tab + t * C];
    uint32_t e = scanned[tab + (t + 1) * C];
    s = s < capacity ? s : capacity;
    e = e < capacity ? e : capacity;
    if (s == e) s = e = 0u;
    ranges[2 * (size_t)g] = s;
    ranges[2 * (size_t)g + 1] = e;
}

}  // namespace gsb

using namespace gsb;

int bin_blocks(int C, int gx, int gy, int tpl) { return C * ((gx + 7) / 8) * ((gy + 8 * tpl - 1) / (8 * tpl)); }

hipError_t launch_bin_count(const BinArgs& a, int blocks, int tpl, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    const size_t VN = (size_t)a.V * a.N;
    bin_rects_kernel<<<(unsigned)((VN + 255) / 256), 256, 0, s>>>(a);
    if (a.diff_len > 0 && a.diff_len <= DIFF_MAX) {
        bin_count_diff_kernel<<<a.V * a.C, 256, (size_t)a.diff_len * sizeof(int), s>>>(a);
        return hipGetLastError();
    }
    switch (tpl) {
        case 1: bin_count_kernel<1><<<blocks, 64, 0, s>>>(a); break;
        case 2: bin_count_kernel<2><<<blocks, 64, 0, s>>>(a); break;
        case 4: bin_count_kernel<4><<<blocks, 64, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bin_place(const BinArgs& a, int blocks, int tpl, hipStream_t s, int block0, const uint32_t* gate) {
    if (blocks <= 0) return hipSuccess;
    switch (tpl) {
        case 1: bin_place_kernel<1><<<blocks, 64, 0, s>>>(a, block0, gate); break;
        case 2: bin_place_kernel<2><<<blocks, 64, 0, s>>>(a, block0, gate); break;
        case 4: bin_place_kernel<4><<<blocks, 64, 0, s>>>(a, block0, gate); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bin_ranges(const BinArgs& a, const uint32_t* scanned, uint32_t total_tiles, uint32_t* ranges,
                             hipStream_t s) {
    if (total_tiles == 0) return hipSuccess;
    bin_ranges_kernel<<<(total_tiles + 255) / 256, 256, 0, s>>>(scanned, a.views, a.view0, a.V, a.C, total_tiles,
                                                                a.capacity, ranges);
    return hipGetLastError();
}
