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
// The list is cut into chunks of G Gaussians.
//   count:  per chunk, LDS counters (one per tile of the view) count how many
//           of its Gaussians cover each tile; the counts go to a table laid out
//           tile-major, (view, tile, chunk), plus one trailing entry.
//   scan:   one exclusive scan of the table gives, for every (tile, chunk), the
//           output slot of the chunk's first pair in that tile (the tile-major
//           layout makes the tile starts part of the same scan), and the total.
//   place:  each chunk walks its Gaussians again in depth order with the
//           counters set from the scanned table and writes each Gaussian index
//           at counter++ -- its stable rank in the tile.
//   ranges: (start, end) of a tile are the scanned entries of its first chunk
//           and of the next tile's first chunk.
// The four waves of a block split the view's tile rows into bands, so every
// counter has a single owner wave and each tile sees its Gaussians in order.
// HBM traffic: 4 B per pair written once, ~16 B per Gaussian read twice, and the
// table (4 B per tile and chunk) written, scanned and read -- no pass over the
// pairs is repeated.
#include "gs_device.h"

namespace gsb {

constexpr int BIN_THREADS = 256;
constexpr int BIN_WAVES = BIN_THREADS / 64;
constexpr int BIN_STAGE = BIN_THREADS;  // Gaussian rectangles staged per round

template <bool PLACE>
__global__ __launch_bounds__(BIN_THREADS) void bin_chunk_kernel(BinArgs a) {
    extern __shared__ uint32_t cnt[];  // one counter per tile of the view
    __shared__ int4 s_rect[BIN_STAGE];  // x0, y0, x1, y1 (empty when x0 == x1)
    __shared__ uint32_t s_id[BIN_STAGE];

    const int q = blockIdx.x;
    const int v = q / a.C, c = q - v * a.C;
    const int4 vi = a.views ? a.views[v] : a.view0;  // (tiles per row, tile rows, table offset, first tile)
    const int gx = vi.x, gy = vi.y, T = gx * gy;
    const size_t tab = (size_t)(uint32_t)vi.z;
    if (!PLACE && q == 0 && threadIdx.x == 0) a.table[a.table_len - 1] = 0u;
    for (int t = threadIdx.x; t < T; t += BIN_THREADS) cnt[t] = PLACE ? a.table[tab + (size_t)t * a.C + c] : 0u;

    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rb = (gy * w) / BIN_WAVES, re = (gy * (w + 1)) / BIN_WAVES;  // this wave's tile rows
    const size_t seg = (size_t)v * a.N;
    const uint32_t r0 = (uint32_t)c * (uint32_t)a.G;
    const uint32_t r1 = (r0 + (uint32_t)a.G < (uint32_t)a.N) ? r0 + (uint32_t)a.G : (uint32_t)a.N;

    for (uint32_t base = r0; base < r1; base += BIN_STAGE) {
        __syncthreads();  // counters initialised / previous stage consumed
        const uint32_t r = base + threadIdx.x;
        int4 rect = make_int4(0, 0, 0, 0);
        uint32_t id = 0u;
        if (r < r1) {
            id = a.sidx[seg + r];
            const int rad = a.radii[seg + id];
            if (rad > 0) {
                const float2 m = a.means2D[seg + id];
                int mn[2], mx[2];
                gsd::get_rect(m.x, m.y, rad, gx, gy, mn, mx);
                rect = make_int4(mn[0], mn[1], mx[0], mx[1]);
            }
        }
        s_rect[threadIdx.x] = rect;
        s_id[threadIdx.x] = id;
        __syncthreads();
        const int nst = (int)((r1 - base) < (uint32_t)BIN_STAGE ? (r1 - base) : (uint32_t)BIN_STAGE);
        for (int j = 0; j < nst; j++) {
            const int4 R = s_rect[j];
            const int x0 = __builtin_amdgcn_readfirstlane(R.x);
            const int x1 = __builtin_amdgcn_readfirstlane(R.z);
            int y0 = __builtin_amdgcn_readfirstlane(R.y);
            int y1 = __builtin_amdgcn_readfirstlane(R.w);
            y0 = y0 > rb ? y0 : rb;
            y1 = y1 < re ? y1 : re;
            if (x0 >= x1 || y0 >= y1) continue;
            const int wd = x1 - x0;
            const int np = wd * (y1 - y0);
            const uint32_t gid = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_id[j]);
            const float inv = 1.0f / (float)wd;
            for (int e = lane; e < np; e += 64) {
                // row of pair e inside the band rectangle: float estimate, then
                // one correction step (the estimate is off by at most one)
                int yy = (int)((float)e * inv);
                if (yy * wd > e) yy--;
                else if ((yy + 1) * wd <= e) yy++;
                const int t = (y0 + yy) * gx + x0 + (e - yy * wd);
                if (PLACE) {
                    const uint32_t p = atomicAdd(&cnt[t], 1u);
                    if (p < a.capacity) a.out[p] = gid;
                    else *a.overflow = 1u;
                } else {
                    atomicAdd(&cnt[t], 1u);
                }
            }
        }
    }
    if (!PLACE) {
        __syncthreads();
        for (int t = threadIdx.x; t < T; t += BIN_THREADS) a.table[tab + (size_t)t * a.C + c] = cnt[t];
    }
}

// One thread per global tile.
__global__ __launch_bounds__(256) void bin_ranges_kernel(const uint32_t* __restrict__ scanned,
                                                         const int4* __restrict__ views, int4 view0, int V, int C,
                                                         uint32_t total_tiles, uint32_t* __restrict__ ranges) {
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
    if (s == e) s = e = 0u;
    ranges[2 * (size_t)g] = s;
    ranges[2 * (size_t)g + 1] = e;
}

}  // namespace gsb

using namespace gsb;

size_t bin_lds_bytes(int max_tiles) { return (size_t)max_tiles * 4; }

hipError_t launch_bin_count(const BinArgs& a, int max_tiles, hipStream_t s) {
    const int blocks = a.V * a.C;
    if (blocks <= 0) return hipSuccess;
    bin_chunk_kernel<false><<<blocks, BIN_THREADS, bin_lds_bytes(max_tiles), s>>>(a);
    return hipGetLastError();
}

hipError_t launch_bin_place(const BinArgs& a, int max_tiles, hipStream_t s) {
    const int blocks = a.V * a.C;
    if (blocks <= 0) return hipSuccess;
    bin_chunk_kernel<true><<<blocks, BIN_THREADS, bin_lds_bytes(max_tiles), s>>>(a);
    return hipGetLastError();
}

hipError_t launch_bin_ranges(const BinArgs& a, const uint32_t* scanned, uint32_t total_tiles, uint32_t* ranges,
                             hipStream_t s) {
    if (total_tiles == 0) return hipSuccess;
    bin_ranges_kernel<<<(total_tiles + 255) / 256, 256, 0, s>>>(scanned, a.views, a.view0, a.V, a.C, total_tiles,
                                                                ranges);
    return hipGetLastError();
}
