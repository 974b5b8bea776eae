// k_work.hip -- the algorithmic work of the 2D rasterizers, counted on the
// device, for their roofline: how many pixel-Gaussian evaluations the
// reference's loops perform on a given frame (cpu/rasterizer_cpu.cpp:141-176,
// cpu/rasterizer_backward_cpu.cpp:34-134).
//
//   counts[0]  forward evaluations: per pixel, the entries of its tile's list
//              that the reference's forward loop examines, up to and including
//              the one that stops it (transmittance below 1e-4), or the whole
//              list when nothing stops it;
//   counts[1]  backward evaluations: per pixel, the entries from the start of
//              its tile's list to its last contributor (the backward must
//              evaluate each of them to know whether it contributed);
//   counts[2]  contributors: the sum of n_contrib (the entries that get the
//              full gradient body).
//
// A replay of each pixel's list with the reference's exact alpha (restated
// expf), one workgroup per 16x16 tile, entries staged in LDS 256 at a time;
// one 64-bit atomic per workgroup and count.  It is a measurement pass, run
// outside any timed region.
#include "gs_device.h"
#include "gs_internal.h"

namespace gsw {
using namespace gsd;

struct WorkEntry {
    float2 xy;
    float ca, cb, cc, op;
};

__global__ __launch_bounds__(256) void raster_work_kernel(int W, int H, int gx, const float2* __restrict__ means2D,
                                                          const float* __restrict__ conics,
                                                          const float* __restrict__ op2d,
                                                          const uint32_t* __restrict__ values,
                                                          const uint32_t* __restrict__ ranges,
                                                          const int* __restrict__ n_contrib,
                                                          const uint32_t* __restrict__ last_contrib,
                                                          unsigned long long* __restrict__ counts) {
    __shared__ WorkEntry sm[256];
    __shared__ unsigned long long s_sum[3];
    const int tile = blockIdx.x;
    const int tx = tile % gx, ty = tile / gx;
    const int t = threadIdx.x;
    const int px = tx * 16 + (t & 15), py = ty * 16 + (t >> 4);
    const bool inside = px < W && py < H;
    const uint32_t r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
    if (t < 3) s_sum[t] = 0ull;
    unsigned long long e_bwd = 0ull, contrib = 0ull;
    if (inside) {
        const uint32_t last = last_contrib[py * W + px];  // sorted index + 1, or 0
        e_bwd = last > r0 ? (unsigned long long)(last - r0) : 0ull;
        contrib = (unsigned long long)n_contrib[py * W + px];
    }
    float T = 1.0f;
    bool done = !inside;
    unsigned long long e_fwd = 0ull;
    const float fx = (float)px, fy = (float)py;
    for (uint32_t base = r0; base < r1; base += 256u) {
        if (__syncthreads_count(done ? 0 : 1) == 0) break;
        const uint32_t j = base + t;
        if (j < r1) {
            const uint32_t id = values[j];
            WorkEntry e;
            e.xy = means2D[id];
            e.ca = conics[id * 3];
            e.cb = conics[id * 3 + 1];
            e.cc = conics[id * 3 + 2];
            e.op = op2d[id];
            sm[t] = e;
        }
        __syncthreads();
        const uint32_t m = (r1 - base) < 256u ? (r1 - base) : 256u;
        for (uint32_t q = 0; q < m && !done; q++) {
            const WorkEntry e = sm[q];
            e_fwd++;
            const float dx = e.xy.x - fx, dy = e.xy.y - fy;
            const float power = -0.5f * (e.ca * dx * dx + e.cc * dy * dy) - e.cb * dx * dy;
            if (power > 0.0f) continue;
            const float alpha = smin(0.99f, e.op * exp_ref(power));
            if (alpha < 1.0f / 255.0f) continue;
            const float test_T = T * (1.0f - alpha);
            if (test_T < 0.0001f) {
                done = true;
                break;
            }
            T = test_T;
        }
        __syncthreads();
    }
    atomicAdd(&s_sum[0], e_fwd);
    atomicAdd(&s_sum[1], e_bwd);
    atomicAdd(&s_sum[2], contrib);
    __syncthreads();
    if (t < 3) atomicAdd(&counts[t], s_sum[t]);
}

}  // namespace gsw

extern "C" int gs_raster_work(gs_ctx* c, int width, int height, const gs_preprocessed* pre,
                              const uint32_t* values_sorted, const uint32_t* ranges, const gs_pixel_state* px,
                              unsigned long long* counts) {
    if (!c) return gs_set_error(GS_ERR_INVALID, "null context");
    if (width <= 0 || height <= 0) return gs_set_error(GS_ERR_INVALID, "bad image size");
    if (!pre || !ranges || !px || !px->last_contrib || !px->n_contrib || !counts)
        return gs_set_error(GS_ERR_INVALID, "null argument (the pixel state's n_contrib and last_contrib are required)");
    if (!pre->means2D || !pre->conics || !pre->opacities) return gs_set_error(GS_ERR_INVALID, "null preprocess array");
    if (reinterpret_cast<uintptr_t>(pre->means2D) % 8) return gs_set_error(GS_ERR_INVALID, "means2D must be 8-byte aligned");
    hipError_t e = launch_zero(counts, 3 * sizeof(unsigned long long), c->stream);
    if (e != hipSuccess) return gs_hip_status(e, "zero counts");
    const int gx = (width + 15) / 16, gy = (height + 15) / 16;
    gsw::raster_work_kernel<<<gx * gy, 256, 0, c->stream>>>(width, height, gx,
                                                            reinterpret_cast<const float2*>(pre->means2D),
                                                            pre->conics, pre->opacities, values_sorted, ranges,
                                                            px->n_contrib, px->last_contrib, counts);
    return gs_hip_status(hipGetLastError(), "raster work kernel");
}
