// k_invdepth.hip -- Rasterizer::rasterize's inverse-depth output for the 2D
// path (cpu/rasterizer_cpu.cpp:170-172,197): per pixel, the sum of
// alpha * T / depth over its contributors in blend order.
//
// A separate pass over the forward's pixel state, so the forward raster
// keeps its lean body when the output is not asked for: each lane replays
// its pixel's list up to the forward's last contributor (last_contrib, the
// sorted-list index + 1), with the reference's exact alpha (restated expf)
// and the 1/255 and transmittance cuts.  One workgroup per 16x16 tile; the
// tile's entries are staged in LDS 256 at a time.  Latency-bound like the
// forward, and only run on request.
#include "gs_device.h"
#include "gs_internal.h"

namespace gsi {
using namespace gsd;

struct InvEntry {
    float2 xy;
    float ca, cb, cc, op, rdepth;
    float pad;
};

__global__ __launch_bounds__(256) void inv_depth_kernel(int W, int H, int gx, const float2* __restrict__ means2D,
                                                        const float* __restrict__ conics,
                                                        const float* __restrict__ op2d,
                                                        const float* __restrict__ depths,
                                                        const uint32_t* __restrict__ values,
                                                        const uint32_t* __restrict__ ranges,
                                                        const uint32_t* __restrict__ last_contrib,
                                                        float* __restrict__ inv_depth) {
    __shared__ InvEntry sm[256];
    const int tile = blockIdx.x;
    const int tx = tile % gx, ty = tile / gx;
    const int t = threadIdx.x;
    const int px = tx * 16 + (t & 15), py = ty * 16 + (t >> 4);
    const bool inside = px < W && py < H;
    const uint32_t r0 = ranges[2 * tile];
    uint32_t end = inside ? last_contrib[py * W + px] : 0u;  // sorted index + 1, or 0
    if (end < r0) end = r0;
    // the workgroup walks up to its furthest last contributor
    uint32_t wend = end;
    {
        __shared__ uint32_t s_max;
        if (t == 0) s_max = 0u;
        __syncthreads();
        atomicMax(&s_max, end);
        __syncthreads();
        wend = s_max;
    }
    float T = 1.0f, inv = 0.0f;
    const float fx = (float)px, fy = (float)py;
    for (uint32_t base = r0; base < wend; base += 256u) {
        const uint32_t j = base + t;
        if (j < wend) {
            const uint32_t id = values[j];
            InvEntry e;
            e.xy = means2D[id];
            e.ca = conics[id * 3];
            e.cb = conics[id * 3 + 1];
            e.cc = conics[id * 3 + 2];
            e.op = op2d[id];
            e.rdepth = 1.0f / depths[id];
            e.pad = 0.0f;
            sm[t] = e;
        }
        __syncthreads();
        const uint32_t m = (wend - base) < 256u ? (wend - base) : 256u;
        for (uint32_t q = 0; q < m; q++) {
            if (base + q >= end) break;
            const InvEntry e = sm[q];
            const float dx = e.xy.x - fx, dy = e.xy.y - fy;
            const float power = -0.5f * (e.ca * dx * dx + e.cc * dy * dy) - e.cb * dx * dy;
            if (power > 0.0f) continue;
            const float alpha = smin(0.99f, e.op * exp_ref(power));
            if (alpha < 1.0f / 255.0f) continue;
            const float test_T = T * (1.0f - alpha);
            if (test_T < 0.0001f) break;
            inv += e.rdepth * alpha * T;
            T = test_T;
        }
        __syncthreads();
    }
    if (inside) inv_depth[py * W + px] = inv;
}

}  // namespace gsi

extern "C" int gs_raster_inv_depth(gs_ctx* c, int width, int height, const gs_preprocessed* pre,
                                   const uint32_t* values_sorted, const uint32_t* ranges, const gs_pixel_state* px,
                                   float* inv_depth) {
    if (!c) return gs_set_error(GS_ERR_INVALID, "null context");
    if (width <= 0 || height <= 0) return gs_set_error(GS_ERR_INVALID, "bad image size");
    if (!pre || !ranges || !px || !px->last_contrib || !inv_depth)
        return gs_set_error(GS_ERR_INVALID, "null argument (the pixel state's last_contrib is required)");
    if (!pre->means2D || !pre->conics || !pre->opacities || !pre->depths)
        return gs_set_error(GS_ERR_INVALID, "null preprocess array");
    if (reinterpret_cast<uintptr_t>(pre->means2D) % 8) return gs_set_error(GS_ERR_INVALID, "means2D must be 8-byte aligned");
    const int gx = (width + 15) / 16, gy = (height + 15) / 16;
    gsi::inv_depth_kernel<<<gx * gy, 256, 0, c->stream>>>(width, height, gx,
                                                          reinterpret_cast<const float2*>(pre->means2D), pre->conics,
                                                          pre->opacities, pre->depths, values_sorted, ranges,
                                                          px->last_contrib, inv_depth);
    return gs_hip_status(hipGetLastError(), "inv_depth kernel");
}
