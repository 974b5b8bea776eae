// k_raster.hip -- tile rasterization, forward and backward.
//
// One 256-thread workgroup (four wave64s) per 16x16 tile, one thread per
// pixel.  The tile's depth-sorted Gaussians are staged in LDS in batches of
// BATCH entries (position, conic+opacity, colour); every lane then reads the
// same LDS slot (a broadcast, no bank conflicts).  The forward stops a pixel
// when its transmittance would fall below 1e-4 and the workgroup when all its
// pixels stopped; it records T_final, the contributor count and the list index
// (+1) of the last contributor.
//
// The backward walks the same list back to front from the largest last
// contributor in the tile.  Each pixel recovers T_i = T/(1-alpha) exactly as
// the reference CPU backward (cpu/rasterizer_backward_cpu.cpp:69-134); the
// nine per-Gaussian gradient terms are summed across the wave with lane
// shuffles, across the four waves in LDS, and land in HBM with one float
// atomic per (tile, Gaussian, component).
#include "gs_device.h"

namespace gsr {

using gsd::smin;

template <int BATCH>
__global__ __launch_bounds__(256) void raster_fwd_kernel(RasterArgs a, float* __restrict__ img,
                                                         float* __restrict__ T_final, int* __restrict__ n_contrib,
                                                         uint32_t* __restrict__ last_contrib) {
    __shared__ float2 s_xy[BATCH];
    __shared__ float4 s_co[BATCH];
    __shared__ float4 s_rgb[BATCH];
    const int tile = blockIdx.x;
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int t = threadIdx.x;
    const int px = tx * gsd::TILE + (t & 15), py = ty * gsd::TILE + (t >> 4);
    const bool inside = px < a.W && py < a.H;
    const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
    const float fpx = (float)px, fpy = (float)py;
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    int cnt = 0;
    uint32_t last = 0;
    bool done = !inside;
    for (uint32_t base = start; base < end; base += BATCH) {
        if (__syncthreads_count(!done) == 0) break;
        for (int k = t; k < BATCH; k += 256) {
            uint32_t j = base + k;
            if (j < end) {
                uint32_t id = a.values[(size_t)j * a.vstride];
                s_xy[k] = a.means2D[id];
                s_co[k] = make_float4(a.conics[id * 3], a.conics[id * 3 + 1], a.conics[id * 3 + 2], a.opac[id]);
                s_rgb[k] = make_float4(a.rgb[id * 3], a.rgb[id * 3 + 1], a.rgb[id * 3 + 2], 0.0f);
            }
        }
        __syncthreads();
        const int n = (int)((end - base) < (uint32_t)BATCH ? (end - base) : (uint32_t)BATCH);
        if (!done) {
            for (int k = 0; k < n; k++) {
                const float2 m = s_xy[k];
                const float4 co = s_co[k];
                const float dx = m.x - fpx, dy = m.y - fpy;
                const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
                if (power > 0.0f) continue;
                const float alpha = smin(0.99f, co.w * expf(power));
                if (alpha < 1.0f / 255.0f) continue;
                const float test_T = T * (1.0f - alpha);
                if (test_T < 0.0001f) {
                    done = true;
                    break;
                }
                const float4 c = s_rgb[k];
                C0 += c.x * alpha * T;
                C1 += c.y * alpha * T;
                C2 += c.z * alpha * T;
                T = test_T;
                cnt++;
                last = base + (uint32_t)k + 1u;
            }
        }
    }
    if (inside) {
        const int pix = py * a.W + px;
        img[pix * 3 + 0] = C0 + T * a.bg0;
        img[pix * 3 + 1] = C1 + T * a.bg1;
        img[pix * 3 + 2] = C2 + T * a.bg2;
        if (T_final) T_final[pix] = T;
        if (n_contrib) n_contrib[pix] = cnt;
        if (last_contrib) last_contrib[pix] = last;
    }
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int NG = 9;  // gradient terms per Gaussian: means2D 2, conic 3, rgb 3, opacity 1

template <int BATCH>
__global__ __launch_bounds__(256) void raster_bwd_kernel(RasterArgs a, const float* __restrict__ T_final,
                                                         const uint32_t* __restrict__ last_contrib,
                                                         const float* __restrict__ d_image, GradArgs g) {
    __shared__ float2 s_xy[BATCH];
    __shared__ float4 s_co[BATCH];
    __shared__ float4 s_rgb[BATCH];
    __shared__ uint32_t s_id[BATCH];
    __shared__ float s_g[BATCH][NG];
    __shared__ uint32_t s_maxlast;

    const int tile = blockIdx.x;
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int px = tx * gsd::TILE + (t & 15), py = ty * gsd::TILE + (t >> 4);
    const bool inside = px < a.W && py < a.H;
    const uint32_t start = a.ranges[2 * tile];
    const int pix = py * a.W + px;
    float T = 0.0f, dC0 = 0.0f, dC1 = 0.0f, dC2 = 0.0f;
    uint32_t last = 0;
    if (inside) {
        T = T_final[pix];
        last = last_contrib[pix];
        dC0 = d_image[pix * 3];
        dC1 = d_image[pix * 3 + 1];
        dC2 = d_image[pix * 3 + 2];
    }
    float acc0 = a.bg0, acc1 = a.bg1, acc2 = a.bg2;
    const float fpx = (float)px, fpy = (float)py;
    if (t == 0) s_maxlast = 0;
    __syncthreads();
    if (last > 0) atomicMax(&s_maxlast, last);
    __syncthreads();
    const uint32_t stop = s_maxlast;

    for (uint32_t hi = stop; hi > start;) {
        const uint32_t lo = (hi - start > (uint32_t)BATCH) ? hi - BATCH : start;
        const int n = (int)(hi - lo);
        __syncthreads();
        for (int k = t; k < BATCH; k += 256) {
            if (k < n) {
                uint32_t id = a.values[(size_t)(lo + k) * a.vstride];
                s_id[k] = id;
                s_xy[k] = a.means2D[id];
                s_co[k] = make_float4(a.conics[id * 3], a.conics[id * 3 + 1], a.conics[id * 3 + 2], a.opac[id]);
                s_rgb[k] = make_float4(a.rgb[id * 3], a.rgb[id * 3 + 1], a.rgb[id * 3 + 2], 0.0f);
            }
#pragma unroll
            for (int c = 0; c < NG; c++) s_g[k][c] = 0.0f;
        }
        __syncthreads();
        for (int k = n - 1; k >= 0; k--) {
            const uint32_t j = lo + (uint32_t)k;
            float gv[NG];
#pragma unroll
            for (int c = 0; c < NG; c++) gv[c] = 0.0f;
            bool contrib = false;
            if (j < last) {
                const float2 m = s_xy[k];
                const float4 co = s_co[k];
                const float dx = m.x - fpx, dy = m.y - fpy;
                const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
                if (power <= 0.0f) {
                    const float ep = expf(power);
                    const float alpha = smin(0.99f, co.w * ep);
                    if (alpha >= 1.0f / 255.0f) {
                        contrib = true;
                        const float4 cl = s_rgb[k];
                        const float oma = 1.0f - alpha;
                        const float Ti = (oma > 1e-6f) ? T / oma : 0.0f;
                        gv[5] = alpha * Ti * dC0;
                        gv[6] = alpha * Ti * dC1;
                        gv[7] = alpha * Ti * dC2;
                        float dal = 0.0f;
                        dal += Ti * (cl.x - acc0) * dC0;
                        dal += Ti * (cl.y - acc1) * dC1;
                        dal += Ti * (cl.z - acc2) * dC2;
                        acc0 = alpha * cl.x + oma * acc0;
                        acc1 = alpha * cl.y + oma * acc1;
                        acc2 = alpha * cl.z + oma * acc2;
                        T = Ti;
                        float dpow = 0.0f, dop = 0.0f;
                        if (alpha < 0.99f) {
                            dpow = dal * co.w * ep;
                            dop = dal * ep;
                        }
                        gv[8] = dop;
                        gv[0] = dpow * (-co.x * dx - co.y * dy);
                        gv[1] = dpow * (-co.z * dy - co.y * dx);
                        gv[2] = dpow * (-0.5f * dx * dx);
                        gv[3] = dpow * (-dx * dy);
                        gv[4] = dpow * (-0.5f * dy * dy);
                    }
                }
            }
            if (__ballot(contrib) != 0ull) {
#pragma unroll
                for (int c = 0; c < NG; c++) gv[c] = wave_sum_f(gv[c]);
                if (lane < NG) {
                    float v = gv[0];
#pragma unroll
                    for (int c = 1; c < NG; c++)
                        if (lane == c) v = gv[c];
                    if (v != 0.0f) atomicAdd(&s_g[k][lane], v);
                }
            }
        }
        __syncthreads();
        for (int k = t; k < n; k += 256) {
            const uint32_t id = s_id[k];
            float v;
            v = s_g[k][0]; if (v != 0.0f) unsafeAtomicAdd(&g.d_means2D[id].x, v);
            v = s_g[k][1]; if (v != 0.0f) unsafeAtomicAdd(&g.d_means2D[id].y, v);
            v = s_g[k][2]; if (v != 0.0f) unsafeAtomicAdd(&g.d_conics[id * 3 + 0], v);
            v = s_g[k][3]; if (v != 0.0f) unsafeAtomicAdd(&g.d_conics[id * 3 + 1], v);
            v = s_g[k][4]; if (v != 0.0f) unsafeAtomicAdd(&g.d_conics[id * 3 + 2], v);
            v = s_g[k][5]; if (v != 0.0f) unsafeAtomicAdd(&g.d_rgb[id * 3 + 0], v);
            v = s_g[k][6]; if (v != 0.0f) unsafeAtomicAdd(&g.d_rgb[id * 3 + 1], v);
            v = s_g[k][7]; if (v != 0.0f) unsafeAtomicAdd(&g.d_rgb[id * 3 + 2], v);
            v = s_g[k][8]; if (v != 0.0f) unsafeAtomicAdd(&g.d_op[id], v);
        }
        hi = lo;
    }
}

}  // namespace gsr

using namespace gsr;

hipError_t launch_raster_fwd(const RasterArgs& a, int num_tiles, int batch, float* img, float* T_final, int* n_contrib,
                             uint32_t* last_contrib, hipStream_t s) {
    if (num_tiles <= 0) return hipSuccess;
    switch (batch) {
        case 64: raster_fwd_kernel<64><<<num_tiles, 256, 0, s>>>(a, img, T_final, n_contrib, last_contrib); break;
        case 128: raster_fwd_kernel<128><<<num_tiles, 256, 0, s>>>(a, img, T_final, n_contrib, last_contrib); break;
        case 256: raster_fwd_kernel<256><<<num_tiles, 256, 0, s>>>(a, img, T_final, n_contrib, last_contrib); break;
        case 512: raster_fwd_kernel<512><<<num_tiles, 256, 0, s>>>(a, img, T_final, n_contrib, last_contrib); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_raster_bwd(const RasterArgs& a, int num_tiles, int batch, const float* T_final,
                             const uint32_t* last_contrib, const float* d_image, const GradArgs& g, hipStream_t s) {
    if (num_tiles <= 0) return hipSuccess;
    switch (batch) {
        case 64: raster_bwd_kernel<64><<<num_tiles, 256, 0, s>>>(a, T_final, last_contrib, d_image, g); break;
        case 128: raster_bwd_kernel<128><<<num_tiles, 256, 0, s>>>(a, T_final, last_contrib, d_image, g); break;
        case 256: raster_bwd_kernel<256><<<num_tiles, 256, 0, s>>>(a, T_final, last_contrib, d_image, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
