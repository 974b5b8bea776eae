// k_density.hip -- adaptive density control on the device
// (DensityController, harmonyos_3dgs/src/density_controller.cpp:21-204).
//
// The parameters stay in HBM; a densification is a handful of passes over the
// Gaussians, with the reference's resulting order reproduced by prefix sums:
//   1. classify: each Gaussian's mean screen-space gradient norm and largest
//      scale decide clone / split (densify_and_clone :47-73, densify_and_split
//      :75-138);
//   2. exclusive scans of the clone and split flags give every Gaussian its
//      place in the intermediate array the reference builds: the originals
//      that do not split (compact keeps their order), then the clones in
//      index order, then two children per split Gaussian in index order;
//   3. prune flags over that intermediate array (:160-186), in its indexing;
//   4. an exclusive scan of the kept flags gives each survivor its final slot
//      (compact, train_types.h:119-136), and one pass writes the new arrays,
//      computing the children's positions from the split noise.
// Every float operation follows the reference's order; expf is exp_ref.
#include "gs_device.h"
#include "gs_internal.h"

namespace gsdc {
using gsd::exp_ref;
using gsd::smax;

// accumulate_stats (:29-41)
__global__ __launch_bounds__(256) void accumulate_kernel(int n, const float2* __restrict__ d_means2D,
                                                         const int* __restrict__ radii, float* __restrict__ accum,
                                                         float* __restrict__ denom, int* __restrict__ max_radii) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 d = d_means2D[i];
    const float norm = sqrtf(d.x * d.x + d.y * d.y);
    accum[i] += norm;
    denom[i] += 1.0f;
    if (radii && radii[i] > max_radii[i]) max_radii[i] = radii[i];
}

__device__ __forceinline__ float max_scale3(const float* ls) {
    float m = 0.0f;
    m = smax(m, exp_ref(ls[0]));
    m = smax(m, exp_ref(ls[1]));
    m = smax(m, exp_ref(ls[2]));
    return m;
}

// the clone and split decisions of the original Gaussians
__global__ __launch_bounds__(256) void classify_kernel(int n, const float* __restrict__ accum,
                                                       const float* __restrict__ denom,
                                                       const float* __restrict__ log_scales, float threshold,
                                                       float size_threshold, uint32_t* __restrict__ clone,
                                                       uint32_t* __restrict__ split) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float avg = denom[i] > 0.0f ? accum[i] / denom[i] : 0.0f;
    uint32_t c = 0u, s = 0u;
    if (!(avg < threshold)) {
        const float ms = max_scale3(log_scales + 3 * (size_t)i);
        c = !(ms > size_threshold) ? 1u : 0u;   // :57
        s = !(ms <= size_threshold) ? 1u : 0u;  // :91
    }
    clone[i] = c;
    split[i] = s;
}

// Source of each element of the intermediate array: .x the original index,
// .y the kind (0 kept original, 1 clone, 2 + q child q) and, for children,
// 4 * the split rank.
__global__ __launch_bounds__(256) void map_kernel(int n, const uint32_t* __restrict__ clone,
                                                  const uint32_t* __restrict__ split,
                                                  const uint32_t* __restrict__ rank_c,
                                                  const uint32_t* __restrict__ rank_s, uint32_t n_clone,
                                                  uint32_t n_split, int2* __restrict__ map) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t base = (uint32_t)n - n_split;
    if (split[i]) {
        const uint32_t k = rank_s[i];
        map[base + n_clone + 2u * k] = make_int2(i, 2 + 4 * (int)k);
        map[base + n_clone + 2u * k + 1u] = make_int2(i, 3 + 4 * (int)k);
    } else {
        map[(uint32_t)i - rank_s[i]] = make_int2(i, 0);
    }
    if (clone[i]) map[base + rank_c[i]] = make_int2(i, 1);
}

// prune decisions over the intermediate array (:160-186); max_radii holds the
// original Gaussians' statistics and is read at the intermediate index, as the
// reference does
__global__ __launch_bounds__(256) void prune_kernel(int m1, int n, const int2* __restrict__ map,
                                                    const float* __restrict__ log_scales,
                                                    const float* __restrict__ opacity, const int* __restrict__ max_radii,
                                                    float scale_factor, float min_opacity, int max_screen_size,
                                                    float world_threshold, uint32_t* __restrict__ keep) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m1) return;
    const int2 e = map[j];
    const int src = e.x;
    const bool child = (e.y & 3) >= 2;
    float ls[3];
#pragma unroll
    for (int c = 0; c < 3; c++) ls[c] = log_scales[3 * (size_t)src + c] - (child ? scale_factor : 0.0f);
    const float op = 1.0f / (1.0f + exp_ref(-opacity[src]));
    bool prune = op < min_opacity;
    if (j < n && max_radii[j] > max_screen_size) prune = true;
    if (max_scale3(ls) > world_threshold) prune = true;
    keep[j] = prune ? 0u : 1u;
}

// the new arrays: survivors of the intermediate array at their compacted slots
__global__ __launch_bounds__(256) void scatter_kernel(int m1, int max_coeffs, const int2* __restrict__ map,
                                                      const uint32_t* __restrict__ keep,
                                                      const uint32_t* __restrict__ slot, const float* __restrict__ pos,
                                                      const float* __restrict__ ls, const float* __restrict__ rot,
                                                      const float* __restrict__ sh, const float* __restrict__ op,
                                                      const float* __restrict__ noise, float scale_factor,
                                                      float* __restrict__ npos, float* __restrict__ nls,
                                                      float* __restrict__ nrot, float* __restrict__ nsh,
                                                      float* __restrict__ nop) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m1 || !keep[j]) return;
    const size_t d = slot[j];
    const int2 e = map[j];
    const size_t s = (size_t)e.x;
    const int kind = e.y & 3;
    const float* p = pos + 3 * s;
    const float* l = ls + 3 * s;
    const float* q = rot + 4 * s;
    if (kind < 2) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            npos[3 * d + c] = p[c];
            nls[3 * d + c] = l[c];
        }
    } else {
        // densify_and_split :87-131: the child's offset ~ R (N(0,1) * scale)
        const int k = e.y >> 2, child = kind - 2;
        float scale[3], r4[4];
#pragma unroll
        for (int c = 0; c < 3; c++) scale[c] = exp_ref(l[c]);
#pragma unroll
        for (int c = 0; c < 4; c++) r4[c] = q[c];
        const float qlen = sqrtf(r4[0] * r4[0] + r4[1] * r4[1] + r4[2] * r4[2] + r4[3] * r4[3]);
        if (qlen > 1e-8f) {
#pragma unroll
            for (int c = 0; c < 4; c++) r4[c] /= qlen;
        }
        const float r = r4[0], x = r4[1], y = r4[2], z = r4[3];
        const float R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + r * z),     2 * (x * z - r * y),
                            2 * (x * y - r * z),     1 - 2 * (x * x + z * z), 2 * (y * z + r * x),
                            2 * (x * z + r * y),     2 * (y * z - r * x),     1 - 2 * (x * x + y * y)};
        const float* nz = noise + 6 * (size_t)k + 3 * child;
        const float sample[3] = {nz[0] * scale[0], nz[1] * scale[1], nz[2] * scale[2]};
#pragma unroll
        for (int row = 0; row < 3; row++) {
            float o = 0.0f;
#pragma unroll
            for (int col = 0; col < 3; col++) o += R[row * 3 + col] * sample[col];
            npos[3 * d + row] = p[row] + o;
            nls[3 * d + row] = l[row] - scale_factor;
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) nrot[4 * d + c] = q[c];
    const int k3 = max_coeffs * 3;
    for (int c = 0; c < k3; c++) nsh[(size_t)k3 * d + c] = sh[(size_t)k3 * s + c];
    nop[d] = op[s];
}

__global__ __launch_bounds__(256) void fill_kernel(int n, float v, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = v;
}

}  // namespace gsdc

static unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_density_accumulate(int n, const float* d_means2D, const int* radii, float* accum, float* denom,
                                     int* max_radii, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gsdc::accumulate_kernel<<<blocks_for(n), 256, 0, s>>>(n, reinterpret_cast<const float2*>(d_means2D), radii, accum,
                                                          denom, max_radii);
    return hipGetLastError();
}

hipError_t launch_density_classify(int n, const float* accum, const float* denom, const float* log_scales,
                                   float threshold, float size_threshold, uint32_t* clone, uint32_t* split,
                                   hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gsdc::classify_kernel<<<blocks_for(n), 256, 0, s>>>(n, accum, denom, log_scales, threshold, size_threshold, clone,
                                                        split);
    return hipGetLastError();
}

hipError_t launch_density_map(int n, const uint32_t* clone, const uint32_t* split, const uint32_t* rank_c,
                              const uint32_t* rank_s, uint32_t n_clone, uint32_t n_split, int2* map, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gsdc::map_kernel<<<blocks_for(n), 256, 0, s>>>(n, clone, split, rank_c, rank_s, n_clone, n_split, map);
    return hipGetLastError();
}

hipError_t launch_density_prune(int m1, int n, const int2* map, const float* log_scales, const float* opacity,
                                const int* max_radii, float scale_factor, float min_opacity, int max_screen_size,
                                float world_threshold, uint32_t* keep, hipStream_t s) {
    if (m1 <= 0) return hipSuccess;
    gsdc::prune_kernel<<<blocks_for(m1), 256, 0, s>>>(m1, n, map, log_scales, opacity, max_radii, scale_factor,
                                                      min_opacity, max_screen_size, world_threshold, keep);
    return hipGetLastError();
}

hipError_t launch_density_scatter(int m1, int max_coeffs, const int2* map, const uint32_t* keep, const uint32_t* slot,
                                  const float* const src[5], const float* noise, float scale_factor,
                                  float* const dst[5], hipStream_t s) {
    if (m1 <= 0) return hipSuccess;
    gsdc::scatter_kernel<<<blocks_for(m1), 256, 0, s>>>(m1, max_coeffs, map, keep, slot, src[0], src[1], src[2], src[3],
                                                        src[4], noise, scale_factor, dst[0], dst[1], dst[2], dst[3],
                                                        dst[4]);
    return hipGetLastError();
}

hipError_t launch_fill(int n, float v, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gsdc::fill_kernel<<<blocks_for(n), 256, 0, s>>>(n, v, out);
    return hipGetLastError();
}
