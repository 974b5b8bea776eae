// gs_density.cpp -- adaptive density control on the trainer's device state
// (DensityController, harmonyos_3dgs/src/density_controller.cpp:9-204, and
// Trainer::step_with_densify, src/trainer.cpp:72-182).
//
// The statistics and the parameters stay on the device.  A densification
// reads back three counts (clones, splits, survivors) to size its arrays; the
// split noise is the reference's own sequence: its LCG and Box-Muller step
// (density_controller.cpp:9-16) run here on the host, with the host's logf /
// cosf / sqrtf that the reference calls, so every split offset is bit for bit
// the reference's.  The k-th split Gaussian (in index order) takes the normals
// 6k .. 6k+5 of the sequence, three per child, as the reference draws them.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <vector>

#include "gs_device.h"
#include "gs_internal.h"
#include "gs_trainer_impl.h"

#define DC_TRY(x)                     \
    do {                              \
        int st_ = (x);                \
        if (st_ != GS_OK) return st_; \
    } while (0)
#define DC_HIP(expr, what)                                    \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return gs_hip_status(e_, what); \
    } while (0)

namespace {

// DensityController::randn (density_controller.cpp:9-16): two LCG steps, the
// top 24 bits of each as a uniform, Box-Muller.
float dc_randn(uint64_t& st) {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    float u1 = (float)((st >> 33) & 0xFFFFFF) / (float)0xFFFFFF;
    if (u1 < 1e-7f) u1 = 1e-7f;
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    const float u2 = (float)((st >> 33) & 0xFFFFFF) / (float)0xFFFFFF;
    return std::sqrt(-2.0f * std::log(u1)) * std::cos(6.28318530718f * u2);
}

int dc_reset(gs_trainer* t) {
    const size_t n = (size_t)t->n;
    DC_TRY(t->dc_accum.ensure(n * 4));
    DC_TRY(t->dc_denom.ensure(n * 4));
    DC_TRY(t->dc_maxr.ensure(n * 4));
    DC_HIP(launch_zero(t->dc_accum.p, n * 4, t->ctx->stream), "zero statistics");
    DC_HIP(launch_zero(t->dc_denom.p, n * 4, t->ctx->stream), "zero statistics");
    DC_HIP(launch_zero(t->dc_maxr.p, n * 4, t->ctx->stream), "zero statistics");
    t->dc_n = t->n;
    return GS_OK;
}

// Reads n uint32 device words (synchronous).
int read_u32(gs_trainer* t, const uint32_t* dev, uint32_t* host, int n) {
    DC_HIP(hipMemcpyAsync(host, dev, (size_t)n * 4, hipMemcpyDeviceToHost, t->ctx->stream), "hipMemcpyAsync");
    DC_HIP(hipStreamSynchronize(t->ctx->stream), "hipStreamSynchronize");
    return GS_OK;
}

}  // namespace

extern "C" int gs_trainer_count(gs_trainer* t, int* n) {
    if (!t || !n) return gs_set_error(GS_ERR_INVALID, "gs_trainer_count: null argument");
    *n = t->n;
    return GS_OK;
}

extern "C" int gs_trainer_density_reset(gs_trainer* t) {
    if (!t) return gs_set_error(GS_ERR_INVALID, "null trainer");
    return dc_reset(t);
}

extern "C" int gs_trainer_density_accumulate(gs_trainer* t, const float* d_means2D, const int* radii) {
    if (!t || (t->n > 0 && !d_means2D)) return gs_set_error(GS_ERR_INVALID, "gs_trainer_density_accumulate: null argument");
    if (reinterpret_cast<uintptr_t>(d_means2D) % 8) return gs_set_error(GS_ERR_INVALID, "d_means2D must be 8-byte aligned");
    if (t->dc_n != t->n) DC_TRY(dc_reset(t));  // accumulate_stats :30
    DC_HIP(launch_density_accumulate(t->n, d_means2D, radii, t->dc_accum.f(), t->dc_denom.f(), t->dc_maxr.i(),
                                     t->ctx->stream),
           "density statistics kernel");
    return GS_OK;
}

extern "C" int gs_trainer_reset_opacity(gs_trainer* t) {
    if (!t) return gs_set_error(GS_ERR_INVALID, "null trainer");
    // inverse_sigmoid(0.01), evaluated as the reference does (std::log of a float)
    const float v = std::log(0.01f / 0.99f);
    DC_HIP(launch_fill(t->n, v, t->param[4].f(), t->ctx->stream), "opacity reset kernel");
    return GS_OK;
}

extern "C" int gs_trainer_densify_and_prune(gs_trainer* t, const gs_densify_cfg* cfg, float extent, int* new_n) {
    if (!t || !cfg) return gs_set_error(GS_ERR_INVALID, "gs_trainer_densify_and_prune: null argument");
    const int N = t->n;
    hipStream_t s = t->ctx->stream;
    if (N == 0) {  // :152
        if (new_n) *new_n = 0;
        return GS_OK;
    }
    if (t->dc_n != N) DC_TRY(dc_reset(t));
    const size_t n = (size_t)N;
    // scratch: clone, split, their ranks, and the scans' partial sums
    const size_t partials = (size_t)scan_partials_needed((uint32_t)n) + 1;
    const size_t words = 4 * n + 8;  // clone, split, rank_c, rank_s, totals
    DC_TRY(t->dc_scratch.ensure(words * 4 + partials * 8 + 256));
    uint32_t* clone = t->dc_scratch.u();
    uint32_t* split = clone + n;
    uint32_t* rank_c = split + n;
    uint32_t* rank_s = rank_c + n;
    uint32_t* totals = rank_s + n;  // [0] clones, [1] splits, [2] survivors
    auto* part = reinterpret_cast<unsigned long long*>(t->dc_scratch.u() + ((words + 63) & ~size_t(63)));
    const float size_threshold = 0.01f * extent;  // percent_dense * extent (:50)
    DC_HIP(launch_density_classify(N, t->dc_accum.f(), t->dc_denom.f(), t->param[1].f(), cfg->grad_threshold,
                                   size_threshold, clone, split, s),
           "density classify kernel");
    DC_HIP(launch_scan_u32_exclusive(clone, (uint32_t)n, rank_c, part, s, totals), "scan");
    DC_HIP(launch_scan_u32_exclusive(split, (uint32_t)n, rank_s, part, s, totals + 1), "scan");
    uint32_t nc_ns[2];
    DC_TRY(read_u32(t, totals, nc_ns, 2));
    const uint32_t nc = nc_ns[0], ns = nc_ns[1];
    const long long m1 = (long long)N + nc + ns;  // after clone and split (:157)
    if (m1 > INT_MAX) return gs_set_error(GS_ERR_CAPACITY, "densification would hold %lld Gaussians", m1);

    // the split noise, three normals per child, children in index order
    std::vector<float> noise((size_t)ns * 6);
    for (float& z : noise) z = dc_randn(t->dc_rng);
    DC_TRY(t->dc_noise.ensure(noise.size() * 4 + 4));
    if (!noise.empty())
        DC_HIP(hipMemcpyAsync(t->dc_noise.p, noise.data(), noise.size() * 4, hipMemcpyHostToDevice, s), "upload noise");

    // the intermediate array's sources, prune flags and final slots
    const size_t M1 = (size_t)m1;
    const size_t partials1 = (size_t)scan_partials_needed((uint32_t)M1) + 1;
    DC_TRY(t->dc_map.ensure(M1 * 16 + partials1 * 8 + 256));
    int2* map = static_cast<int2*>(t->dc_map.p);
    uint32_t* keep = reinterpret_cast<uint32_t*>(map + M1);
    uint32_t* slot = keep + M1;
    auto* part1 = reinterpret_cast<unsigned long long*>(t->dc_map.u() + ((M1 * 4 + 63) & ~size_t(63)));
    DC_HIP(launch_density_map(N, clone, split, rank_c, rank_s, nc, ns, map, s), "density map kernel");
    const float scale_factor = std::log(0.8f * 2.0f);  // :108
    DC_HIP(launch_density_prune((int)M1, N, map, t->param[1].f(), t->param[4].f(), t->dc_maxr.i(), scale_factor,
                                cfg->min_opacity, cfg->max_screen_size, 0.1f * extent, keep, s),
           "density prune kernel");
    DC_HIP(launch_scan_u32_exclusive(keep, (uint32_t)M1, slot, part1, s, totals + 2), "scan");
    uint32_t M = 0;
    DC_TRY(read_u32(t, totals + 2, &M, 1));

    // write the survivors into the second set of arrays, then swap
    const size_t per[5] = {3, 3, 4, (size_t)t->max_coeffs * 3, 1};
    float* dst[5];
    const float* src[5];
    for (int k = 0; k < 5; k++) {
        DC_TRY(t->param_alt[k].ensure((size_t)M * per[k] * 4 + 4));
        dst[k] = t->param_alt[k].f();
        src[k] = t->param[k].f();
    }
    DC_HIP(launch_density_scatter((int)M1, t->max_coeffs, map, keep, slot, src,
                                  static_cast<const float*>(t->dc_noise.p), scale_factor, dst, s),
           "density scatter kernel");
    for (int k = 0; k < 5; k++) std::swap(t->param[k], t->param_alt[k]);
    const bool changed = (int)M != N;
    t->n = (int)M;
    if (changed) {
        // AdamOptimizer::reset (optimizer.cpp:45-58) through trainer.cpp:159-163;
        // gradients sized for the new count
        for (int k = 0; k < 5; k++) {
            const size_t b = t->tensor_elems(k) * 4;
            DC_TRY(t->adam_m[k].ensure(b + 4));
            DC_TRY(t->adam_v[k].ensure(b + 4));
            DC_TRY(t->grad[k].ensure(b + 4));
            DC_HIP(launch_zero(t->adam_m[k].p, b, s), "zero Adam moments");
            DC_HIP(launch_zero(t->adam_v[k].p, b, s), "zero Adam moments");
            DC_HIP(launch_zero(t->grad[k].p, b, s), "zero gradients");
        }
        t->adam_steps = 0;
    }
    for (auto& g : t->graphs) g.release();  // captured with the old arrays
    DC_TRY(dc_reset(t));                    // :192
    DC_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (new_n) *new_n = (int)M;
    return GS_OK;
}

extern "C" int gs_trainer_step_densify(gs_trainer* t, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt,
                                       const gs_train_cfg* tc, const gs_densify_cfg* dc, float extent, int iteration,
                                       gs_step_result* out) {
    if (!t || !dc) return gs_set_error(GS_ERR_INVALID, "gs_trainer_step_densify: null argument");
    const bool densify_now = dc->enabled && iteration >= dc->densify_from_iter && iteration < dc->densify_until_iter;
    if (densify_now && dc->densify_interval <= 0)
        return gs_set_error(GS_ERR_INVALID, "densify_interval must be positive");
    gs_step_result r{};
    DC_TRY(gs_trainer_step(t, cam, cfg, gt, tc, iteration, &r));
    // trainer.cpp:110-146: the screen-space gradients and radii of a forward
    // and raster backward at the parameters the step just updated
    long long K = 0;
    DC_TRY(tr_forward_raster_bwd(t, cam, cfg, gt, &K));
    gs::ViewBufs& v = t->views[0];
    DC_TRY(gs_trainer_density_accumulate(t, v.rgrads().d_means2D, v.radii.i()));
    if (densify_now && iteration % dc->densify_interval == 0) DC_TRY(gs_trainer_densify_and_prune(t, dc, extent, nullptr));
    if (dc->enabled && dc->opacity_reset_interval > 0 && iteration > 0 && iteration % dc->opacity_reset_interval == 0)
        DC_TRY(gs_trainer_reset_opacity(t));
    DC_HIP(hipStreamSynchronize(t->ctx->stream), "hipStreamSynchronize");
    r.n_gaussians = t->n;
    if (out) *out = r;
    return GS_OK;
}
