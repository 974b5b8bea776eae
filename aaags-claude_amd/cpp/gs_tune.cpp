// gs_tune.cpp -- single-view autotuner (gs_autotune).
//
// Each knob only changes the cost of its own stage, so the tuner runs one
// stage at a time over that knob's candidates (coordinate descent), timing
// each candidate with HIP events on the context's stream (candidates
// interleaved, median of `reps` runs each), and keeps the fastest:
//   binning:         tiles per lane {1, 2, 4}, then group size {1024 .. 65536},
//                    then the depth-sort digit width {4 .. 8};
//   raster forward:  LDS batch depth {64, 128, 256, 512};
//   raster backward: LDS batch depth {64, 128, 256};
//   forward:         depth groups placed before the first raster round
//                    {0 (one round), 2, 4, 8, 16, 32}, on whole forward renders.
// Only forward, loss and backward stages run: the trainer's parameters and
// optimizer state are untouched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gs_internal.h"
#include "gs_trainer_impl.h"

#define TU_TRY(x)                     \
    do {                              \
        int st_ = (x);                \
        if (st_ != GS_OK) return st_; \
    } while (0)

namespace {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    int init() {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            return gs_set_error(GS_ERR_HIP, "autotune: hipEventCreate failed");
        return GS_OK;
    }
};

// Device time (ms) of one run of f().
template <typename F>
int time_once(gs_ctx* c, Timer& tm, F&& f, float* ms) {
    if (hipEventRecord(tm.a, c->stream) != hipSuccess) return gs_set_error(GS_ERR_HIP, "hipEventRecord");
    TU_TRY(f());
    if (hipEventRecord(tm.b, c->stream) != hipSuccess) return gs_set_error(GS_ERR_HIP, "hipEventRecord");
    if (hipEventSynchronize(tm.b) != hipSuccess) return gs_set_error(GS_ERR_HIP, "hipEventSynchronize");
    if (hipEventElapsedTime(ms, tm.a, tm.b) != hipSuccess) return gs_set_error(GS_ERR_HIP, "hipEventElapsedTime");
    return GS_OK;
}

// Times every candidate `reps` times, round-robin (one untimed run each
// first), so that slow drifts of clocks or temperature hit all candidates
// alike; keeps the candidate with the smallest median and leaves it set.
template <typename Set, typename Run>
int pick(gs_ctx* c, Timer& tm, int reps, const std::vector<int>& cands, Set&& set, Run&& run, int* best,
         float* best_ms) {
    std::vector<std::vector<float>> t(cands.size());
    for (size_t i = 0; i < cands.size(); i++) {
        TU_TRY(set(cands[i]));
        TU_TRY(run());
    }
    for (int r = 0; r < reps; r++) {
        for (size_t i = 0; i < cands.size(); i++) {
            TU_TRY(set(cands[i]));
            float ms = 0.0f;
            TU_TRY(time_once(c, tm, run, &ms));
            t[i].push_back(ms);
        }
    }
    size_t bi = 0;
    float bm = 0.0f;
    for (size_t i = 0; i < cands.size(); i++) {
        std::sort(t[i].begin(), t[i].end());
        const float med = t[i][t[i].size() / 2];
        if (i == 0 || med < bm) {
            bm = med;
            bi = i;
        }
    }
    TU_TRY(set(cands[bi]));
    *best = cands[bi];
    *best_ms = bm;
    return GS_OK;
}

}  // namespace

extern "C" int gs_autotune(gs_trainer* t, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt, int reps,
                           gs_tuning* out) {
    if (!t || !cam || !cfg || !gt) return gs_set_error(GS_ERR_INVALID, "gs_autotune: null argument");
    if (reps < 1) reps = 1;
    gs_ctx* c = t->ctx;
    // one forward sizes the view and pair buffers and fills the preprocess outputs
    long long K = 0;
    TU_TRY(gs_trainer_render(t, cam, cfg, &K));
    gs::ViewBufs& v = t->views[0];
    gs_preprocessed pre = v.pre();
    const int W = cam->width, H = cam->height;
    const bool direct = (W + 15) / 16 <= GS_BIN_MAX_TILES_XY && (H + 15) / 16 <= GS_BIN_MAX_TILES_XY;
    if (!direct) return gs_set_error(GS_ERR_UNSUPPORTED, "gs_autotune: view too large for direct binning");
    // room for the pairs whatever the knobs (the count does not depend on them)
    TU_TRY(t->keys_sorted.ensure((size_t)(K + K / 8 + 1024) * 4));
    const uint32_t cap = (uint32_t)std::min<size_t>(t->keys_sorted.bytes / 4, 0xFFFFFFFFu);
    Timer tm;
    TU_TRY(tm.init());
    gs_tuning res{};
    long long k2 = 0;
    auto bin = [&]() { return gs_bin_tiles(c, t->n, &pre, W, H, t->keys_sorted.u(), cap, v.ranges.u(), &k2); };
    TU_TRY(pick(c, tm, reps, {1, 2, 4}, [&](int x) { return gs_ctx_set_binning(c, 0, x); }, bin,
                &res.bin_tiles_per_lane, &res.binning_ms));
    TU_TRY(pick(c, tm, reps, {1024, 2048, 4096, 8192, 16384, 32768, 65536}, [&](int x) { return gs_ctx_set_binning(c, x, 0); },
                bin, &res.bin_group, &res.binning_ms));
    TU_TRY(pick(c, tm, reps, {4, 5, 6, 7, 8}, [&](int x) { return gs_ctx_set_tuning(c, x, 0, 0); }, bin,
                &res.radix_bits, &res.binning_ms));
    TU_TRY(bin());  // the pair lists of the chosen setting
    t->pair_stride = 1;
    gs_pixel_state px{v.T_final.f(), v.n_contrib.i(), v.last.u()};
    auto fwd = [&]() {
        return gs_raster_fwd(c, W, H, &pre, t->keys_sorted.u(), v.ranges.u(), cfg->bg_color, v.image.f(), &px);
    };
    TU_TRY(pick(c, tm, reps, {64, 128, 256, 512}, [&](int x) { return gs_ctx_set_tuning(c, 0, x, 0); }, fwd,
                &res.raster_batch, &res.raster_fwd_ms));
    TU_TRY(fwd());
    TU_TRY(gs_l1_loss(c, W, H, v.image.f(), gt, v.d_image.f(), static_cast<double*>(v.loss.p)));
    gs_raster_grads rg = v.rgrads();
    auto bwd = [&]() {
        return gs_raster_bwd(c, W, H, &pre, t->keys_sorted.u(), v.ranges.u(), cfg->bg_color, &px, v.d_image.f(), &rg);
    };
    TU_TRY(pick(c, tm, reps, {64, 128, 256}, [&](int x) { return gs_ctx_set_tuning(c, 0, 0, x); }, bwd,
                &res.raster_bwd_batch, &res.raster_bwd_ms));
    res.persistent_per_cu = c->persistent_per_cu;
    // the two-round forward: groups placed before the first raster round, on
    // the trainer's whole forward with the knobs chosen above
    long long k3 = 0;
    auto render = [&]() { return gs_trainer_render(t, cam, cfg, &k3); };
    TU_TRY(pick(c, tm, reps, {0, 16, 32, 64, 128},
                [&](int x) {
                    c->split_groups = x;
                    return GS_OK;
                },
                render, &res.split_groups, &res.forward_ms));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return gs_set_error(GS_ERR_HIP, "hipStreamSynchronize");
    if (out) *out = res;
    return GS_OK;
}
