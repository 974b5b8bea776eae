// gs_trainer.cpp -- native single-view training driver over the C ABI.
//
// Mirrors Trainer::step (harmonyos_3dgs/src/trainer.cpp:7-70): activation is
// fused into the preprocess kernel, the forward keeps T_final / contributor
// state on the device, the backward and the optimizer run on the same stream,
// and the only host round trips are the pair-count readback used to size the
// pair buffers (as TileBinnerGPU::bin does, gpu/tile_binner_gpu.cpp:292-316)
// and the loss readback of StepResult.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "gs_device.h"
#include "gs_internal.h"
#include "gs_trainer_impl.h"

gsd::CamArgs make_cam_args(const gs_camera* cam, const gs_render_cfg* cfg);

using gs::DevBuf;

namespace gs {

// Grows the buffer to at least `need` bytes; always leaves an allocation
// (256 bytes at least), so that buffers sized by an empty scene or view are
// valid, non-null pointers for the argument checks.
int DevBuf::ensure(size_t need) {
    if (p && need <= bytes) return GS_OK;
    if (p) {
        hipError_t e = hipFree(p);
        if (e != hipSuccess) return gs_hip_status(e, "hipFree");
        p = nullptr;
        bytes = 0;
    }
    size_t nb = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(&p, nb);
    if (e != hipSuccess) return gs_hip_status(e, "hipMalloc");
    bytes = nb;
    return GS_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

}  // namespace gs

#define TR_TRY(x)                 \
    do {                          \
        int st_ = (x);            \
        if (st_ != GS_OK) return st_; \
    } while (0)

static int tr_alloc_view(gs_trainer* t, gs::ViewBufs& v, int W, int H) {
    size_t n = (size_t)t->n;
    size_t npix = (size_t)W * H;
    int tiles = ((W + 15) / 16) * ((H + 15) / 16);
    TR_TRY(v.means2D.ensure(n * 8));
    TR_TRY(v.depths.ensure(n * 4));
    TR_TRY(v.conics.ensure(n * 12));
    TR_TRY(v.op2d.ensure(n * 4));
    TR_TRY(v.rgb.ensure(n * 12));
    TR_TRY(v.radii.ensure(n * 4));
    TR_TRY(v.tiles.ensure(n * 4));
    TR_TRY(v.offsets.ensure(n * 4 + 4));
    TR_TRY(v.ranges.ensure((size_t)tiles * 8));
    TR_TRY(v.image.ensure(npix * 12));
    TR_TRY(v.d_image.ensure(npix * 12));
    TR_TRY(v.T_final.ensure(npix * 4));
    TR_TRY(v.n_contrib.ensure(npix * 4));
    TR_TRY(v.last.ensure(npix * 4));
    TR_TRY(v.loss.ensure(8));
    TR_TRY(v.rg.ensure(n * 36));
    v.rg_n = n;
    return GS_OK;
}

hipError_t gs::StageTimer::collect() {
    for (int s = 0; s < GS_NUM_STAGES; s++) {
        if (!on(s)) continue;
        float ms = 0.0f;
        hipError_t e = hipEventElapsedTime(&ms, beg[s], end[s]);
        if (e != hipSuccess) return e;
        total_ms[s] += ms;
    }
    steps++;
    return hipSuccess;
}

void gs::StageTimer::release() {
    for (int s = 0; s < GS_NUM_STAGES; s++) {
        if (beg[s]) (void)hipEventDestroy(beg[s]);
        if (end[s]) (void)hipEventDestroy(end[s]);
        beg[s] = end[s] = nullptr;
    }
    mask = 0;
}

// Event mark at the start (begin) or end of a timed stage.
static int tr_mark(gs_trainer* t, int stage, bool begin) {
    hipStream_t st = t->ctx->stream;
    hipError_t e = begin ? t->timer.mark_begin(stage, st) : t->timer.mark_end(stage, st);
    return e == hipSuccess ? GS_OK : gs_hip_status(e, "hipEventRecord(stage timer)");
}

gs_preprocessed gs::ViewBufs::pre() {
    gs_preprocessed p;
    p.means2D = means2D.f();
    p.depths = depths.f();
    p.conics = conics.f();
    p.opacities = op2d.f();
    p.rgb = rgb.f();
    p.radii = radii.i();
    p.tiles_touched = tiles.i();
    return p;
}

gs_raster_grads gs::ViewBufs::rgrads() {
    gs_raster_grads g;
    float* base = rg.f();
    g.d_means2D = base;
    g.d_conics = base + 2 * rg_n;
    g.d_rgb = base + 5 * rg_n;
    g.d_opacities_2d = base + 8 * rg_n;
    return g;
}

gs_gaussians gs_trainer::gaussians() const {
    gs_gaussians g;
    g.count = n;
    g.sh_degree = sh_degree;
    g.max_coeffs = max_coeffs;
    g.raw = 1;
    g.positions = static_cast<const float*>(param[0].p);
    g.scales = static_cast<const float*>(param[1].p);
    g.rotations = static_cast<const float*>(param[2].p);
    g.sh_coeffs = static_cast<const float*>(param[3].p);
    g.opacities = static_cast<const float*>(param[4].p);
    return g;
}

size_t gs_trainer::tensor_elems(int k) const {
    const size_t per[5] = {3, 3, 4, (size_t)max_coeffs * 3, 1};
    return (size_t)n * per[k];
}

extern "C" int gs_trainer_create(gs_ctx* ctx, int n, int sh_degree, int max_coeffs, gs_trainer** out) {
    if (!ctx || !out) return gs_set_error(GS_ERR_INVALID, "gs_trainer_create: null argument");
    if (n < 0) return gs_set_error(GS_ERR_INVALID, "gs_trainer_create: negative count");
    if (!(max_coeffs == 1 || max_coeffs == 4 || max_coeffs == 9 || max_coeffs == 16) || sh_degree < 0 ||
        (sh_degree + 1) * (sh_degree + 1) > max_coeffs)
        return gs_set_error(GS_ERR_INVALID, "gs_trainer_create: bad SH layout (degree %d, %d coefficients)", sh_degree,
                            max_coeffs);
    gs_trainer* t = new gs_trainer();
    t->ctx = ctx;
    t->n = n;
    t->sh_degree = sh_degree;
    t->max_coeffs = max_coeffs;
    for (int k = 0; k < 5; k++) {
        size_t b = t->tensor_elems(k) * 4;
        int st;
        if ((st = t->param[k].ensure(b)) || (st = t->grad[k].ensure(b)) || (st = t->adam_m[k].ensure(b)) ||
            (st = t->adam_v[k].ensure(b))) {
            gs_trainer_destroy(t);
            return st;
        }
        gs_zero(ctx, t->adam_m[k].p, b);
        gs_zero(ctx, t->adam_v[k].p, b);
        gs_zero(ctx, t->grad[k].p, b);
    }
    *out = t;
    return GS_OK;
}

gs_trainer::~gs_trainer() {
    for (int k = 0; k < 5; k++) {
        param[k].release();
        grad[k].release();
        adam_m[k].release();
        adam_v[k].release();
        param_alt[k].release();
    }
    gs::DevBuf* dc[] = {&dc_accum, &dc_denom, &dc_maxr, &dc_scratch, &dc_map, &dc_noise};
    for (gs::DevBuf* b : dc) b->release();
    for (auto& v : views) v.release_all();
    keys.release();
    vals.release();
    keys_sorted.release();
    vals_sorted.release();
    timer.release();
    if (k_event) (void)hipEventDestroy(k_event);
    if (k_pinned) (void)hipHostFree(k_pinned);
    for (auto& g : graphs) g.release();
}

void gs::ViewBufs::release_all() {
    DevBuf* all[] = {&means2D, &depths, &conics, &op2d, &rgb, &radii, &tiles, &offsets, &ranges, &image, &d_image,
                     &T_final, &n_contrib, &last, &loss, &rg, &split_flags, &split_rc};
    for (DevBuf* b : all) b->release();
}

extern "C" int gs_trainer_destroy(gs_trainer* t) {
    if (!t) return GS_OK;
    if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
    delete t;
    return GS_OK;
}

extern "C" int gs_trainer_set_params(gs_trainer* t, const float* pos, const float* ls, const float* rot,
                                     const float* sh, const float* op, int from_host) {
    if (!t) return gs_set_error(GS_ERR_INVALID, "null trainer");
    const float* src[5] = {pos, ls, rot, sh, op};
    for (int k = 0; k < 5; k++) {
        if (!src[k]) return gs_set_error(GS_ERR_INVALID, "gs_trainer_set_params: null array %d", k);
        hipError_t e = hipMemcpyAsync(t->param[k].p, src[k], t->tensor_elems(k) * 4,
                                      from_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, t->ctx->stream);
        if (e != hipSuccess) return gs_hip_status(e, "hipMemcpyAsync(params)");
    }
    hipError_t e = hipStreamSynchronize(t->ctx->stream);
    if (e != hipSuccess) return gs_hip_status(e, "hipStreamSynchronize");
    return GS_OK;
}

extern "C" int gs_trainer_get_params(gs_trainer* t, float* pos, float* ls, float* rot, float* sh, float* op,
                                     int to_host) {
    if (!t) return gs_set_error(GS_ERR_INVALID, "null trainer");
    float* dst[5] = {pos, ls, rot, sh, op};
    for (int k = 0; k < 5; k++) {
        if (!dst[k]) continue;
        hipError_t e = hipMemcpyAsync(dst[k], t->param[k].p, t->tensor_elems(k) * 4,
                                      to_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, t->ctx->stream);
        if (e != hipSuccess) return gs_hip_status(e, "hipMemcpyAsync(params)");
    }
    hipError_t e = hipStreamSynchronize(t->ctx->stream);
    if (e != hipSuccess) return gs_hip_status(e, "hipStreamSynchronize");
    return GS_OK;
}

extern "C" void* gs_trainer_buffer(gs_trainer* t, const char* which) {
    if (!t || !which) return nullptr;
    std::string w(which);
    const char* pn[5] = {"pos", "scale", "rot", "sh", "op"};
    for (int k = 0; k < 5; k++) {
        if (w == std::string("param_") + pn[k]) return t->param[k].p;
        if (w == std::string("grad_") + pn[k]) return t->grad[k].p;
        if (w == std::string("adam_m_") + pn[k]) return t->adam_m[k].p;
        if (w == std::string("adam_v_") + pn[k]) return t->adam_v[k].p;
    }
    if (t->views.empty()) return nullptr;
    gs::ViewBufs& v = t->views[0];
    if (w == "image") return v.image.p;
    if (w == "d_image") return v.d_image.p;
    if (w == "T_final") return v.T_final.p;
    if (w == "n_contrib") return v.n_contrib.p;
    if (w == "last_contrib") return v.last.p;
    if (w == "means2D") return v.means2D.p;
    if (w == "depths") return v.depths.p;
    if (w == "conics") return v.conics.p;
    if (w == "opacities_2d") return v.op2d.p;
    if (w == "rgb") return v.rgb.p;
    if (w == "radii") return v.radii.p;
    if (w == "tiles_touched") return v.tiles.p;
    if (w == "ranges") return v.ranges.p;
    if (w == "pairs") return t->keys_sorted.p;  // uint32 indices (pair_stride 1) or 64-bit words (2)
    if (w == "loss") return v.loss.p;
    return nullptr;
}

// Binning of one view into the trainer's pair buffer: direct placement
// (gs_bin_tiles, 4-byte Gaussian indices) when the view has at most
// GS_BIN_MAX_TILES_XY tiles per side, else the depth-ordered word sort
// (gs_bin_sorted_words, 8-byte words).  The buffer grows when a view needs more
// pairs than it holds.  Sets t->pair_stride (uint32 entries per pair).
static int tr_bin(gs_trainer* t, gs::ViewBufs& v, const gs_camera* cam, const gs_preprocessed& pre, long long* K) {
    gs_ctx* c = t->ctx;
    const bool words = (cam->width + 15) / 16 > GS_BIN_MAX_TILES_XY || (cam->height + 15) / 16 > GS_BIN_MAX_TILES_XY;
    const size_t per = words ? 8 : 4;
    for (int attempt = 0; attempt < 2; attempt++) {
        const uint32_t cap = (uint32_t)std::min<size_t>(t->keys_sorted.bytes / per, 0xFFFFFFFFu);
        int st = words ? gs_bin_sorted_words(c, t->n, &pre, cam->width, cam->height,
                                             static_cast<uint64_t*>(t->keys_sorted.p), cap, v.ranges.u(), K)
                       : gs_bin_tiles_impl(c, t->n, &pre, cam->width, cam->height, t->keys_sorted.u(), cap,
                                           v.ranges.u(), K, nullptr, nullptr, c->split_groups);
        if (st == GS_ERR_CAPACITY && *K > (long long)cap && attempt == 0) {
            TR_TRY(t->keys_sorted.ensure((size_t)(*K + *K / 8 + 1024) * per));
            continue;
        }
        TR_TRY(st);
        t->pair_stride = words ? 2 : 1;
        return GS_OK;
    }
    return gs_set_error(GS_ERR_CAPACITY, "trainer binning: pair buffer did not grow");
}

// Forward of one view into view buffers.  When the direct-placement path has a
// pair buffer already (a previous step sized it), binning does not wait for
// the pair count: it is copied to pinned memory and *pending is set, and the
// caller checks it with tr_pairs() before the results are used for an update;
// otherwise the count is read back first and *pairs is set.
static int tr_forward(gs_trainer* t, gs::ViewBufs& v, const gs_camera* cam, const gs_render_cfg* cfg,
                      long long* pairs, bool* pending) {
    gs_ctx* c = t->ctx;
    TR_TRY(tr_alloc_view(t, v, cam->width, cam->height));
    gs_gaussians g = t->gaussians();
    gs_preprocessed pre = v.pre();
    TR_TRY(tr_mark(t, GS_STAGE_PREPROCESS, true));
    TR_TRY(gs_preprocess_fwd(c, &g, cam, cfg, &pre));
    TR_TRY(tr_mark(t, GS_STAGE_PREPROCESS, false));
    long long K = 0;
    *pending = false;
    TR_TRY(tr_mark(t, GS_STAGE_BINNING, true));
    const bool direct = (cam->width + 15) / 16 <= GS_BIN_MAX_TILES_XY && (cam->height + 15) / 16 <= GS_BIN_MAX_TILES_XY;
    if (direct && t->keys_sorted.bytes >= 4 && t->allow_async) {
        if (!t->k_pinned) {
            // mapped + coherent: the scan kernel stores the count straight into it
            hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&t->k_pinned), 8,
                                         hipHostMallocMapped | hipHostMallocCoherent);
            if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&t->k_pinned_dev), t->k_pinned, 0);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&t->k_event, hipEventDisableTiming);
            if (e != hipSuccess) return gs_hip_status(e, "pinned pair count");
        }
        const uint32_t cap = (uint32_t)std::min<size_t>(t->keys_sorted.bytes / 4, 0xFFFFFFFFu);
        TR_TRY(gs_bin_tiles_impl(c, t->n, &pre, cam->width, cam->height, t->keys_sorted.u(), cap, v.ranges.u(), nullptr,
                                 t->k_pinned_dev, t->k_event, c->split_groups));
        t->pair_stride = 1;
        *pending = true;
    } else {
        TR_TRY(tr_bin(t, v, cam, pre, &K));
    }
    TR_TRY(tr_mark(t, GS_STAGE_BINNING, false));
    gs_pixel_state px{v.T_final.f(), v.n_contrib.i(), v.last.u()};
    TR_TRY(tr_mark(t, GS_STAGE_RASTER_FWD, true));
    gsr::RasterSplit sp;
    TR_TRY(gs_bin_split_info(c, &sp, nullptr));
    if (t->pair_stride == 1 && sp.round == 1) {
        // two rounds: render what the binning placed (the nearest depth groups),
        // place the rest for the bins with unfinished tiles, resume those tiles
        const size_t tiles = (size_t)((cam->width + 15) / 16) * ((cam->height + 15) / 16);
        TR_TRY(v.split_flags.ensure(tiles * 4));
        TR_TRY(v.split_rc.ensure((size_t)cam->width * cam->height * 12));
        sp.tile_flag = v.split_flags.u();  // written by round 1 for every tile
        sp.rc = v.split_rc.f();
        TR_TRY(gs_raster_fwd_split(c, cam->width, cam->height, &pre, t->keys_sorted.u(), v.ranges.u(),
                                   cfg->bg_color, v.image.f(), &px, &sp));
        TR_TRY(gs_bin_place_rest(c, sp.tile_flag));
        sp.round = 2;
        TR_TRY(gs_raster_fwd_split(c, cam->width, cam->height, &pre, t->keys_sorted.u(), v.ranges.u(),
                                   cfg->bg_color, v.image.f(), &px, &sp));
    } else if (t->pair_stride == 1)
        TR_TRY(gs_raster_fwd(c, cam->width, cam->height, &pre, t->keys_sorted.u(), v.ranges.u(), cfg->bg_color,
                             v.image.f(), &px));
    else
        TR_TRY(gs_raster_fwd_words(c, cam->width, cam->height, &pre, static_cast<const uint64_t*>(t->keys_sorted.p),
                                   v.ranges.u(), cfg->bg_color, v.image.f(), &px));
    TR_TRY(tr_mark(t, GS_STAGE_RASTER_FWD, false));
    *pairs = K;
    return GS_OK;
}

// The pair count of a pending binning (waits for its copy only); *fits is
// false when the pairs overflowed the buffer, which is then grown.
static int tr_pairs(gs_trainer* t, long long* K, bool* fits) {
    hipError_t e = hipEventSynchronize(t->k_event);
    if (e != hipSuccess) return gs_hip_status(e, "hipEventSynchronize");
    *K = (long long)*static_cast<volatile unsigned long long*>(t->k_pinned);
    const size_t cap = t->keys_sorted.bytes / 4;
    *fits = (size_t)*K <= cap;
    if (!*fits) {
        // the dropped pairs raised the context's overflow flag: clear it, grow
        void* sm = nullptr;
        TR_TRY(gs_small_get(t->ctx, &sm));
        e = launch_zero(static_cast<char*>(sm) + GS_SMALL_OVERFLOW, 4, t->ctx->stream);
        if (e != hipSuccess) return gs_hip_status(e, "zero kernel");
        TR_TRY(t->keys_sorted.ensure((size_t)(*K + *K / 8 + 1024) * 4));
    }
    return GS_OK;
}

static int tr_optimizer(gs_trainer* t, const gs_train_cfg* tc, int iteration) {
    gs_ctx* c = t->ctx;
    float lr_pos = gs_lr_schedule(tc->lr_position_init, tc->lr_position_final, iteration, tc->max_steps);
    const float lrs[5] = {lr_pos, tc->lr_scaling, tc->lr_rotation, tc->lr_feature, tc->lr_opacity};
    float* P[5];
    const float* G[5];
    float* M[5];
    float* V[5];
    for (int k = 0; k < 5; k++) {
        P[k] = t->param[k].f();
        G[k] = t->grad[k].f();
        M[k] = t->adam_m[k].f();
        V[k] = t->adam_v[k].f();
    }
    if (tc->use_adam) {
        t->adam_steps++;  // AdamOptimizer::step_count_ (optimizer.cpp:72)
        gs_adam_hparams hp;
        for (int k = 0; k < 5; k++) hp.lr[k] = lrs[k];
        hp.beta1 = tc->adam_beta1;
        hp.beta2 = tc->adam_beta2;
        hp.eps = tc->adam_eps;
        gs_adam_bias_corrections(hp.beta1, hp.beta2, t->adam_steps, &hp.bc1, &hp.bc2);
        return gs_adam_step(c, t->n, t->max_coeffs, P, G, M, V, &hp);
    }
    return gs_sgd_step(c, t->n, t->max_coeffs, P, G, lrs);
}

// Preprocess backward with Adam applied in place from the gradients held in
// registers (the multi-view kernel with one view): the same arithmetic as
// gs_preprocess_bwd followed by gs_adam_step, without writing the gradients.
static int tr_fused_backward_adam(gs_trainer* t, const gs_camera* cam, const gs_render_cfg* cfg,
                                  const gs_train_cfg* tc, int iteration, const int* radii, const gs_raster_grads& rg) {
    gs_ctx* c = t->ctx;
    const gsd::CamArgs ca = make_cam_args(cam, cfg);
    gso::AdamDev h{};
    h.lr[0] = gs_lr_schedule(tc->lr_position_init, tc->lr_position_final, iteration, tc->max_steps);
    h.lr[1] = tc->lr_scaling;
    h.lr[2] = tc->lr_rotation;
    h.lr[3] = tc->lr_feature;
    h.lr[4] = tc->lr_opacity;
    h.b1 = tc->adam_beta1;
    h.b2 = tc->adam_beta2;
    h.eps = tc->adam_eps;
    t->adam_steps++;  // AdamOptimizer::step_count_ (optimizer.cpp:72)
    gs_adam_bias_corrections(h.b1, h.b2, t->adam_steps, &h.bc1, &h.bc2);
    float* P[5];
    float* M[5];
    float* V[5];
    for (int k = 0; k < 5; k++) {
        P[k] = t->param[k].f();
        M[k] = t->adam_m[k].f();
        V[k] = t->adam_v[k].f();
    }
    gs_param_grads pg{t->grad[0].f(), t->grad[1].f(), t->grad[2].f(), t->grad[3].f(), t->grad[4].f()};
    // camera and hyper-parameters travel as kernel arguments: no copies on the stream
    hipError_t e = launch_preprocess_bwd_multi(t->gaussians(), nullptr, &ca, 1, radii, rg, pg, P, M, V, nullptr, &h,
                                               t->sh_degree, c->stream);
    if (e != hipSuccess) return gs_hip_status(e, "preprocess backward + adam kernel");
    return GS_OK;
}

extern "C" int gs_trainer_set_timing(gs_trainer* t, unsigned stage_mask) {
    if (!t) return gs_set_error(GS_ERR_INVALID, "null trainer");
    stage_mask &= (1u << GS_NUM_STAGES) - 1u;
    for (int s = 0; s < GS_NUM_STAGES; s++) {
        t->timer.total_ms[s] = 0.0;
        if (((stage_mask >> s) & 1u) && !t->timer.beg[s]) {
            hipError_t e = hipEventCreate(&t->timer.beg[s]);
            if (e == hipSuccess) e = hipEventCreate(&t->timer.end[s]);
            if (e != hipSuccess) return gs_hip_status(e, "hipEventCreate(stage timer)");
        }
    }
    t->timer.steps = 0;
    t->timer.mask = stage_mask;
    return GS_OK;
}

extern "C" int gs_trainer_stage_times(gs_trainer* t, double* ms, int n, long long* steps) {
    if (!t || (n > 0 && !ms)) return gs_set_error(GS_ERR_INVALID, "gs_trainer_stage_times: null argument");
    for (int s = 0; s < n && s < GS_NUM_STAGES; s++)
        ms[s] = (t->timer.steps > 0 && t->timer.on(s)) ? t->timer.total_ms[s] / (double)t->timer.steps : 0.0;
    for (int s = GS_NUM_STAGES; s < n; s++) ms[s] = 0.0;
    if (steps) *steps = t->timer.steps;
    return GS_OK;
}

extern "C" int gs_trainer_set_options(gs_trainer* t, int fused_adam) {
    if (!t) return gs_set_error(GS_ERR_INVALID, "null trainer");
    t->fused_adam = fused_adam ? 1 : 0;
    return GS_OK;
}

// Forward, L1 loss and raster backward of one view into the trainer's view
// buffers.  With a pending pair count the host checks it here, while the
// rasterizers run, and redoes the three stages in the rare case that the pairs
// outgrew the buffer.  *K: the pair count when it was read back.
int tr_forward_raster_bwd(gs_trainer* t, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt,
                          long long* Kout) {
    gs_ctx* c = t->ctx;
    if (t->views.empty()) t->views.resize(1);
    gs::ViewBufs& v = t->views[0];
    long long K = 0;
    auto fwd_bwd = [&](bool* pending) -> int {
        TR_TRY(tr_forward(t, v, cam, cfg, &K, pending));
        TR_TRY(tr_mark(t, GS_STAGE_LOSS, true));
        TR_TRY(gs_l1_loss(c, cam->width, cam->height, v.image.f(), gt, v.d_image.f(), static_cast<double*>(v.loss.p)));
        TR_TRY(tr_mark(t, GS_STAGE_LOSS, false));
        TR_TRY(tr_mark(t, GS_STAGE_CLEAR, true));
        TR_TRY(gs_zero(c, v.rg.p, (size_t)t->n * 36));
        TR_TRY(tr_mark(t, GS_STAGE_CLEAR, false));
        gs_preprocessed pre = v.pre();
        gs_pixel_state px{v.T_final.f(), v.n_contrib.i(), v.last.u()};
        gs_raster_grads rg = v.rgrads();
        TR_TRY(tr_mark(t, GS_STAGE_RASTER_BWD, true));
        if (t->pair_stride == 1)
            TR_TRY(gs_raster_bwd(c, cam->width, cam->height, &pre, t->keys_sorted.u(), v.ranges.u(), cfg->bg_color,
                                 &px, v.d_image.f(), &rg));
        else
            TR_TRY(gs_raster_bwd_words(c, cam->width, cam->height, &pre,
                                       static_cast<const uint64_t*>(t->keys_sorted.p), v.ranges.u(), cfg->bg_color,
                                       &px, v.d_image.f(), &rg));
        TR_TRY(tr_mark(t, GS_STAGE_RASTER_BWD, false));
        return GS_OK;
    };
    bool pending = false;
    TR_TRY(fwd_bwd(&pending));
    if (pending) {
        bool fits = true;
        TR_TRY(tr_pairs(t, &K, &fits));
        if (!fits) {
            t->allow_async = false;
            const int st = fwd_bwd(&pending);
            t->allow_async = true;
            TR_TRY(st);
        }
    }
    *Kout = K;
    return GS_OK;
}

extern "C" int gs_trainer_step(gs_trainer* t, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt,
                               const gs_train_cfg* tc, int iteration, gs_step_result* out) {
    if (!t || !cam || !cfg || !gt || !tc) return gs_set_error(GS_ERR_INVALID, "gs_trainer_step: null argument");
    TR_TRY(check_cam_cfg(cam, cfg));
    gs_ctx* c = t->ctx;
    long long K = 0;
    TR_TRY(tr_forward_raster_bwd(t, cam, cfg, gt, &K));
    gs::ViewBufs& v = t->views[0];
    gs_preprocessed pre = v.pre();
    gs_raster_grads rg = v.rgrads();
    gs_gaussians g = t->gaussians();
    gs_param_grads pg{t->grad[0].f(), t->grad[1].f(), t->grad[2].f(), t->grad[3].f(), t->grad[4].f()};
    TR_TRY(tr_mark(t, GS_STAGE_BACKWARD, true));
    if (t->fused_adam && tc->use_adam) {
        TR_TRY(tr_fused_backward_adam(t, cam, cfg, tc, iteration, pre.radii, rg));
    } else {
        TR_TRY(gs_preprocess_bwd(c, &g, cam, cfg, pre.radii, &rg, &pg, 0));
        TR_TRY(tr_optimizer(t, tc, iteration));
    }
    TR_TRY(tr_mark(t, GS_STAGE_BACKWARD, false));
    // out == NULL and no stage timing: the step stays queued on the stream (no
    // read-back, no wait), so the next step's launches follow it back to back;
    // the loss stays in the "loss" buffer until the next step
    if (!out && !t->timer.mask) return GS_OK;
    double loss = 0.0;
    hipError_t e = hipMemcpyAsync(&loss, v.loss.p, 8, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return gs_hip_status(e, "hipMemcpyAsync(loss)");
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return gs_hip_status(e, "hipStreamSynchronize");
    if (t->timer.mask) {
        e = t->timer.collect();
        if (e != hipSuccess) return gs_hip_status(e, "hipEventElapsedTime(stage timer)");
    }
    if (out) {
        out->loss = loss;
        out->iteration = iteration;
        out->n_gaussians = t->n;
        out->pairs = K;
    }
    return GS_OK;
}

extern "C" int gs_trainer_render(gs_trainer* t, const gs_camera* cam, const gs_render_cfg* cfg, long long* pairs) {
    if (!t || !cam || !cfg) return gs_set_error(GS_ERR_INVALID, "gs_trainer_render: null argument");
    TR_TRY(check_cam_cfg(cam, cfg));
    if (t->views.empty()) t->views.resize(1);
    long long K = 0;
    bool pending = false;
    TR_TRY(tr_forward(t, t->views[0], cam, cfg, &K, &pending));
    if (pending) {
        bool fits = true;
        TR_TRY(tr_pairs(t, &K, &fits));
        if (!fits) {
            t->allow_async = false;
            const int st = tr_forward(t, t->views[0], cam, cfg, &K, &pending);
            t->allow_async = true;
            TR_TRY(st);
        }
    }
    if (pairs) *pairs = K;
    return GS_OK;
}
