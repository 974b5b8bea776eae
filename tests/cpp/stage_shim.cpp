// stage_shim.cpp -- test-only extern "C" shim over the C++ drop-in classes
// (gs3d_hip.h), so the tests can run the reference-shaped stage objects on
// host arrays and compare them with the oracle.  Built into
// tests/cpp/_build/libgs_stage_shim.so, linked against libgs_hip.so.
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <memory>
#include <vector>

#include "gs3d_hip.h"

static Camera cam_of(const gs_camera* c) {
    Camera k;
    std::memcpy(&k, c, sizeof(Camera));
    return k;
}
static RenderConfig cfg_of(const gs_render_cfg* c) {
    RenderConfig r;
    for (int i = 0; i < 3; i++) r.bg_color[i] = c->bg_color[i];
    r.scale_modifier = c->scale_modifier;
    r.antialiasing = c->antialiasing != 0;
    r.training = c->training != 0;
    r.sh_degree = c->sh_degree;
    r.tile_w = c->tile_w;
    r.tile_h = c->tile_h;
    return r;
}
static GaussianData gd(int n, int deg, int mc, float* pos, float* sh, float* sc, float* rot, float* op) {
    GaussianData g;
    g.count = n; g.sh_degree = deg; g.max_coeffs = mc;
    g.positions = pos; g.sh_coeffs = sh; g.scales = sc; g.rotations = rot; g.opacities = op; g.filter_3D = nullptr;
    return g;
}

static std::string g_err;

extern "C" const char* shim_error() { return g_err.c_str(); }

#define SHIM_TRY(...)                   \
    try {                               \
        __VA_ARGS__;                    \
        return 0;                       \
    } catch (const std::exception& e) { \
        g_err = e.what();               \
        return 1;                       \
    }

// Renderer(PreprocessorHIP, TileBinnerHIP, SorterHIP, RasterizerHIP)::render on activated host arrays.
extern "C" int shim_render(int n, int deg, int mc, float* pos, float* sh, float* sc, float* rot, float* op,
                           const gs_camera* cam, const gs_render_cfg* cfg, float* img) {
    SHIM_TRY({
        HipContext ctx(0);
        Renderer r(std::unique_ptr<Preprocessor>(new PreprocessorHIP(ctx)),
                   std::unique_ptr<TileBinner>(new TileBinnerHIP(ctx)), std::unique_ptr<Sorter>(new SorterHIP(ctx)),
                   std::unique_ptr<Rasterizer>(new RasterizerHIP(ctx)), 512ull << 20);
        r.render(gd(n, deg, mc, pos, sh, sc, rot, op), cam_of(cam), cfg_of(cfg), img);
    })
}

// Renderer::render with the inverse-depth output (RasterizerHIP out_depth).
extern "C" int shim_render_depth(int n, int deg, int mc, float* pos, float* sh, float* sc, float* rot, float* op,
                                 const gs_camera* cam, const gs_render_cfg* cfg, float* img, float* depth) {
    SHIM_TRY({
        HipContext ctx(0);
        Renderer r(std::unique_ptr<Preprocessor>(new PreprocessorHIP(ctx)),
                   std::unique_ptr<TileBinner>(new TileBinnerHIP(ctx)), std::unique_ptr<Sorter>(new SorterHIP(ctx)),
                   std::unique_ptr<Rasterizer>(new RasterizerHIP(ctx)), 512ull << 20);
        r.render(gd(n, deg, mc, pos, sh, sc, rot, op), cam_of(cam), cfg_of(cfg), img, depth);
    })
}

// The free functions l1_loss / l1_loss_double (loss.h:6-10) on host arrays.
extern "C" int shim_l1(const float* rendered, const float* gt, int H, int W, float* d_image, float* d_image2,
                       float* loss_f, double* loss_d) {
    SHIM_TRY({
        *loss_f = l1_loss(rendered, gt, H, W, d_image);
        *loss_d = l1_loss_double(rendered, gt, H, W, d_image2);
    })
}

// The eval_3D forward through the stage objects (RenderConfig::eval_3D set),
// with the ForwardCache's T_final / n_contrib; filter may be NULL.
extern "C" int shim_render_3d(int n, int deg, int mc, float* pos, float* sh, float* sc, float* rot, float* op,
                              float* filter, const gs_camera* cam, const gs_render_cfg* cfg, float* img,
                              float* T_final, int* n_contrib, int* radii, float* depth) {
    SHIM_TRY({
        HipContext ctx(0);
        PreprocessorHIP pp(ctx);
        TileBinnerHIP bn(ctx);
        SorterHIP sr(ctx);
        RasterizerHIP rs(ctx);
        FrameAllocator alloc(512ull << 20);
        GaussianData g = gd(n, deg, mc, pos, sh, sc, rot, op);
        g.filter_3D = filter;
        Camera k = cam_of(cam);
        RenderConfig f = cfg_of(cfg);
        f.eval_3D = true;
        PreprocessOutput pre = pp.process(g, k, f, alloc, nullptr);
        if (!pre.eval_3D || !pre.gauss2screen || !pre.cov3D_inv || !pre.mean_offset)
            throw std::runtime_error("eval_3D arrays missing from the preprocess output");
        std::memcpy(radii, pre.radii, (size_t)n * 4);
        BinningOutput bin = bn.bin(pre, n, k, f, alloc);
        sr.sort(bin, alloc);
        ForwardCache cache{};
        rs.rasterize(pre, bin, k, f, img, depth, &cache, &alloc);
        const size_t npix = (size_t)k.width * k.height;
        std::memcpy(T_final, cache.T_final, npix * 4);
        std::memcpy(n_contrib, cache.n_contrib, npix * 4);
    })
}

// Each stage object in turn; host copies of every stage output.  keys/values
// buffers hold `cap` pairs; *K receives the pair count.
extern "C" int shim_stages(int n, int deg, int mc, float* pos, float* sh, float* sc, float* rot, float* op,
                           const gs_camera* cam, const gs_render_cfg* cfg, float* means2D, float* depths,
                           float* conics, float* op2d, float* rgb, int* radii, int* tiles, long long cap,
                           long long* K, uint64_t* keys_sorted, uint32_t* values_sorted, uint32_t* ranges,
                           float* img, float* T_final, int* n_contrib) {
    SHIM_TRY({
        HipContext ctx(0);
        FrameAllocator alloc(1ull << 30);
        PreprocessorHIP pp(ctx);
        TileBinnerHIP tb(ctx);
        tb.fill_host_outputs = true;
        SorterHIP so(ctx);
        RasterizerHIP rs(ctx);
        Camera c = cam_of(cam);
        RenderConfig f = cfg_of(cfg);
        PreprocessOutput pre = pp.process(gd(n, deg, mc, pos, sh, sc, rot, op), c, f, alloc);
        std::memcpy(means2D, pre.means2D, (size_t)n * 8);
        std::memcpy(depths, pre.depths, (size_t)n * 4);
        std::memcpy(conics, pre.conics, (size_t)n * 12);
        std::memcpy(op2d, pre.opacities_2d, (size_t)n * 4);
        std::memcpy(rgb, pre.rgb, (size_t)n * 12);
        std::memcpy(radii, pre.radii, (size_t)n * 4);
        std::memcpy(tiles, pre.tiles_touched, (size_t)n * 4);
        BinningOutput bin = tb.bin(pre, n, c, f, alloc);
        *K = bin.total_pairs;
        if (bin.total_pairs > cap) throw std::runtime_error("shim: pair buffer too small");
        if (bin.total_pairs > 0) so.sort(bin, alloc);
        if (bin.total_pairs > 0) {
            std::memcpy(keys_sorted, bin.keys_sorted, (size_t)bin.total_pairs * 8);
            std::memcpy(values_sorted, bin.values_sorted, (size_t)bin.total_pairs * 4);
        }
        std::memcpy(ranges, bin.tile_ranges, (size_t)bin.num_tiles * 8);
        ForwardCache cache{};
        rs.rasterize(pre, bin, c, f, img, nullptr, &cache, &alloc);
        std::memcpy(T_final, cache.T_final, (size_t)c.width * c.height * 4);
        std::memcpy(n_contrib, cache.n_contrib, (size_t)c.width * c.height * 4);
    })
}

// Host-produced stage inputs (no device_data): the binner, sorter and
// rasterizer upload them.  pre arrays as from a CPU preprocessor.
extern "C" int shim_from_host(int n, float* means2D, float* depths, float* conics, float* op2d, float* rgb,
                              int* radii, int* tiles, const gs_camera* cam, const gs_render_cfg* cfg, float* img) {
    SHIM_TRY({
        HipContext ctx(0);
        FrameAllocator alloc(1ull << 30);
        PreprocessOutput pre{};
        pre.means2D = means2D; pre.depths = depths; pre.conics = conics; pre.opacities_2d = op2d; pre.rgb = rgb;
        pre.radii = radii; pre.tiles_touched = tiles;
        Camera c = cam_of(cam);
        RenderConfig f = cfg_of(cfg);
        TileBinnerHIP tb(ctx);
        tb.fill_host_outputs = true;
        SorterHIP so(ctx);
        RasterizerHIP rs(ctx);
        BinningOutput bin = tb.bin(pre, n, c, f, alloc);
        // drop the device data: the sorter and rasterizer must work from host arrays
        bin.device_data = nullptr;
        if (bin.total_pairs > 0) so.sort(bin, alloc);
        bin.device_data = nullptr;
        rs.rasterize(pre, bin, c, f, img);
    })
}

// RasterizerBackwardHIP + PreprocessorBackwardHIP after PreprocessorHIP /
// TileBinnerHIP / SorterHIP, with the caller's d_image; raw parameters in.
extern "C" int shim_backward(int n, int deg, int mc, float* raw_pos, float* raw_sc, float* raw_rot, float* raw_sh,
                             float* raw_op, const gs_camera* cam, const gs_render_cfg* cfg, const float* d_image,
                             float* d_means2D, float* d_conics, float* d_rgb, float* d_op2d, float* g_pos,
                             float* g_sc, float* g_rot, float* g_sh, float* g_op) {
    SHIM_TRY({
        HipContext ctx(0);
        FrameAllocator alloc(1ull << 30);
        RawGaussianParams raw{n, deg, mc, raw_pos, raw_sc, raw_rot, raw_sh, raw_op};
        std::vector<float> ap(n * 3), as(n * 3), ar(n * 4), ash((size_t)n * mc * 3), ao(n);
        GaussianData g = gd(n, deg, mc, ap.data(), ash.data(), as.data(), ar.data(), ao.data());
        raw.activate(g);
        Camera c = cam_of(cam);
        RenderConfig f = cfg_of(cfg);
        PreprocessorHIP pp(ctx);
        TileBinnerHIP tb(ctx);
        SorterHIP so(ctx);
        PreprocessOutput pre = pp.process(g, c, f, alloc);
        BinningOutput bin = tb.bin(pre, n, c, f, alloc);
        if (bin.total_pairs > 0) so.sort(bin, alloc);
        ForwardCache cache{};
        cache.pre = &pre;
        cache.bin = &bin;
        RasterGradOutput rg{d_means2D, d_conics, d_rgb, d_op2d};
        RasterizerBackwardHIP rb(ctx);
        rb.backward(pre, bin, c, f, cache, d_image, n, rg);
        GradientOutput go{g_pos, g_sc, g_rot, g_sh, g_op};
        PreprocessorBackwardHIP pb(ctx);
        pb.backward(g, c, f, cache, rg, raw, go);
    })
}

// TrainerHIP::step `steps` times on host raw parameters (updated in place).
extern "C" int shim_trainer(int n, int deg, int mc, float* raw_pos, float* raw_sc, float* raw_rot, float* raw_sh,
                            float* raw_op, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt,
                            int use_adam, int steps, float* losses, float* img) {
    SHIM_TRY({
        HipContext ctx(0);
        TrainerHIP tr(ctx);
        RawGaussianParams raw{n, deg, mc, raw_pos, raw_sc, raw_rot, raw_sh, raw_op};
        TrainConfig tc;
        tc.use_adam = use_adam != 0;
        for (int s = 0; s < steps; s++) losses[s] = tr.step(raw, cam_of(cam), gt, cfg_of(cfg), tc, s).loss;
        std::memcpy(img, tr.last_rendered(), (size_t)cam->width * cam->height * 12);
    })
}

// AdamOptimizerHIP::step on host arrays.
extern "C" int shim_adam(int n, int deg, int mc, float* p_pos, float* p_sc, float* p_rot, float* p_sh, float* p_op,
                         float* g_pos, float* g_sc, float* g_rot, float* g_sh, float* g_op, int steps) {
    SHIM_TRY({
        HipContext ctx(0);
        AdamOptimizerHIP adam(ctx);
        RawGaussianParams p{n, deg, mc, p_pos, p_sc, p_rot, p_sh, p_op};
        GradientOutput g{g_pos, g_sc, g_rot, g_sh, g_op};
        TrainConfig tc;
        tc.use_adam = true;
        for (int s = 0; s < steps; s++) adam.step(p, g, tc, s);
    })
}

// ---- DensityControllerHIP, OwnedRawParams and TrainerHIP::step_with_densify
// behind opaque handles (the same calls as the reference's, for the tests). ----
static HipContext& shim_ctx() {
    static HipContext ctx(0);
    return ctx;
}

extern "C" void* shim_params_new(int N, int sh_degree, int max_coeffs, const float* pos, const float* sc,
                                 const float* rot, const float* sh, const float* op) {
    OwnedRawParams* p = new OwnedRawParams();
    p->sh_degree = sh_degree;
    p->max_coeffs = max_coeffs;
    p->resize(N);
    std::memcpy(p->positions.data(), pos, (size_t)N * 12);
    std::memcpy(p->scales.data(), sc, (size_t)N * 12);
    std::memcpy(p->rotations.data(), rot, (size_t)N * 16);
    std::memcpy(p->sh_coeffs.data(), sh, (size_t)N * max_coeffs * 12);
    std::memcpy(p->opacities.data(), op, (size_t)N * 4);
    return p;
}
extern "C" void shim_params_free(void* h) { delete static_cast<OwnedRawParams*>(h); }
extern "C" int shim_params_count(void* h) { return static_cast<OwnedRawParams*>(h)->count(); }
extern "C" void shim_params_get(void* h, float* pos, float* sc, float* rot, float* sh, float* op) {
    OwnedRawParams* p = static_cast<OwnedRawParams*>(h);
    const size_t n = (size_t)p->count();
    std::memcpy(pos, p->positions.data(), n * 12);
    std::memcpy(sc, p->scales.data(), n * 12);
    std::memcpy(rot, p->rotations.data(), n * 16);
    std::memcpy(sh, p->sh_coeffs.data(), n * p->max_coeffs * 12);
    std::memcpy(op, p->opacities.data(), n * 4);
}

static DensifyConfig shim_dcfg(const int* icfg, const float* fcfg) {
    DensifyConfig c;
    c.densify_from_iter = icfg[0]; c.densify_until_iter = icfg[1]; c.densify_interval = icfg[2];
    c.opacity_reset_interval = icfg[3]; c.max_screen_size = icfg[4]; c.enabled = icfg[5] != 0;
    c.grad_threshold = fcfg[0]; c.min_opacity = fcfg[1]; c.percent_dense = fcfg[2];
    return c;
}

extern "C" void* shim_dc_new() {
    try {
        return new DensityControllerHIP(shim_ctx());
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
extern "C" void shim_dc_free(void* h) { delete static_cast<DensityControllerHIP*>(h); }
extern "C" int shim_dc_reset(void* h, int N) { SHIM_TRY(static_cast<DensityControllerHIP*>(h)->reset(N)) }
extern "C" int shim_dc_accumulate(void* h, const float* d_means2D, const int* radii, int N) {
    SHIM_TRY(static_cast<DensityControllerHIP*>(h)->accumulate_stats(d_means2D, radii, N))
}
extern "C" int shim_dc_densify(void* h, void* params, const int* icfg, const float* fcfg, float extent, int* n_out) {
    SHIM_TRY(*n_out = static_cast<DensityControllerHIP*>(h)->densify_and_prune(*static_cast<OwnedRawParams*>(params),
                                                                                shim_dcfg(icfg, fcfg), extent))
}
extern "C" void shim_dc_reset_opacity(void* params) {
    DensityControllerHIP::reset_opacity(*static_cast<OwnedRawParams*>(params));
}

extern "C" void* shim_trainer_new() {
    try {
        return new TrainerHIP(shim_ctx());
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
extern "C" void shim_trainer_free(void* h) { delete static_cast<TrainerHIP*>(h); }
extern "C" int shim_trainer_step_densify(void* h, void* params, const gs_camera* cam, const gs_render_cfg* cfg,
                                         const float* gt, const float* tcf, const int* tci, const int* icfg,
                                         const float* fcfg, float extent, int iteration, float* loss, int* n_out) {
    SHIM_TRY({
        TrainConfig tc;
        tc.lr_position_init = tcf[0]; tc.lr_position_final = tcf[1]; tc.lr_feature = tcf[2];
        tc.lr_opacity = tcf[3]; tc.lr_scaling = tcf[4]; tc.lr_rotation = tcf[5];
        tc.adam_beta1 = tcf[6]; tc.adam_beta2 = tcf[7]; tc.adam_eps = tcf[8];
        tc.max_steps = tci[0]; tc.use_adam = tci[1] != 0;
        TrainerHIP::StepResult r = static_cast<TrainerHIP*>(h)->step_with_densify(
            *static_cast<OwnedRawParams*>(params), cam_of(cam), gt, cfg_of(cfg), tc, shim_dcfg(icfg, fcfg), extent,
            iteration);
        *loss = r.loss;
        *n_out = r.n_gaussians;
    })
}

// FrameAllocator's contract as the reference's tests/test_types.cpp:10-41
// states it (host only, no device call): 64-byte aligned blocks, a
// std::runtime_error on exhaustion, the whole arena reusable after reset().
// Returns 0 when every check holds, else the number of the failed check.
extern "C" int shim_frame_allocator_contract() {
    FrameAllocator a(1024);
    if (a.capacity() < 1024 || a.used() != 0) return 1;
    void* p = a.allocate(10);
    void* q = a.allocate(7);
    if (reinterpret_cast<uintptr_t>(p) % 64 || reinterpret_cast<uintptr_t>(q) % 64) return 2;
    if (static_cast<char*>(q) - static_cast<char*>(p) != 64 || a.used() != 64 + 7) return 3;
    float* f = a.allocate_array<float>(16);
    if (reinterpret_cast<uintptr_t>(f) % 64) return 4;
    a.reset();
    if (a.used() != 0 || a.allocate(1) != p) return 5;
    a.reset();
    if (a.allocate(a.capacity()) != p || a.used() != a.capacity()) return 6;
    try {
        a.allocate(1);
        return 7;
    } catch (const std::runtime_error&) {
    }
    FrameAllocator small(64);
    try {
        small.allocate(128);
        return 8;
    } catch (const std::runtime_error&) {
    }
    if (small.used() != 0) return 9;
    try {
        small.allocate(SIZE_MAX - 8);  // no wrap-around past the end
        return 10;
    } catch (const std::runtime_error&) {
    }
    return 0;
}
