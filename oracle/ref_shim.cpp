// ref_shim.cpp -- extern "C" driver around the reference's own CPU classes.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product).  oracle/Makefile.ref
// compiles this file together with the reference's CPU sources where they lie
// under harmonyos_3dgs/src (nothing is copied), into oracle/_ref/libgs_ref.so.
// It is used to validate the restatement in gs_oracle.c and as the "reference"
// CPU baseline.  It follows the call sequence of harmonyos_3dgs/src/
// verify_grads_main.cpp:83-177 (forward + L1 + backward) and of
// Trainer::step (trainer.cpp:7-70) for full training steps.
#include "types.h"
#include "train_types.h"
#include "loss.h"
#include "optimizer.h"
#include "trainer.h"
#include "density_controller.h"
#include "cpu/preprocessor_cpu.h"
#include "cpu/tile_binner_cpu.h"
#include "cpu/sorter_cpu.h"
#include "cpu/rasterizer_cpu.h"
#include "cpu/rasterizer_backward_cpu.h"
#include "cpu/preprocessor_backward_cpu.h"
#include "ply_loader.h"
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

struct gs_camera {
    float view_matrix[16], viewproj_matrix[16], cam_pos[3], tan_fovx, tan_fovy;
    int width, height;
};
struct gs_render_cfg {
    float bg_color[3];
    float scale_modifier;
    int antialiasing, training, sh_degree, tile_w, tile_h;
};

static Camera to_cam(const gs_camera* c) {
    Camera cam{};
    std::memcpy(cam.view_matrix, c->view_matrix, 64);
    std::memcpy(cam.viewproj_matrix, c->viewproj_matrix, 64);
    std::memcpy(cam.cam_pos, c->cam_pos, 12);
    cam.tan_fovx = c->tan_fovx; cam.tan_fovy = c->tan_fovy;
    cam.width = c->width; cam.height = c->height;
    return cam;
}
static RenderConfig to_cfg(const gs_render_cfg* c) {
    RenderConfig cfg;
    for (int i = 0; i < 3; i++) cfg.bg_color[i] = c->bg_color[i];
    cfg.scale_modifier = c->scale_modifier;
    cfg.antialiasing = c->antialiasing != 0;
    cfg.eval_3D = false;
    cfg.training = c->training != 0;
    cfg.sh_degree = c->sh_degree;
    cfg.tile_w = c->tile_w; cfg.tile_h = c->tile_h;
    return cfg;
}

extern "C" {

// Forward (activation, preprocess with cache, bin, sort, raster), L1 loss in
// double, raster backward and preprocess backward.  Returns the pair count, or
// -1 if an output buffer of `max_pairs` entries is too small.
long long ref_forward_backward(int N, int sh_degree, int max_coeffs, float* raw_pos, float* raw_sc,
                               float* raw_rot, float* raw_sh, float* raw_op, const gs_camera* camp,
                               const gs_render_cfg* cfgp, const float* gt, size_t arena_bytes,
                               float* img, double* loss, float* T_final, int* n_contrib,
                               float* means2D, float* depths, float* conics, float* op2d, float* rgb,
                               int* radii, int* tiles, float* g_pos, float* g_sc, float* g_rot,
                               float* g_sh, float* g_op, float* d_means2D, float* d_conics,
                               float* d_rgb, float* d_op2d) {
    RawGaussianParams raw;
    raw.count = N; raw.sh_degree = sh_degree; raw.max_coeffs = max_coeffs;
    raw.raw_positions = raw_pos; raw.raw_scales = raw_sc; raw.raw_rotations = raw_rot;
    raw.raw_sh_coeffs = raw_sh; raw.raw_opacities = raw_op;
    std::vector<float> ap(N * 3), as(N * 3), ar(N * 4), ash((size_t)N * max_coeffs * 3), ao(N);
    GaussianData g;
    g.count = N; g.sh_degree = sh_degree; g.max_coeffs = max_coeffs;
    g.positions = ap.data(); g.scales = as.data(); g.rotations = ar.data();
    g.sh_coeffs = ash.data(); g.opacities = ao.data(); g.filter_3D = nullptr;
    raw.activate(g);
    Camera cam = to_cam(camp);
    RenderConfig cfg = to_cfg(cfgp);
    FrameAllocator alloc(arena_bytes);
    PreprocessorCPU pp; TileBinnerCPU bn; SorterCPU sr; RasterizerCPU rs;
    ForwardCache cache{};
    PreprocessOutput pre = pp.process(g, cam, cfg, alloc, &cache);
    BinningOutput bin = bn.bin(pre, N, cam, cfg, alloc);
    if (bin.total_pairs > 0) sr.sort(bin, alloc);
    cache.pre = &pre; cache.bin = &bin;
    int npix = cam.width * cam.height;
    rs.rasterize(pre, bin, cam, cfg, img, nullptr, &cache, &alloc);
    float* d_image = alloc.allocate_array<float>((size_t)npix * 3);
    *loss = l1_loss_double(img, gt, cam.height, cam.width, d_image);
    RasterGradOutput rg; rg.allocate_and_zero(alloc, N);
    RasterizerBackwardCPU rb; rb.backward(pre, bin, cam, cfg, cache, d_image, rg);
    GradientOutput gr; gr.allocate_and_zero(alloc, N, max_coeffs);
    PreprocessorBackwardCPU pb; pb.backward(g, cam, cfg, cache, rg, raw, gr);

    std::memcpy(T_final, cache.T_final, npix * 4);
    std::memcpy(n_contrib, cache.n_contrib, npix * 4);
    for (int i = 0; i < N; i++) {
        bool vis = pre.radii[i] > 0;
        radii[i] = pre.radii[i]; tiles[i] = pre.tiles_touched[i];
        means2D[2 * i] = vis ? pre.means2D[2 * i] : 0; means2D[2 * i + 1] = vis ? pre.means2D[2 * i + 1] : 0;
        depths[i] = vis ? pre.depths[i] : 0; op2d[i] = vis ? pre.opacities_2d[i] : 0;
        for (int k = 0; k < 3; k++) {
            conics[3 * i + k] = vis ? pre.conics[3 * i + k] : 0;
            rgb[3 * i + k] = vis ? pre.rgb[3 * i + k] : 0;
        }
    }
    std::memcpy(g_pos, gr.d_raw_positions, N * 12);
    std::memcpy(g_sc, gr.d_raw_scales, N * 12);
    std::memcpy(g_rot, gr.d_raw_rotations, N * 16);
    std::memcpy(g_sh, gr.d_raw_sh_coeffs, (size_t)N * max_coeffs * 12);
    std::memcpy(g_op, gr.d_raw_opacities, N * 4);
    std::memcpy(d_means2D, rg.d_means2D, N * 8);
    std::memcpy(d_conics, rg.d_conics, N * 12);
    std::memcpy(d_rgb, rg.d_rgb, N * 12);
    std::memcpy(d_op2d, rg.d_opacities_2d, N * 4);
    return bin.total_pairs;
}

// `steps` calls of Trainer::step with Adam on one camera; raw params are
// updated in place.  Returns the loss of the last step.
float ref_train_steps(int N, int sh_degree, int max_coeffs, float* raw_pos, float* raw_sc,
                      float* raw_rot, float* raw_sh, float* raw_op, const gs_camera* camp,
                      const gs_render_cfg* cfgp, const float* gt, float lr_pos_init,
                      float lr_pos_final, float lr_feature, float lr_opacity, float lr_scaling,
                      float lr_rotation, int max_steps, float beta1, float beta2, float eps,
                      int first_iteration, int steps, size_t arena_bytes, float* img_out) {
    RawGaussianParams raw;
    raw.count = N; raw.sh_degree = sh_degree; raw.max_coeffs = max_coeffs;
    raw.raw_positions = raw_pos; raw.raw_scales = raw_sc; raw.raw_rotations = raw_rot;
    raw.raw_sh_coeffs = raw_sh; raw.raw_opacities = raw_op;
    Camera cam = to_cam(camp);
    RenderConfig cfg = to_cfg(cfgp);
    TrainConfig tc;
    tc.lr_position_init = lr_pos_init; tc.lr_position_final = lr_pos_final;
    tc.lr_feature = lr_feature; tc.lr_opacity = lr_opacity; tc.lr_scaling = lr_scaling;
    tc.lr_rotation = lr_rotation; tc.max_steps = max_steps; tc.use_adam = true;
    tc.adam_beta1 = beta1; tc.adam_beta2 = beta2; tc.adam_eps = eps;
    std::unique_ptr<Trainer> tr(new Trainer(arena_bytes));
    float loss = 0.f;
    for (int s = 0; s < steps; s++) loss = tr->step(raw, cam, gt, cfg, tc, first_iteration + s).loss;
    if (img_out) std::memcpy(img_out, tr->last_rendered(), (size_t)cam.width * cam.height * 12);
    return loss;
}

// loadPly (ply_loader.cpp:32-189) into caller arrays sized from *count /
// *max_coeffs (call with null arrays first to get them).  Returns 0, or -1 on
// a loader exception.
int ref_load_ply(const char* path, int* count, int* max_coeffs, int* has_filter, float* pos, float* sc,
                 float* rot, float* sh, float* op, float* filter) {
    try {
        LoadedModel m = loadPly(path);
        const GaussianData& g = m.data;
        *count = g.count; *max_coeffs = g.max_coeffs; *has_filter = g.filter_3D != nullptr;
        if (pos) {
            const size_t n = (size_t)g.count;
            std::memcpy(pos, g.positions, n * 12);
            std::memcpy(sc, g.scales, n * 12);
            std::memcpy(rot, g.rotations, n * 16);
            std::memcpy(sh, g.sh_coeffs, n * g.max_coeffs * 12);
            std::memcpy(op, g.opacities, n * 4);
            if (filter && g.filter_3D) std::memcpy(filter, g.filter_3D, n * 4);
        }
        m.free();
        return 0;
    } catch (...) {
        return -1;
    }
}

// The 2D forward with Rasterizer::rasterize's inverse-depth output
// (rasterizer_cpu.cpp:170-172,197) on activated parameters; the pairs are
// stably sorted as in ref_render_3d.
int ref_inv_depth(int N, int sh_degree, int max_coeffs, float* pos, float* sc, float* rot, float* sh, float* op,
                  const gs_camera* camp, const gs_render_cfg* cfgp, size_t arena_bytes, float* img,
                  float* inv_depth) {
    GaussianData g;
    g.count = N; g.sh_degree = sh_degree; g.max_coeffs = max_coeffs;
    g.positions = pos; g.scales = sc; g.rotations = rot; g.sh_coeffs = sh; g.opacities = op;
    g.filter_3D = nullptr;
    Camera cam = to_cam(camp);
    RenderConfig cfg = to_cfg(cfgp);
    FrameAllocator alloc(arena_bytes);
    PreprocessorCPU pp; TileBinnerCPU bn; SorterCPU sr; RasterizerCPU rs;
    PreprocessOutput pre = pp.process(g, cam, cfg, alloc, nullptr);
    BinningOutput bin = bn.bin(pre, N, cam, cfg, alloc);
    const long long K = bin.total_pairs;
    if (K > 0) {
        sr.sort(bin, alloc);
        std::vector<uint32_t> ord(K);
        for (long long i = 0; i < K; i++) ord[i] = (uint32_t)i;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](uint32_t a, uint32_t b) { return bin.keys_unsorted[a] < bin.keys_unsorted[b]; });
        for (long long i = 0; i < K; i++) {
            bin.keys_sorted[i] = bin.keys_unsorted[ord[i]];
            bin.values_sorted[i] = bin.values_unsorted[ord[i]];
        }
    }
    rs.rasterize(pre, bin, cam, cfg, img, inv_depth, nullptr, nullptr);
    return 0;
}

// The eval_3D forward (PreprocessorCPU with RenderConfig::eval_3D, the
// per-tile depth keys of TileBinnerCPU, the k-buffer RasterizerCPU) on
// activated parameters.  The keys/values are returned in emission order; the
// sorted values use a stable order of the keys (SorterCPU's std::sort can
// only differ on key ties; its tile ranges are kept).  Returns the pair count,
// or -1 if `max_pairs` is too small.
long long ref_render_3d(int N, int sh_degree, int max_coeffs, float* pos, float* sc, float* rot, float* sh,
                        float* op, float* filter3d, const gs_camera* camp, const gs_render_cfg* cfgp,
                        size_t arena_bytes, long long max_pairs, float* means2D, float* depths, float* op2d,
                        float* rgb, int* radii, int* tiles, float* g2s, float* cinv, float* moff, uint64_t* keys,
                        uint32_t* values, uint32_t* values_sorted, uint32_t* ranges, float* img, float* T_final,
                        int* n_contrib) {
    GaussianData g;
    g.count = N; g.sh_degree = sh_degree; g.max_coeffs = max_coeffs;
    g.positions = pos; g.scales = sc; g.rotations = rot; g.sh_coeffs = sh; g.opacities = op;
    g.filter_3D = filter3d;
    Camera cam = to_cam(camp);
    RenderConfig cfg = to_cfg(cfgp);
    cfg.eval_3D = true;
    FrameAllocator alloc(arena_bytes);
    PreprocessorCPU pp; TileBinnerCPU bn; SorterCPU sr; RasterizerCPU rs;
    PreprocessOutput pre = pp.process(g, cam, cfg, alloc, nullptr);
    BinningOutput bin = bn.bin(pre, N, cam, cfg, alloc);
    const long long K = bin.total_pairs;
    if (K > max_pairs) return -1;
    if (K > 0) {
        std::memcpy(keys, bin.keys_unsorted, K * 8);
        std::memcpy(values, bin.values_unsorted, K * 4);
        sr.sort(bin, alloc);
        std::vector<uint32_t> ord(K);
        for (long long i = 0; i < K; i++) ord[i] = (uint32_t)i;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](uint32_t a, uint32_t b) { return bin.keys_unsorted[a] < bin.keys_unsorted[b]; });
        for (long long i = 0; i < K; i++) {
            bin.keys_sorted[i] = bin.keys_unsorted[ord[i]];
            bin.values_sorted[i] = bin.values_unsorted[ord[i]];
        }
        std::memcpy(values_sorted, bin.values_sorted, K * 4);
    }
    std::memcpy(ranges, bin.tile_ranges, (size_t)bin.num_tiles * 8);
    ForwardCache cache{};
    rs.rasterize(pre, bin, cam, cfg, img, nullptr, &cache, &alloc);
    const int npix = cam.width * cam.height;
    std::memcpy(T_final, cache.T_final, npix * 4);
    std::memcpy(n_contrib, cache.n_contrib, npix * 4);
    for (int i = 0; i < N; i++) {
        const bool vis = pre.radii[i] > 0;
        radii[i] = pre.radii[i]; tiles[i] = pre.tiles_touched[i];
        means2D[2 * i] = vis ? pre.means2D[2 * i] : 0; means2D[2 * i + 1] = vis ? pre.means2D[2 * i + 1] : 0;
        depths[i] = vis ? pre.depths[i] : 0; op2d[i] = vis ? pre.opacities_2d[i] : 0;
        for (int k = 0; k < 3; k++) {
            rgb[3 * i + k] = vis ? pre.rgb[3 * i + k] : 0;
            moff[3 * i + k] = vis ? pre.mean_offset[3 * i + k] : 0;
        }
        for (int k = 0; k < 6; k++) cinv[6 * i + k] = vis ? pre.cov3D_inv[6 * i + k] : 0;
        for (int k = 0; k < 16; k++) g2s[16 * i + k] = vis ? pre.gauss2screen[16 * i + k] : 0;
    }
    return K;
}

// ---- DensityController and OwnedRawParams (density_controller.cpp:9-204,
// train_types.h:85-157) behind opaque handles, and Trainer::step_with_densify
// (trainer.cpp:72-182). ----
void* ref_params_new(int N, int sh_degree, int max_coeffs, const float* pos, const float* sc, const float* rot,
                     const float* sh, const float* op) {
    OwnedRawParams* p = new OwnedRawParams();
    p->sh_degree = sh_degree;
    p->max_coeffs = max_coeffs;
    // OwnedRawParams::append copies from pointers into the vectors it grows
    // (train_types.h:108-116): with this much capacity no push_back reallocates
    // during a densification, so those reads stay valid
    const size_t cap = (size_t)N * 16 + 64;
    p->positions.reserve(cap * 3);
    p->scales.reserve(cap * 3);
    p->rotations.reserve(cap * 4);
    p->sh_coeffs.reserve(cap * max_coeffs * 3);
    p->opacities.reserve(cap);
    p->resize(N);
    std::memcpy(p->positions.data(), pos, (size_t)N * 12);
    std::memcpy(p->scales.data(), sc, (size_t)N * 12);
    std::memcpy(p->rotations.data(), rot, (size_t)N * 16);
    std::memcpy(p->sh_coeffs.data(), sh, (size_t)N * max_coeffs * 12);
    std::memcpy(p->opacities.data(), op, (size_t)N * 4);
    return p;
}
void ref_params_free(void* h) { delete static_cast<OwnedRawParams*>(h); }
int ref_params_count(void* h) { return static_cast<OwnedRawParams*>(h)->count(); }
void ref_params_get(void* h, float* pos, float* sc, float* rot, float* sh, float* op) {
    OwnedRawParams* p = static_cast<OwnedRawParams*>(h);
    const size_t n = (size_t)p->count();
    std::memcpy(pos, p->positions.data(), n * 12);
    std::memcpy(sc, p->scales.data(), n * 12);
    std::memcpy(rot, p->rotations.data(), n * 16);
    std::memcpy(sh, p->sh_coeffs.data(), n * p->max_coeffs * 12);
    std::memcpy(op, p->opacities.data(), n * 4);
}
void* ref_dc_new() { return new DensityController(); }
void ref_dc_free(void* h) { delete static_cast<DensityController*>(h); }
void ref_dc_reset(void* h, int N) { static_cast<DensityController*>(h)->reset(N); }
void ref_dc_accumulate(void* h, const float* d_means2D, const int* radii, int N) {
    static_cast<DensityController*>(h)->accumulate_stats(d_means2D, radii, N);
}
// icfg: from, until, interval, opacity_reset_interval, max_screen_size, enabled;
// fcfg: grad_threshold, min_opacity, percent_dense
static DensifyConfig dcfg_of(const int* icfg, const float* fcfg) {
    DensifyConfig c;
    c.densify_from_iter = icfg[0]; c.densify_until_iter = icfg[1]; c.densify_interval = icfg[2];
    c.opacity_reset_interval = icfg[3]; c.max_screen_size = icfg[4]; c.enabled = icfg[5] != 0;
    c.grad_threshold = fcfg[0]; c.min_opacity = fcfg[1]; c.percent_dense = fcfg[2];
    return c;
}
int ref_dc_densify(void* h, void* params, const int* icfg, const float* fcfg, float extent) {
    return static_cast<DensityController*>(h)->densify_and_prune(*static_cast<OwnedRawParams*>(params),
                                                                  dcfg_of(icfg, fcfg), extent);
}
void ref_dc_reset_opacity(void* params) { DensityController::reset_opacity(*static_cast<OwnedRawParams*>(params)); }

void* ref_trainer_new(size_t arena_bytes) { return new Trainer(arena_bytes); }
void ref_trainer_free(void* h) { delete static_cast<Trainer*>(h); }
// tcf: lr_pos_init, lr_pos_final, lr_feature, lr_opacity, lr_scaling, lr_rotation, beta1, beta2, eps;
// tci: max_steps, use_adam.  Returns the loss; *n_out the count after the step.
float ref_trainer_step_densify(void* h, void* params, const gs_camera* camp, const gs_render_cfg* cfgp,
                               const float* gt, const float* tcf, const int* tci, const int* icfg, const float* fcfg,
                               float extent, int iteration, int* n_out) {
    TrainConfig tc;
    tc.lr_position_init = tcf[0]; tc.lr_position_final = tcf[1]; tc.lr_feature = tcf[2];
    tc.lr_opacity = tcf[3]; tc.lr_scaling = tcf[4]; tc.lr_rotation = tcf[5];
    tc.adam_beta1 = tcf[6]; tc.adam_beta2 = tcf[7]; tc.adam_eps = tcf[8];
    tc.max_steps = tci[0]; tc.use_adam = tci[1] != 0;
    Camera cam = to_cam(camp);
    RenderConfig cfg = to_cfg(cfgp);
    Trainer::StepResult r = static_cast<Trainer*>(h)->step_with_densify(
        *static_cast<OwnedRawParams*>(params), cam, gt, cfg, tc, dcfg_of(icfg, fcfg), extent, iteration);
    *n_out = r.n_gaussians;
    return r.loss;
}

}  // extern "C"
