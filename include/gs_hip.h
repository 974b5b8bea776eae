/*
 * gs_hip.h -- C ABI of the MI355X (gfx950) 3D Gaussian Splatting render-and-train
 * hot path.
 *
 * Every entry point is asynchronous on the context's HIP stream, takes device
 * pointers and plain sizes, and returns a status code (GS_OK == 0).  Nothing
 * here allocates caller-visible memory: outputs are caller-owned device
 * buffers; temporaries live in a workspace owned by the context.
 *
 * Data layout (identical to the reference's public structs, so host arrays of
 * the reference can be copied across unchanged):
 *   positions  float[N*3]            (GaussianData::positions, types.h:39)
 *   scales     float[N*3]            (types.h:41; raw: log-space, train_types.h:15)
 *   rotations  float[N*4] (r,x,y,z)  (types.h:42; raw: unnormalised)
 *   sh_coeffs  float[N*max_coeffs*3] interleaved RGB per basis (types.h:40)
 *   opacities  float[N]              (types.h:43; raw: logit)
 *   images     float[H*W*3] HWC
 *   keys       uint64[K] = tile_id << 32 | float_bits(view depth)
 *   values     uint32[K] = Gaussian index
 *   ranges     uint32[num_tiles*2] = [start, end) into the sorted pair list
 *
 * Array pointers are never null, also for an empty scene or view (N == 0,
 * K == 0): pass any valid device allocation there; it is not read or written
 * through a count of zero.  The raster entry points take no Gaussian count,
 * so a null per-Gaussian array is rejected (GS_ERR_INVALID) rather than
 * taken to mean an empty scene.  Entry points that take the count return at
 * once for N == 0, before their array checks.
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef GS_HIP_H
#define GS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define GS_OK 0
#define GS_ERR_INVALID 1     /* bad argument (null pointer, bad size, unsupported mode) */
#define GS_ERR_HIP 2         /* a HIP runtime call failed */
#define GS_ERR_CAPACITY 3    /* a pre-sized buffer was too small (pairs > capacity) */
#define GS_ERR_UNSUPPORTED 4 /* feature outside this path */

/* Camera: same fields and order as the reference Camera (types.h:48-54).
 * Matrices are column-major. */
typedef struct gs_camera {
    float view_matrix[16];
    float viewproj_matrix[16];
    float cam_pos[3];
    float tan_fovx, tan_fovy;
    int width, height;
} gs_camera;

/* The 2D-path fields of the reference RenderConfig (types.h:57-66).
 * eval_3D is requested by calling the *_3d entry points below. */
typedef struct gs_render_cfg {
    float bg_color[3];
    float scale_modifier;
    int antialiasing;
    int training;
    int sh_degree;
    int tile_w, tile_h; /* must be 16 x 16 */
} gs_render_cfg;

/* Gaussian parameter arrays.  raw != 0: pre-activation parameters
 * (RawGaussianParams, train_types.h:9-18), activated in-kernel exactly as
 * RawGaussianParams::activate (train_types.cpp:6-36).  raw == 0: activated
 * GaussianData (types.h:34-45). */
typedef struct gs_gaussians {
    int count;
    int sh_degree;
    int max_coeffs;
    int raw;
    const float* positions;
    const float* scales;
    const float* rotations;
    const float* sh_coeffs;
    const float* opacities;
} gs_gaussians;

/* Per-Gaussian preprocess outputs (PreprocessOutput, types.h:69-82, 2D path). */
typedef struct gs_preprocessed {
    float* means2D;   /* [N*2] */
    float* depths;    /* [N]   */
    float* conics;    /* [N*3] */
    float* opacities; /* [N]   */
    float* rgb;       /* [N*3] */
    int* radii;       /* [N]   */
    int* tiles_touched; /* [N] */
} gs_preprocessed;

/* The eval_3D arrays of PreprocessOutput (types.h:77-79): the per-Gaussian
 * gauss-to-screen matrix (row-major 4x4), the upper triangle of the inverse
 * dilated 3D covariance, and position - camera position. */
typedef struct gs_preprocessed_3d {
    float* gauss2screen; /* [N*16], 16-byte aligned */
    float* cov3D_inv;    /* [N*6], 8-byte aligned */
    float* mean_offset;  /* [N*3]  */
} gs_preprocessed_3d;

/* Screen-space gradients (RasterGradOutput, train_types.h:25-31). */
typedef struct gs_raster_grads {
    float* d_means2D;      /* [N*2] */
    float* d_conics;       /* [N*3] */
    float* d_rgb;          /* [N*3] */
    float* d_opacities_2d; /* [N]   */
} gs_raster_grads;

/* Raw-parameter gradients (GradientOutput, train_types.h:35-43). */
typedef struct gs_param_grads {
    float* d_positions;  /* [N*3] */
    float* d_scales;     /* [N*3] */
    float* d_rotations;  /* [N*4] */
    float* d_sh_coeffs;  /* [N*max_coeffs*3] */
    float* d_opacities;  /* [N] */
} gs_param_grads;

/* Per-pixel forward state kept for the backward (ForwardCache::T_final and
 * ::n_contrib, types.h:97-99, plus the sorted-list index + 1 of each pixel's
 * last contributor, which lets the backward start there without a replay). */
typedef struct gs_pixel_state {
    float* T_final;          /* [H*W] */
    int* n_contrib;          /* [H*W] */
    uint32_t* last_contrib;  /* [H*W] */
} gs_pixel_state;

typedef struct gs_ctx gs_ctx;

/* Context: device, stream and a growable device workspace.
 * Replaces OpenCLContext::init (gpu/opencl_context.cpp:13-80). */
int gs_ctx_create(int device, void* hip_stream, gs_ctx** out);
int gs_ctx_destroy(gs_ctx* ctx);
int gs_ctx_set_stream(gs_ctx* ctx, void* hip_stream);
int gs_ctx_synchronize(gs_ctx* ctx);
const char* gs_last_error(void);

/* Fills `bytes` bytes with zero using a kernel (safe inside graph capture). */
int gs_zero(gs_ctx* ctx, void* ptr, size_t bytes);

/* Activation only: RawGaussianParams::activate (train_types.cpp:6-36). */
int gs_activate(gs_ctx* ctx, const gs_gaussians* raw, float* positions, float* scales,
                float* rotations, float* sh_coeffs, float* opacities);

/* Preprocess forward: Preprocessor::process (preprocessor.h:7-12), 2D branch
 * of PreprocessorCPU::process (cpu/preprocessor_cpu.cpp:11-62,188-265) and of
 * the preprocess kernel (gpu/kernels/preprocess.cl:583-620).  Activation is
 * fused when g->raw != 0.  Culled Gaussians get zeros in every output. */
int gs_preprocess_fwd(gs_ctx* ctx, const gs_gaussians* g, const gs_camera* cam,
                      const gs_render_cfg* cfg, const gs_preprocessed* out);

/* Inclusive prefix sum of tiles_touched into uint32 offsets; the total pair
 * count is written to *total_dev (device uint32) and, if total_host != NULL,
 * copied to the host after a stream sync.  Replaces the host scan of
 * TileBinnerGPU::bin (gpu/tile_binner_gpu.cpp:282-292) and
 * tile_binner_cpu.cpp:33-41. */
int gs_scan_tiles(gs_ctx* ctx, int n, const int* tiles_touched, uint32_t* offsets,
                  uint32_t* total_dev, long long* total_host);

/* Duplicate-with-keys: TileBinner::bin (tile_binner.h:7-12), the scatter
 * kernel (gpu/kernels/scatter.cl:73-154) and tile_binner_cpu.cpp:59-99.
 * `offsets` is the inclusive scan from gs_scan_tiles.  Pairs at positions >=
 * capacity are dropped and GS_ERR_CAPACITY is reported by gs_check_overflow. */
int gs_duplicate_keys(gs_ctx* ctx, int n, const float* means2D, const float* depths,
                      const int* radii, const uint32_t* offsets, int width, int height,
                      int tile_w, int tile_h, uint64_t* keys, uint32_t* values,
                      uint32_t capacity);

/* Stable LSD radix sort of (uint64 key, uint32 value) pairs on key bits
 * [0, key_bits), radix_bits (4..8) bits per pass.  Inputs are left unchanged.
 * Replaces rs_histogram/rs_compute_offsets/rs_scatter_kv
 * (gpu/kernels/radix_sort_kv_cl.h:24-181) driven by SorterGPU::sort
 * (gpu/sorter_gpu.cpp:175-303).  Equal keys keep their input order
 * (std::stable_sort semantics). */
int gs_radix_sort_kv(gs_ctx* ctx, uint32_t count, const uint64_t* keys_in,
                     const uint32_t* values_in, uint64_t* keys_out, uint32_t* values_out,
                     int key_bits, int radix_bits);

/* Tile ranges from sorted keys: identify_tile_ranges
 * (gpu/kernels/radix_sort.cl:14-46) and sorter_cpu.cpp:28-43.  Empty tiles get
 * (0,0). */
int gs_tile_ranges(gs_ctx* ctx, uint32_t count, const uint64_t* keys_sorted, int num_tiles,
                   uint32_t* ranges);

/* Sorter::sort (sorter.h:7): radix sort on tile-id + depth bits, then tile
 * ranges. */
int gs_sort_pairs(gs_ctx* ctx, uint32_t count, const uint64_t* keys_in, const uint32_t* values_in,
                  uint64_t* keys_out, uint32_t* values_out, int num_tiles, uint32_t* ranges);

/* Depth-ordered binning for the render/train fast path: replaces
 * TileBinner::bin + Sorter::sort (tile_binner.h:7-12, sorter.h:7;
 * tile_binner_cpu.cpp:59-99, sorter_cpu.cpp:6-44) with the same result in
 * less traffic.  The visible Gaussians are stably sorted by the float bits of
 * their depth, their pairs are emitted in that order as 64-bit words
 * tile_id << 32 | gaussian_index, and the words are stably sorted by tile id.
 * Each tile's run then lists the same Gaussians in the same order as
 * gs_sort_pairs produces from the duplicate-with-keys pairs (ties in
 * (tile, depth) keep index order in both).  words[0..K) and ranges are
 * outputs; K (read back) must not exceed `capacity`. */
int gs_bin_sorted_words(gs_ctx* ctx, int n, const gs_preprocessed* pre, int width, int height,
                        uint64_t* words, uint32_t capacity, uint32_t* ranges, long long* total_host);

/* Binning by direct placement, the trainer's path: replaces TileBinner::bin +
 * Sorter::sort (tile_binner.h:7-12, sorter.h:7; tile_binner_cpu.cpp:59-99,
 * sorter_cpu.cpp:6-44).  values[0..K) and ranges are exactly the values_out
 * and ranges of gs_sort_pairs on the duplicate-with-keys pairs; the keys are
 * not produced.  The Gaussians are stably sorted by depth bits, then each
 * index is written once, straight to its slot: per-tile counts over groups of
 * depth-ordered Gaussians, scanned tile-major, give every pair's position.
 * K is read back and must not exceed `capacity`; `values` must be 16-byte
 * aligned.  Up to 255 x 255 tiles (GS_ERR_UNSUPPORTED beyond). */
int gs_bin_tiles(gs_ctx* ctx, int n, const gs_preprocessed* pre, int width, int height, uint32_t* values,
                 uint32_t capacity, uint32_t* ranges, long long* total_host);

/* Forward rasterization: Rasterizer::rasterize (rasterizer.h:7-15), the
 * rasterize kernel (gpu/kernels/rasterize.cl:10-274, 2D lines) and
 * RasterizerCPU::rasterize (cpu/rasterizer_cpu.cpp:137-198).  One workgroup
 * per 16x16 tile.  `px` may be NULL fields if not needed. */
int gs_raster_fwd(gs_ctx* ctx, int width, int height, const gs_preprocessed* pre,
                  const uint32_t* values_sorted, const uint32_t* ranges, const float bg[3],
                  float* image, const gs_pixel_state* px);

/* Rasterizer::rasterize's inverse-depth output, 2D path
 * (cpu/rasterizer_cpu.cpp:170-172,197): per pixel, sum(alpha T / depth) over
 * its contributors.  Run after gs_raster_fwd on the same inputs, with its
 * pixel state (last_contrib is required); inv_depth is [H*W]. */
int gs_raster_inv_depth(gs_ctx* ctx, int width, int height, const gs_preprocessed* pre,
                        const uint32_t* values_sorted, const uint32_t* ranges, const gs_pixel_state* px,
                        float* inv_depth);

/* gs_raster_fwd over the packed words of gs_bin_sorted_words. */
int gs_raster_fwd_words(gs_ctx* ctx, int width, int height, const gs_preprocessed* pre,
                        const uint64_t* words, const uint32_t* ranges, const float bg[3], float* image,
                        const gs_pixel_state* px);

/* ---- 3DGS PLY input: loadPly (include/ply_loader.h:12, src/ply_loader.cpp:32-189) ----
 * gs_ply_info parses the header: vertex count, SH degree and max_coeffs
 * (from the number of f_rest_* properties: 45/24/9 -> 3/2/1) and whether a
 * filter_3D property is present.  gs_ply_load reads the vertices into
 * device arrays ([N*3], [N*3], [N*4] 16-byte aligned, [N*max_coeffs*3], [N],
 * filter_3d [N] or NULL), activated as loadPly does (exp scale, sigmoid
 * opacity, normalised rotation) with the SH reordered to sh[k][ch].  n and
 * max_coeffs are the sizes the arrays were allocated for (from gs_ply_info);
 * a file whose header differs, or that has no filter_3D when filter_3d is
 * non-NULL, is refused with GS_ERR_INVALID before any device work.  Returns
 * after the data is on the device. */
int gs_ply_info(const char* path, int* count, int* sh_degree, int* max_coeffs, int* has_filter_3d);
int gs_ply_load(gs_ctx* ctx, const char* path, int n, int max_coeffs, float* positions, float* scales,
                float* rotations, float* sh_coeffs, float* opacities, float* filter_3d);
/* loadPly followed by initRawFromGaussianData (src/train_main.cpp:31-63):
 * the raw parameters a training run starts from (log of the activated scale,
 * inverse sigmoid of the clamped opacity, normalised rotations, SH as read).
 * Same arrays and checks as gs_ply_load; returns after the data is ready. */
int gs_ply_load_raw(gs_ctx* ctx, const char* path, int n, int max_coeffs, float* positions, float* log_scales,
                    float* rotations, float* sh_coeffs, float* opacity_logits);
/* savePly (src/train_main.cpp:82-146): raw (pre-activation) device arrays
 * written as a 3DGS PLY, SH reordered to the file's channel-major f_rest. */
int gs_ply_save(gs_ctx* ctx, const char* path, int n, int max_coeffs, const float* positions,
                const float* scales, const float* rotations, const float* sh_coeffs, const float* opacities);

/* ---- eval_3D forward (AAA-Gaussians 3D evaluation, StopThePop k-buffer) ----
 * The stage-by-stage path of Renderer::render with RenderConfig::eval_3D
 * set (types.h:61): preprocess -> gs_scan_tiles -> gs_duplicate_keys_3d ->
 * gs_sort_pairs -> gs_raster_fwd_3d.  Forward only, as in the reference.
 *
 * Preprocess, eval_3D branch: PreprocessorCPU::process
 * (cpu/preprocessor_cpu.cpp:64-187; math_utils.cpp:474-583 computeGauss2Screen,
 * :253-338 maxContribGaussianFrustum3D, :342-472 computeAABBScreen/View,
 * :590-613 computeCov3DInv) and the preprocess kernel's 3D branch
 * (gpu/kernels/preprocess.cl:394-582).  filter_3d (GaussianData::filter_3D,
 * types.h:44) may be NULL.  out->conics is written with zeros; culled
 * Gaussians get radius 0 and zeros elsewhere. */
int gs_preprocess_fwd_3d(gs_ctx* ctx, const gs_gaussians* g, const float* filter_3d, const gs_camera* cam,
                         const gs_render_cfg* cfg, const gs_preprocessed* out, const gs_preprocessed_3d* out3d);

/* Duplicate-with-keys, eval_3D: TileBinnerCPU::bin (cpu/tile_binner_cpu.cpp:14-20,
 * 59-99) and scatter.cl:118-145.  The low key bits are the depth along the
 * ray through the tile's centre (depthAlongRay, math_utils.cpp:616-630) where
 * that is > 0, else the view depth.  offsets as for gs_duplicate_keys. */
int gs_duplicate_keys_3d(gs_ctx* ctx, int n, const gs_preprocessed* pre, const gs_preprocessed_3d* pre3d,
                         const uint32_t* offsets, const gs_camera* cam, int tile_w, int tile_h,
                         uint64_t* keys, uint32_t* values, uint32_t capacity);

/* Raster forward, eval_3D: RasterizerCPU::rasterize's k-buffer branch
 * (cpu/rasterizer_cpu.cpp:16-21,61-135,185-193; rasterize.cl:148-222):
 * per pixel, the ray's Mahalanobis minimum gives alpha and the 16 nearest
 * entries by per-pixel depth are kept sorted and blended front to back.
 * T_final and n_contrib (ForwardCache, types.h:97-99) may be NULL.  Where
 * the view-projection matrix is singular (the reference then reads an
 * uninitialised inverse) the view depth orders the k-buffer, as it orders
 * the binner's keys. */
int gs_raster_fwd_3d(gs_ctx* ctx, const gs_camera* cam, const gs_preprocessed* pre,
                     const gs_preprocessed_3d* pre3d, const uint32_t* values_sorted, const uint32_t* ranges,
                     const float bg[3], float* image, float* T_final, int* n_contrib);
/* The fast path for gs_duplicate_keys_3d + gs_sort_pairs (tile_binner_cpu.cpp:
 * 59-99, sorter_cpu.cpp:6-43 on eval_3D keys): values and ranges equal theirs
 * bit for bit, and keys (optional, K entries) their sorted keys.  Direct
 * placement of every tile's pairs in Gaussian order, then a stable sort of
 * each tile's list by its per-tile depth key (in LDS for lists of up to 16384
 * pairs, by the radix sort beyond).  *total receives the pair count K; with K
 * above capacity nothing is placed and GS_ERR_CAPACITY is returned.  Views up
 * to 255 x 255 tiles.  Synchronises the stream (K is read back). */
int gs_bin_tiles_3d(gs_ctx* ctx, int n, const gs_preprocessed* pre, const gs_preprocessed_3d* pre3d,
                    const gs_camera* cam, uint32_t* values, uint64_t* keys, uint32_t capacity, uint32_t* ranges,
                    long long* total);

/* L1 loss and its image gradient: l1_loss / l1_loss_double (loss.cpp:4-30).
 * loss_dev: device double, the mean absolute difference (sum in FP64). */
int gs_l1_loss(gs_ctx* ctx, int width, int height, const float* rendered, const float* gt,
               float* d_image, double* loss_dev);

/* Backward rasterization: RasterizerBackwardCPU::backward
 * (cpu/rasterizer_backward_cpu.cpp:7-138) and rasterize_backward
 * (gpu/kernels/rasterize_backward.cl:23-203; the CPU recurrence is the one
 * followed).  Gradients are ACCUMULATED into `grads` (zero them first). */
int gs_raster_bwd(gs_ctx* ctx, int width, int height, const gs_preprocessed* pre,
                  const uint32_t* values_sorted, const uint32_t* ranges, const float bg[3],
                  const gs_pixel_state* px, const float* d_image, const gs_raster_grads* grads);

/* gs_raster_bwd over the packed words of gs_bin_sorted_words. */
int gs_raster_bwd_words(gs_ctx* ctx, int width, int height, const gs_preprocessed* pre,
                        const uint64_t* words, const uint32_t* ranges, const float bg[3],
                        const gs_pixel_state* px, const float* d_image, const gs_raster_grads* grads);

/* Preprocess backward: PreprocessorBackwardCPU::backward
 * (cpu/preprocessor_backward_cpu.cpp:205-645) and preprocess_backward
 * (gpu/kernels/preprocess_backward.cl:27-410; the CPU chain is the one
 * followed).  g must hold RAW parameters.  accumulate != 0 adds into `out`
 * (multi-view gradient accumulation); otherwise `out` is overwritten
 * (culled Gaussians get 0). */
int gs_preprocess_bwd(gs_ctx* ctx, const gs_gaussians* g, const gs_camera* cam,
                      const gs_render_cfg* cfg, const int* radii, const gs_raster_grads* rg,
                      const gs_param_grads* out, int accumulate);

/* Adam hyper-parameters for one step (TrainConfig, train_types.h:46-60;
 * AdamOptimizer::step, optimizer.cpp:60-101).  lr[] is per tensor in the order
 * positions, scales, rotations, sh_coeffs, opacities (positions already
 * scheduled); bc1/bc2 = 1/(1-beta^t). */
typedef struct gs_adam_hparams {
    float lr[5];
    float beta1, beta2, eps;
    float bc1, bc2;
} gs_adam_hparams;

/* Fused Adam over the five parameter tensors (one launch). */
int gs_adam_step(gs_ctx* ctx, int n, int max_coeffs, float* const params[5],
                 const float* const grads[5], float* const m[5], float* const v[5],
                 const gs_adam_hparams* hp);

/* SGDOptimizer::step (optimizer.cpp:9-26): p -= lr[k] * g over the five
 * tensors (used by TrainerGPU, trainer_gpu.h:44). */
int gs_sgd_step(gs_ctx* ctx, int n, int max_coeffs, float* const params[5], const float* const grads[5],
                const float lr[5]);

/* Host helpers with the reference's host arithmetic: lr_schedule
 * (train_types.cpp:64-68) and Adam bias corrections (optimizer.cpp:73-76). */
float gs_lr_schedule(float lr_init, float lr_final, int step, int max_steps);
void gs_adam_bias_corrections(float beta1, float beta2, int step, float* bc1, float* bc2);


/* ------------------------------------------------------------------------
 * Native training driver: the whole step (preprocess -> bin -> sort -> raster
 * -> L1 -> raster backward -> preprocess backward -> optimizer) on one stream,
 * with parameters, gradients and optimizer state resident in HBM.
 * ---------------------------------------------------------------------- */

/* TrainConfig (train_types.h:46-60). */
typedef struct gs_train_cfg {
    float lr_position_init, lr_position_final;
    float lr_feature, lr_opacity, lr_scaling, lr_rotation;
    int max_steps;
    int use_adam;
    float adam_beta1, adam_beta2, adam_eps;
} gs_train_cfg;

/* Trainer::StepResult (trainer.h:18) plus the pair count of the step. */
typedef struct gs_step_result {
    double loss;     /* mean over views for a batched step */
    int iteration;
    int n_gaussians;
    long long pairs; /* tile-Gaussian pairs (summed over views) */
} gs_step_result;

typedef struct gs_trainer gs_trainer;

int gs_trainer_create(gs_ctx* ctx, int n, int sh_degree, int max_coeffs, gs_trainer** out);
int gs_trainer_destroy(gs_trainer* tr);
/* Raw parameters in (from_host != 0: host arrays) / out (to_host != 0). */
int gs_trainer_set_params(gs_trainer* tr, const float* pos, const float* log_scales, const float* rot,
                          const float* sh, const float* opacity_logits, int from_host);
int gs_trainer_get_params(gs_trainer* tr, float* pos, float* log_scales, float* rot, float* sh,
                          float* opacity_logits, int to_host);
/* Device pointers of the trainer's resident buffers: which = "image",
 * "grad_pos", "grad_scale", "grad_rot", "grad_sh", "grad_op", "means2D",
 * "radii", ... (NULL if unknown). */
void* gs_trainer_buffer(gs_trainer* tr, const char* which);

/* Trainer::step (trainer.h:21-23) / TrainerGPU::step (trainer_gpu.h:27-29):
 * one single-view step.  The pair count is read back to size the pair
 * buffers, as the reference GPU stages do; gt is a device HWC image.  With
 * out == NULL (and no stage timing) the step returns without reading the
 * loss back or waiting for the stream, so consecutive steps run back to back;
 * the loss stays in gs_trainer_buffer(tr, "loss") (double) until the next. */
int gs_trainer_step(gs_trainer* tr, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt,
                    const gs_train_cfg* tc, int iteration, gs_step_result* out);

/* fused_adam != 0: with use_adam, the step applies Adam inside the
 * preprocess backward from the gradients held in registers (the same
 * arithmetic as gs_preprocess_bwd + gs_adam_step; the gradient buffers are
 * then not written).  Off by default, as Trainer::step writes the gradients
 * (trainer.cpp:44-62). */
int gs_trainer_set_options(gs_trainer* tr, int fused_adam);

/* Stage timing of gs_trainer_step (no counterpart in the reference; a
 * profiling hook).  For each stage whose bit is set in stage_mask, the step
 * records a HIP event pair around it on the context's stream and adds the
 * elapsed time to a running total; 0 turns timing off.  Each call resets the
 * totals.  gs_trainer_stage_times returns the per-step mean (ms) of the first
 * n stages and the number of steps timed. */
enum {
    GS_STAGE_PREPROCESS = 0, /* preprocess forward */
    GS_STAGE_BINNING = 1,    /* depth sort + direct placement (or the word sort) + ranges */
    GS_STAGE_RASTER_FWD = 2,
    GS_STAGE_LOSS = 3,       /* L1 and its image gradient */
    GS_STAGE_CLEAR = 4,      /* zeroing of the screen-space gradients */
    GS_STAGE_RASTER_BWD = 5,
    GS_STAGE_BACKWARD = 6,   /* preprocess backward + optimizer (fused or not) */
    GS_NUM_STAGES = 7
};
int gs_trainer_set_timing(gs_trainer* tr, unsigned stage_mask);
int gs_trainer_stage_times(gs_trainer* tr, double* ms, int n, long long* steps);

/* Renderer::render (renderer.h:17-21) on the trainer's parameters: forward
 * only, image into the trainer's "image" buffer. */
int gs_trainer_render(gs_trainer* tr, const gs_camera* cam, const gs_render_cfg* cfg, long long* pairs);

/* ------------------------------------------------------------------------
 * Batched multi-view step (no counterpart in the reference, which trains one
 * camera per Trainer::step; the gradients are the sum of the per-view
 * Trainer::step gradients and the loss the mean of the per-view losses).
 * One launch per stage over all views: multi-view preprocess reading the
 * parameters once, joint depth sort and tile sort, forward/backward raster
 * over the global (view, tile) list (persistent work queue), multi-view
 * preprocess backward with Adam fused in, optionally captured in a hipGraph.
 * ---------------------------------------------------------------------- */
typedef struct gs_batch gs_batch;

/* Views are fixed for the batch's lifetime (a new camera set = a new batch). */
int gs_batch_create(gs_trainer* tr, int n_views, const gs_camera* cams, const gs_render_cfg* cfg,
                    gs_batch** out);
int gs_batch_destroy(gs_batch* b);
/* persistent: raster over an atomic tile queue; fused_adam: the preprocess
 * backward applies Adam in place (else gradients go to the trainer's grad
 * buffers and a separate Adam kernel runs); use_graph: capture once, replay. */
int gs_batch_set_options(gs_batch* b, int persistent, int fused_adam, int use_graph);
/* Per-view HWC targets (host or device arrays, one per view). */
int gs_batch_set_targets(gs_batch* b, const float* const* targets, int from_host);
/* One batched step with Adam.  out == NULL: no host synchronisation (results
 * and the capacity check are read by a later gs_batch_result / step with out).
 * A step whose pairs exceed the pre-sized buffers updates nothing (its
 * backward and Adam kernels skip on the device's overflow word, and its Adam
 * step count is rolled back); the next check returns GS_ERR_CAPACITY, regrows
 * the buffers, and the caller repeats the steps that were not applied. */
int gs_batch_step(gs_batch* b, const gs_train_cfg* tc, int iteration, gs_step_result* out);
int gs_batch_result(gs_batch* b, gs_step_result* out, double* view_losses);
/* Forward only for all views into the batch's concatenated image buffer. */
int gs_batch_render(gs_batch* b, long long* pairs);
void* gs_batch_buffer(gs_batch* b, const char* which);
int gs_batch_info(gs_batch* b, int* total_tiles, long long* total_pixels, long long* capacity);

/* ------------------------------------------------------------------------
 * Autotuning (no counterpart in the reference, whose kernels use fixed
 * work-group sizes and a fixed 4-bit radix).  The tuner times the candidates
 * of each knob on the caller's own workload with HIP events on the library's
 * stream, one knob at a time (each knob only affects its own stage), and
 * leaves the fastest setting on the context.  It runs forward, loss and
 * backward stages only: parameters and optimizer state are not modified.
 * The tile size is not a knob: tile lists and ranges are defined on the
 * reference's 16 x 16 tiles, which parity requires.
 * ---------------------------------------------------------------------- */
typedef struct gs_tuning {
    int radix_bits;         /* depth-sort digit width: 4..8 */
    int raster_batch;       /* forward LDS batch depth: 64, 128, 256, 512 */
    int raster_bwd_batch;   /* backward LDS batch depth: 64, 128, 256 */
    int bin_group;          /* Gaussians per binning group: power of two, 256..65536 */
    int bin_tiles_per_lane; /* binning bin shape: 1, 2 or 4 (bin = 8 x 8*tpl tiles) */
    int persistent_per_cu;  /* batched raster: persistent groups per CU (0: one group per tile) */
    float binning_ms, raster_fwd_ms, raster_bwd_ms; /* best times measured */
    int split_groups;       /* single-view forward: groups of 1024 nearest Gaussians placed before the
                             * first raster round; the rest is placed only for bins with unfinished
                             * tiles (0: one round) */
    float forward_ms;       /* best forward (preprocess, binning, raster) measured with that split */
} gs_tuning;

/* Current knobs of a context (times zero) / set all knobs at once (for a
 * caller that stores a tuning and applies it later). */
int gs_ctx_get_tuning(gs_ctx* ctx, gs_tuning* out);
int gs_ctx_apply_tuning(gs_ctx* ctx, const gs_tuning* t);

/* Single-view tuning on the trainer's parameters for one camera; gt is the
 * device target image used to obtain an image gradient for the backward. */
int gs_autotune(gs_trainer* tr, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt, int reps,
                gs_tuning* out);

/* Batched tuning: binning knobs and the persistent grid / LDS batch depths of
 * the multi-view rasterizers, on the batch's views and targets.  The batch's
 * captured graph, if any, is rebuilt on the next step. */
int gs_batch_autotune(gs_batch* b, int reps, gs_tuning* out);

/* ------------------------------------------------------------------------
 * Adaptive density control on the trainer's device parameters
 * (DensityController, include/density_controller.h:7-40;
 * src/density_controller.cpp:9-204; Trainer::step_with_densify,
 * src/trainer.cpp:72-182).  The split noise continues the reference's own
 * generator (an LCG seeded 12345 with Box-Muller normals, :9-16) from the
 * trainer's state, so densifications reproduce the reference's parameters
 * bit for bit.  A densification changes the trainer's Gaussian count: buffers
 * returned by gs_trainer_buffer and batches created on the trainer before it
 * are invalid afterwards.
 * ---------------------------------------------------------------------- */
typedef struct gs_densify_cfg { /* DensifyConfig, include/train_types.h:71-81 (same defaults) */
    int densify_from_iter;      /* 500 */
    int densify_until_iter;     /* 15000 */
    int densify_interval;       /* 100 */
    float grad_threshold;       /* 0.0002 */
    float min_opacity;          /* 0.005 */
    float percent_dense;        /* 0.01; as in the reference, the clone/split size cut is 0.01 * extent */
    int opacity_reset_interval; /* 3000 */
    int max_screen_size;        /* 20 */
    int enabled;                /* 0 */
} gs_densify_cfg;

/* Gaussians currently held by the trainer. */
int gs_trainer_count(gs_trainer* tr, int* n);
/* DensityController::reset (:21-26): zero the statistics for the current count. */
int gs_trainer_density_reset(gs_trainer* tr);
/* DensityController::accumulate_stats (:29-41): adds |d_means2D| (device
 * float[N*2]) to each Gaussian's sum, 1 to its count, and raises its largest
 * screen radius to radii (device int[N], may be NULL).  Resets first when the
 * statistics are for another count. */
int gs_trainer_density_accumulate(gs_trainer* tr, const float* d_means2D, const int* radii);
/* DensityController::densify_and_prune (:148-197): clone, split (noise from
 * the generator), prune, compact, reset the statistics; the Adam moments are
 * zeroed and its step count reset when the count changes (trainer.cpp:159-163).
 * *new_n receives the new count (may be NULL). */
int gs_trainer_densify_and_prune(gs_trainer* tr, const gs_densify_cfg* cfg, float scene_extent, int* new_n);
/* DensityController::reset_opacity (:199-204): every opacity logit to
 * log(0.01 / 0.99). */
int gs_trainer_reset_opacity(gs_trainer* tr);
/* Trainer::step_with_densify (trainer.cpp:72-182): gs_trainer_step, then the
 * statistics of a forward and raster backward at the updated parameters, then
 * densification and opacity reset on their schedules.  out->n_gaussians is the
 * count after the step. */
int gs_trainer_step_densify(gs_trainer* tr, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt,
                            const gs_train_cfg* tc, const gs_densify_cfg* dcfg, float scene_extent, int iteration,
                            gs_step_result* out);

/* ------------------------------------------------------------------------
 * Camera files, images and the multi-view training loop (host side).
 * ---------------------------------------------------------------------- */
/* loadCamerasJson (src/train_main.cpp:262-377; src/main.cpp:146-264 reads one
 * camera of the same file): the 3DGS cameras.json -- objects with id, width,
 * height, fx, fy, position (camera centre) and rotation (camera-to-world
 * 3x3) -- as Cameras: view = [R^T | -R^T c], viewproj = perspective
 * (znear 0.01, zfar 100) * view, tan_fov = size / (2 f).  With cams NULL only
 * *count is set; img_names (optional) receives each camera's img_name in slots
 * of name_stride bytes. */
int gs_cameras_json(const char* path, gs_camera* cams, int max_cams, int* count, char* img_names, int name_stride);
/* readPPM / writePPM (src/image_io.cpp): binary 8-bit PPM (P6) to and from
 * host HWC float images in [0, 1].  gs_read_ppm with image NULL only reads the
 * size. */
int gs_read_ppm(const char* path, float* image, size_t max_floats, int* width, int* height);
int gs_write_ppm(const char* path, const float* image, int width, int height);
/* The training loop of src/train_main.cpp:633-673 on the trainer: each
 * iteration trains on one view, drawn by std::mt19937(seed) through
 * std::uniform_int_distribution over the views (view 0 when there is one;
 * the reference seeds 42), with iteration numbers from first_iteration;
 * every save_every iterations the raw parameters go to
 * "<output_path>.iter<N>.ply", and at the end to output_path (if not NULL).
 * targets: per-view device HWC images.  losses / views (optional, host,
 * `iterations` entries) receive each step's loss and view. */
int gs_trainer_train_views(gs_trainer* tr, int n_views, const gs_camera* cams, const float* const* targets,
                           const gs_render_cfg* cfg, const gs_train_cfg* tc, int first_iteration, int iterations,
                           unsigned seed, int save_every, const char* output_path, int log_every, double* losses,
                           int* views);

/* gs_scan_tiles without any allocation (no counterpart in the reference; for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same inclusive scan of
 * tiles_touched into offsets, and the pair count into total_dev (device
 * uint32, may be NULL), through a fixed per-device scratch of the library
 * (33 KB, held until the process ends) instead of the context's workspace;
 * no host read-back.  At most GS_NOALLOC_SCAN_MAX Gaussians (GS_ERR_INVALID
 * beyond).  Calls that overlap on one device share that scratch and race:
 * serialise them (one stream, or events). */
#define GS_NOALLOC_SCAN_MAX 16777216u
int gs_scan_tiles_noalloc(gs_ctx* ctx, int n, const int* tiles, uint32_t* offsets, uint32_t* total_dev);

/* ------------------------------------------------------------------------
 * Measurement (no counterpart in the reference).
 * ---------------------------------------------------------------------- */

/* The algorithmic work of the 2D rasterizers on one frame, counted on the
 * device by a replay with the reference's exact arithmetic
 * (cpu/rasterizer_cpu.cpp:141-176, cpu/rasterizer_backward_cpu.cpp:34-134):
 * counts[0] = pixel-Gaussian evaluations of the forward loop (up to and
 * including each pixel's stopping entry), counts[1] = entries from each tile's
 * start to each pixel's last contributor (what the backward evaluates),
 * counts[2] = the sum of n_contrib.  counts: device uint64[3], overwritten.
 * px comes from the forward of the same inputs. */
int gs_raster_work(gs_ctx* ctx, int width, int height, const gs_preprocessed* pre,
                   const uint32_t* values_sorted, const uint32_t* ranges, const gs_pixel_state* px,
                   unsigned long long* counts);

/* Measured ceilings of the context's device (synchronous; allocates and frees
 * 2 GiB for the copy): HBM copy bandwidth (read + write bytes), LDS read
 * bandwidth (conflict-free 16-byte reads), wave64 FP32 VALU issue rate
 * (multiply chains) and v_exp_f32 issue rate. */
typedef struct gs_peaks {
    int compute_units;
    double clock_mhz;
    double hbm_copy_gbs;
    double lds_read_gbs;
    double valu_wave_insts_per_s;
    double exp_wave_insts_per_s;
} gs_peaks;
int gs_device_peaks(gs_ctx* ctx, gs_peaks* out);

#ifdef __cplusplus
}
#endif
#endif /* GS_HIP_H */
