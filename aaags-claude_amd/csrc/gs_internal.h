// gs_internal.h -- host-side declarations shared by the C ABI layer, the
// kernel launchers and the C++ stage classes.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "gs_hip.h"

namespace gsd { struct CamArgs; }
namespace gsr { struct RasterArgs; struct RasterSplit; struct GradArgs; struct MultiRaster; struct MultiPix; struct MultiGrads; }
namespace gso { struct AdamArgs; struct AdamDev; }
namespace gsb { struct BinArgs; }

// Small fixed scratch owned by each context.
constexpr size_t GS_SMALL_BYTES = 4096;
constexpr size_t GS_SMALL_TOTAL64 = 0;    // uint64 pair total of the last scan
constexpr size_t GS_SMALL_OVERFLOW = 64;  // uint32 capacity-overflow flag

struct gs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void* ws = nullptr;  // pair-count-sized scratch (sorts, pair words)
    size_t ws_bytes = 0;
    void* ws2 = nullptr;  // Gaussian-count-sized scratch (depth order, ranked offsets)
    size_t ws2_bytes = 0;
    void* small = nullptr;
    bool capturing = false;
    int radix_bits = 8;
    int raster_batch = 256;
    int raster_bwd_batch = 256;
    int bin_chunk = 4096;   // Gaussians per group in direct-placement binning
#define GS_SPLIT_GROUP 1024     // Gaussians per group of the part placed before the first raster round
    int bin_tpl = 1;        // tiles per lane (bin = 8 x 8*tpl tiles)
    int persistent_per_cu = 6;  // batched raster: persistent groups per CU (0: one group per tile)
    int split_groups = 64;      // trainer forward: groups of GS_SPLIT_GROUP Gaussians placed before the
                                // first raster round (0: one round)
    void* bin_split = nullptr;  // the split left pending by the last gs_bin_tiles_impl (BinSplit, capi.hip)
};

int gs_set_error(int code, const char* fmt, ...);
int gs_hip_status(hipError_t e, const char* what);
int gs_ws_get(gs_ctx* c, size_t bytes, void** out);
int gs_ws2_get(gs_ctx* c, size_t bytes, void** out);
int gs_small_get(gs_ctx* c, void** out);
int check_cam_cfg(const gs_camera* cam, const gs_render_cfg* cfg);
int gs_tile_key_bits(int num_tiles);

extern "C" int gs_ctx_set_tuning(gs_ctx* c, int radix_bits, int raster_batch, int raster_bwd_batch);
extern "C" int gs_ctx_reserve(gs_ctx* c, size_t bytes);
extern "C" int gs_ctx_set_binning(gs_ctx* c, int gaussians_per_group, int tiles_per_lane);
int gs_bin_tiles_impl(gs_ctx* c, int n, const gs_preprocessed* pre, int width, int height, uint32_t* values,
                      uint32_t capacity, uint32_t* ranges, long long* total_host, unsigned long long* total_async,
                      hipEvent_t ev_total, int split = 0);
// After gs_bin_tiles_impl with 0 < split < groups, which placed only the
// first `split` depth groups: the description of that split for the forward
// raster (tile_flag and rc left to the caller), or sp->round = 0 when nothing
// is pending; then the placement of the remaining groups for the bins with a
// tile whose gate word is nonzero.
int gs_bin_split_info(gs_ctx* c, gsr::RasterSplit* sp, int* num_bins);
int gs_bin_place_rest(gs_ctx* c, const uint32_t* tile_gate);
int gs_raster_fwd_split(gs_ctx* c, int width, int height, const gs_preprocessed* pre, const uint32_t* values,
                        const uint32_t* ranges, const float bg[3], float* image, const gs_pixel_state* px,
                        const gsr::RasterSplit* sp);
extern "C" int gs_check_overflow(gs_ctx* c, int* overflowed);

// kernel launchers
hipError_t launch_preprocess_fwd(const gs_gaussians& G, const gsd::CamArgs& c, const gs_preprocessed& out, int deg,
                                 hipStream_t s);
hipError_t launch_activate(const gs_gaussians& G, float* pos, float* sc, float* rot, float* sh, float* op,
                           hipStream_t s);
hipError_t launch_preprocess_bwd(const gs_gaussians& G, const gsd::CamArgs& c, const int* radii,
                                 const gs_raster_grads& rg, const gs_param_grads& out, int accumulate, int sh_deg,
                                 hipStream_t s);
uint32_t scan_partials_needed(uint32_t n);
hipError_t launch_scan_tiles(const int* in, uint32_t n, uint32_t* out, unsigned long long* partial, uint32_t* total32,
                             unsigned long long* total64, hipStream_t s);
hipError_t launch_duplicate(int n, const float* means2D, const float* depths, const int* radii,
                            const uint32_t* offsets, int gx, int gy, uint64_t* keys, uint32_t* values,
                            uint32_t capacity, uint32_t* overflow, hipStream_t s);
size_t radix_workspace_bytes(uint32_t n, int radix_bits);
hipError_t launch_radix_sort(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out, uint32_t* vals_out,
                             uint32_t n, int key_bits, int radix_bits, void* workspace, hipStream_t s);
hipError_t launch_radix_sort_bits(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                                  uint32_t* vals_out, uint32_t n, int begin_bit, int end_bit, int radix_bits,
                                  void* workspace, hipStream_t s, const uint32_t* n_dev);
hipError_t launch_tile_ranges(const uint64_t* keys_sorted, uint32_t n, uint32_t num_tiles, uint32_t* ranges,
                              hipStream_t s, const uint32_t* n_dev = nullptr);
hipError_t launch_depth_keys(int n, const float* depths, const int* radii, uint64_t* keys, uint32_t* idx,
                             hipStream_t s);
hipError_t launch_depth_words(int n, const float* depths, const int* radii, uint64_t* words, hipStream_t s,
                              const float* means2D = nullptr, int gx = 0, int gy = 0, uint32_t* rects = nullptr);
hipError_t launch_gather_tiles(int n, const uint32_t* order, const int* tiles, int* out, hipStream_t s);
hipError_t launch_duplicate_ranked(int n, const uint32_t* order, const float* means2D, const int* radii,
                                   const uint32_t* offsets, int gx, int gy, uint64_t* words, uint32_t capacity,
                                   uint32_t* overflow, hipStream_t s);
hipError_t launch_raster_fwd(const gsr::RasterArgs& a, int num_tiles, int batch, float* img, float* T_final,
                             int* n_contrib, uint32_t* last_contrib, hipStream_t s,
                             const gsr::RasterSplit* split = nullptr);
hipError_t launch_raster_bwd(const gsr::RasterArgs& a, int num_tiles, int batch, const float* T_final,
                             const uint32_t* last_contrib, const float* d_image, const gsr::GradArgs& g,
                             hipStream_t s);
int l1_partials_needed(uint32_t n);
hipError_t launch_l1(const float* r, const float* g, float* d, uint32_t n, double* partial, double* loss,
                     hipStream_t s);
hipError_t launch_adam(const gso::AdamArgs& a, hipStream_t s);
// One thread: *counter += 1 when *flag is nonzero (counts steps whose update a
// capacity overflow suppressed).
hipError_t launch_count_if(const uint32_t* flag, uint32_t* counter, hipStream_t s);
hipError_t launch_sgd(const gso::AdamArgs& a, hipStream_t s);
hipError_t launch_zero(void* ptr, size_t bytes, hipStream_t s);

// multi-view
hipError_t launch_preprocess_fwd_multi(const gs_gaussians& G, const gsd::CamArgs* cams_dev, int V,
                                       const gs_preprocessed& out, uint64_t* dkeys, uint32_t* didx, int deg,
                                       hipStream_t s, uint32_t* rects = nullptr);
hipError_t launch_preprocess_bwd_multi(const gs_gaussians& G, const gsd::CamArgs* cams_dev,
                                       const gsd::CamArgs* cam_val, int V, const int* radii, const gs_raster_grads& rg,
                                       const gs_param_grads& out, float* const adam_p[5], float* const adam_m[5],
                                       float* const adam_v[5], const gso::AdamDev* hp, const gso::AdamDev* hp_val,
                                       int sh_deg, hipStream_t s, const uint32_t* skip = nullptr);
hipError_t launch_gather_tiles_multi(uint32_t total, int N, const uint64_t* skeys, const uint32_t* sidx,
                                     const int* tiles, int* out, hipStream_t s);
hipError_t launch_duplicate_multi(uint32_t total, int N, const uint64_t* skeys, const uint32_t* sidx,
                                  const float* means2D, const int* radii, const uint32_t* offsets, const int4* views,
                                  uint64_t* words, uint32_t capacity, uint32_t* overflow, hipStream_t s);
// direct-placement binning (k_bin.hip); larger views use the word sort
constexpr int GS_BIN_MAX_TILES_XY = 255;  // tiles per row and per column (4080 pixels)
int bin_blocks(int C, int gx, int gy, int tpl);
hipError_t launch_bin_count(const gsb::BinArgs& a, int blocks, int tpl, hipStream_t s);
hipError_t launch_bin_place(const gsb::BinArgs& a, int blocks, int tpl, hipStream_t s, int block0 = 0,
                            const uint32_t* gate = nullptr);
hipError_t launch_bin_ranges(const gsb::BinArgs& a, const uint32_t* scanned, uint32_t total_tiles, uint32_t* ranges,
                             hipStream_t s);
// total32 / total64 (optional, device): the sum of all n inputs
hipError_t launch_scan_u32_exclusive(const uint32_t* in, uint32_t n, uint32_t* out, unsigned long long* partial,
                                     hipStream_t s, uint32_t* total32 = nullptr, unsigned long long* total64 = nullptr);
hipError_t launch_tile_ranges_multi(const uint64_t* words, uint32_t n_cap, const uint32_t* n_dev, const int4* views,
                                    uint32_t* ranges, hipStream_t s);
hipError_t launch_raster_fwd_multi(const gsr::MultiRaster& m, const gsr::MultiPix& o, int batch,
                                   int persistent_groups, hipStream_t s);
hipError_t launch_raster_bwd_multi(const gsr::MultiRaster& m, const gsr::MultiPix& o, const gsr::MultiGrads& g,
                                   int batch, int persistent_groups, hipStream_t s);
// adaptive density control (k_density.hip)
hipError_t launch_density_accumulate(int n, const float* d_means2D, const int* radii, float* accum, float* denom,
                                     int* max_radii, hipStream_t s);
hipError_t launch_density_classify(int n, const float* accum, const float* denom, const float* log_scales,
                                   float threshold, float size_threshold, uint32_t* clone, uint32_t* split,
                                   hipStream_t s);
hipError_t launch_density_map(int n, const uint32_t* clone, const uint32_t* split, const uint32_t* rank_c,
                              const uint32_t* rank_s, uint32_t n_clone, uint32_t n_split, int2* map, hipStream_t s);
hipError_t launch_density_prune(int m1, int n, const int2* map, const float* log_scales, const float* opacity,
                                const int* max_radii, float scale_factor, float min_opacity, int max_screen_size,
                                float world_threshold, uint32_t* keep, hipStream_t s);
hipError_t launch_density_scatter(int m1, int max_coeffs, const int2* map, const uint32_t* keep, const uint32_t* slot,
                                  const float* const src[5], const float* noise, float scale_factor,
                                  float* const dst[5], hipStream_t s);
hipError_t launch_fill(int n, float v, float* out, hipStream_t s);
hipError_t launch_l1_multi(const float* r, const float* g, float* d, const int* pix_off_dev, int V, int nb,
                           double* partial, double* loss, hipStream_t s);
