// gs_trainer_impl.h -- the trainer's device-resident state (not public ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "gs_hip.h"
#include "gs_internal.h"

namespace gs {

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    float* f() const { return static_cast<float*>(p); }
    int* i() const { return static_cast<int*>(p); }
    uint32_t* u() const { return static_cast<uint32_t*>(p); }
};

// Everything one camera view needs: preprocess outputs, pixel state and
// screen-space gradients.
struct ViewBufs {
    DevBuf means2D, depths, conics, op2d, rgb, radii, tiles, offsets, ranges;
    DevBuf image, d_image, T_final, n_contrib, last, loss;
    DevBuf rg;  // screen-space gradients, one block: means2D (8 B/G), conics (12), rgb (12), opacities (4)
    DevBuf split_flags, split_rc;  // two-round forward: tile and bin flags; per-pixel colour
    size_t rg_n = 0;
    gs_preprocessed pre();
    gs_raster_grads rgrads();
    void release_all();
};

struct GraphSlot {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

// Optional HIP-event timing of a step's stages (gs_trainer_set_timing).
struct StageTimer {
    unsigned mask = 0;
    hipEvent_t beg[GS_NUM_STAGES] = {}, end[GS_NUM_STAGES] = {};
    double total_ms[GS_NUM_STAGES] = {};
    long long steps = 0;
    bool on(int s) const { return (mask >> s) & 1u; }
    hipError_t mark_begin(int s, hipStream_t st) { return on(s) ? hipEventRecord(beg[s], st) : hipSuccess; }
    hipError_t mark_end(int s, hipStream_t st) { return on(s) ? hipEventRecord(end[s], st) : hipSuccess; }
    hipError_t collect();  // after the step's stream synchronisation
    void release();
};

}  // namespace gs

struct gs_trainer {
    gs_ctx* ctx = nullptr;
    int n = 0, sh_degree = 0, max_coeffs = 1;
    int adam_steps = 0;
    gs::DevBuf param[5], grad[5], adam_m[5], adam_v[5];  // positions, scales, rotations, sh, opacities
    std::vector<gs::ViewBufs> views;
    gs::DevBuf keys, vals, keys_sorted, vals_sorted;
    int pair_stride = 1;  // uint32 entries per pair in keys_sorted: 1 (indices) or 2 (words)
    int fused_adam = 0;   // Adam applied inside the preprocess backward (gradients not written)
    unsigned long long* k_pinned = nullptr;      // pair count of a pending binning (mapped pinned host memory)
    unsigned long long* k_pinned_dev = nullptr;  // its device address
    hipEvent_t k_event = nullptr;             // recorded after its copy
    bool allow_async = true;
    gs::StageTimer timer;
    std::vector<gs::GraphSlot> graphs;
    // adaptive density control (DensityController): per-Gaussian statistics,
    // the split noise generator's state, and densification scratch
    gs::DevBuf dc_accum, dc_denom, dc_maxr;
    int dc_n = -1;                   // Gaussian count the statistics are for (-1: none yet)
    uint64_t dc_rng = 12345;         // DensityController::rng_state_ (density_controller.h:39)
    gs::DevBuf param_alt[5];         // the arrays a densification writes, swapped in afterwards
    gs::DevBuf dc_scratch, dc_map, dc_noise;
    gs_gaussians gaussians() const;
    size_t tensor_elems(int k) const;
    ~gs_trainer();
};

// Forward, L1 and raster backward of one view into views[0] (gs_trainer.cpp).
int tr_forward_raster_bwd(gs_trainer* t, const gs_camera* cam, const gs_render_cfg* cfg, const float* gt,
                          long long* pairs);

extern "C" int gs_sgd_step(gs_ctx* ctx, int n, int max_coeffs, float* const params[5], const float* const grads[5],
                           const float lr[5]);
