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
    DevBuf rg_means2D, rg_conics, rg_rgb, rg_op;
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
    gs::DevBuf cam_dev, hp_dev;  // device copies of the camera and Adam hyper-parameters (fused step)
    unsigned long long* k_pinned = nullptr;  // pair count of a pending binning (pinned host memory)
    hipEvent_t k_event = nullptr;             // recorded after its copy
    bool allow_async = true;
    std::vector<gs::GraphSlot> graphs;
    gs_gaussians gaussians() const;
    size_t tensor_elems(int k) const;
    ~gs_trainer();
};

extern "C" int gs_sgd_step(gs_ctx* ctx, int n, int max_coeffs, float* const params[5], const float* const grads[5],
                           const float lr[5]);
