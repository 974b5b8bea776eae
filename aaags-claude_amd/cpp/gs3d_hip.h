// gs3d_hip.h -- HIP implementations of the reference's stage interfaces.
//
//   HipContext .............. replaces OpenCLContext (gpu/opencl_context.h)
//   PreprocessorHIP ......... Preprocessor      (PreprocessorGPU, gpu/preprocessor_gpu.h:25-37)
//   TileBinnerHIP ........... TileBinner        (TileBinnerGPU)
//   SorterHIP ............... Sorter            (SorterGPU)
//   RasterizerHIP ........... Rasterizer        (RasterizerGPU)
//   RasterizerBackwardHIP ... RasterizerBackwardGPU::backward (gpu/rasterizer_backward_gpu.h:16-19)
//   PreprocessorBackwardHIP . PreprocessorBackwardGPU::backward (gpu/preprocessor_backward_gpu.h:18-21)
//   AdamOptimizerHIP ........ AdamOptimizer::step (optimizer.h:13-14)
//   DensityControllerHIP .... DensityController (density_controller.h:7-40)
//   TrainerHIP .............. Trainer::step / step_with_densify / TrainerGPU::step (trainer.h:21-30,
//                             trainer_gpu.h:27-29)
//
// Conventions follow the GPU stages of the reference: each stage owns its
// device buffers and exposes them to the next stage through device_data
// (valid until the stage's next call); host outputs come from the caller's
// FrameAllocator.  Every stage also accepts inputs that carry no device_data
// (produced by CPU stages) and uploads their host arrays, so CPU and HIP
// stages mix freely.  Errors throw std::runtime_error with the library's
// message, as CL_CHECK does.
#pragma once
#include <memory>
#include <vector>

#include "gs3d_api.h"
#include "gs_hip.h"

class HipContext {
public:
    explicit HipContext(int device = 0, void* stream = nullptr);
    ~HipContext();
    HipContext(const HipContext&) = delete;
    HipContext& operator=(const HipContext&) = delete;
    gs_ctx* get() const { return ctx_; }
    void synchronize();

private:
    gs_ctx* ctx_ = nullptr;
};

void gs_check(int status, const char* what);  // throws std::runtime_error on failure

// l1_loss / l1_loss_double (loss.h:6-10) on the device of `ctx`: d_image is
// elementwise and bit-exact; the sum runs in double on the device, so the
// float result is the mean rounded once (the reference's float version sums
// serially in float).  The context-free declarations in gs3d_api.h use one
// process-wide context on device 0.
float l1_loss(HipContext& ctx, const float* rendered, const float* gt, int H, int W, float* d_image);
double l1_loss_double(HipContext& ctx, const float* rendered, const float* gt, int H, int W, float* d_image);

namespace gs3d_detail {
struct DeviceArray;  // grow-only device allocation
struct PreprocessDevice;
struct BinningDevice;
}  // namespace gs3d_detail

class PreprocessorHIP : public Preprocessor {
public:
    explicit PreprocessorHIP(HipContext& ctx);
    ~PreprocessorHIP() override;
    PreprocessorHIP(const PreprocessorHIP&) = delete;
    PreprocessorHIP& operator=(const PreprocessorHIP&) = delete;
    // Fills radii and tiles_touched on the host (as the reference GPU stage
    // does), and every other output array too when fill_host_outputs is set.
    // A ForwardCache is not filled: the HIP backward recomputes those values.
    PreprocessOutput process(const GaussianData& g, const Camera& cam, const RenderConfig& cfg, FrameAllocator& alloc,
                             ForwardCache* cache = nullptr) override;
    bool fill_host_outputs = true;

private:
    HipContext& ctx_;
    std::unique_ptr<gs3d_detail::PreprocessDevice> dev_;
};

class TileBinnerHIP : public TileBinner {
public:
    explicit TileBinnerHIP(HipContext& ctx);
    ~TileBinnerHIP() override;
    TileBinnerHIP(const TileBinnerHIP&) = delete;
    TileBinnerHIP& operator=(const TileBinnerHIP&) = delete;
    // Keys/values stay on the device (host pointers null, as TileBinnerGPU),
    // unless fill_host_outputs is set.
    BinningOutput bin(const PreprocessOutput& pre, int num_gaussians, const Camera& cam, const RenderConfig& cfg,
                      FrameAllocator& alloc) override;
    bool fill_host_outputs = false;

private:
    HipContext& ctx_;
    std::unique_ptr<gs3d_detail::BinningDevice> dev_;
};

class SorterHIP : public Sorter {
public:
    explicit SorterHIP(HipContext& ctx);
    ~SorterHIP() override;
    // Stable radix sort of the pairs on the device; tile ranges are copied to
    // binning.tile_ranges; keys_sorted / values_sorted are filled on the host
    // when those pointers are set.
    void sort(BinningOutput& binning, FrameAllocator& alloc) override;

private:
    HipContext& ctx_;
    std::unique_ptr<gs3d_detail::BinningDevice> own_;  // for inputs that came from the host
};

class RasterizerHIP : public Rasterizer {
public:
    explicit RasterizerHIP(HipContext& ctx);
    ~RasterizerHIP() override;
    // output_depth receives the inverse depth sum(alpha T / depth) of each
    // pixel's contributors (rasterizer_cpu.cpp:170-172,197); the eval_3D path
    // writes 0 there, as the reference's 3D branch leaves it at 0.
    void rasterize(const PreprocessOutput& pre, const BinningOutput& bin, const Camera& cam, const RenderConfig& cfg,
                   float* output_image, float* output_depth = nullptr, ForwardCache* cache = nullptr,
                   FrameAllocator* allocator = nullptr) override;

private:
    HipContext& ctx_;
    std::unique_ptr<gs3d_detail::DeviceArray> img_, T_, nc_, last_, depth_;
    std::unique_ptr<gs3d_detail::PreprocessDevice> pre_up_;
    std::unique_ptr<gs3d_detail::BinningDevice> bin_up_;
};

class RasterizerBackwardHIP {
public:
    explicit RasterizerBackwardHIP(HipContext& ctx);
    ~RasterizerBackwardHIP();
    // Replays the forward on the device to recover each pixel's last
    // contributor (as the reference CPU backward replays it), then runs the
    // back-to-front adjoint.  rgrad is overwritten (host arrays).
    void backward(const PreprocessOutput& pre, const BinningOutput& bin, const Camera& cam, const RenderConfig& cfg,
                  const ForwardCache& cache, const float* d_image, int num_gaussians, RasterGradOutput& rgrad);

private:
    HipContext& ctx_;
    std::unique_ptr<RasterizerHIP> fwd_;
    std::unique_ptr<gs3d_detail::DeviceArray> dimg_, grads_, scratch_img_, T_, nc_, last_;
    std::unique_ptr<gs3d_detail::PreprocessDevice> pre_up_;
    std::unique_ptr<gs3d_detail::BinningDevice> bin_up_;
};

class PreprocessorBackwardHIP {
public:
    explicit PreprocessorBackwardHIP(HipContext& ctx);
    ~PreprocessorBackwardHIP();
    // Gradients with respect to the raw parameters (host arrays, overwritten).
    void backward(const GaussianData& g, const Camera& cam, const RenderConfig& cfg, const ForwardCache& cache,
                  const RasterGradOutput& rgrad, const RawGaussianParams& raw, GradientOutput& grads);

private:
    HipContext& ctx_;
    std::unique_ptr<gs3d_detail::DeviceArray> buf_;
};

class AdamOptimizerHIP {
public:
    explicit AdamOptimizerHIP(HipContext& ctx);
    ~AdamOptimizerHIP();
    void step(RawGaussianParams& params, const GradientOutput& grads, const TrainConfig& cfg, int iteration);
    void reset(int N, int max_coeffs);
    int step_count() const { return step_count_; }

private:
    HipContext& ctx_;
    std::unique_ptr<gs3d_detail::DeviceArray> state_;
    int step_count_ = 0, last_N_ = 0, last_mc_ = 0;
};

// DensityController (include/density_controller.h:7-40) on the device: the
// statistics live in HBM; parameters cross PCIe per densification, as the
// reference's host arrays require.  The split noise continues the reference's
// generator, so results equal the reference's bit for bit.
class DensityControllerHIP {
public:
    explicit DensityControllerHIP(HipContext& ctx);
    ~DensityControllerHIP();
    DensityControllerHIP(const DensityControllerHIP&) = delete;
    DensityControllerHIP& operator=(const DensityControllerHIP&) = delete;
    void accumulate_stats(const float* d_means2D, const int* radii, int N);  // host arrays
    int densify_and_prune(OwnedRawParams& params, const DensifyConfig& cfg, float scene_extent);
    static void reset_opacity(OwnedRawParams& params);
    void reset(int N);
    int current_N() const { return current_N_; }

private:
    void ensure_state(int N, int sh_degree, int max_coeffs);
    HipContext& ctx_;
    gs_trainer* tr_ = nullptr;  // holds the device statistics and the parameters while densifying
    int deg_ = 0, mc_ = 1, current_N_ = 0;
    uint64_t rng_ = 12345;  // density_controller.h:39
    std::unique_ptr<gs3d_detail::DeviceArray> up_;
};

class TrainerHIP {
public:
    struct StepResult {
        float loss;
        int iteration;
        int n_gaussians;
    };
    explicit TrainerHIP(HipContext& ctx);
    ~TrainerHIP();
    TrainerHIP(const TrainerHIP&) = delete;
    TrainerHIP& operator=(const TrainerHIP&) = delete;
    // One step on host parameters (uploaded, stepped on the device, copied
    // back); optimizer state stays on the device between steps and resets
    // when the Gaussian count changes.
    StepResult step(RawGaussianParams& params, const Camera& camera, const float* gt_image, const RenderConfig& cfg,
                    const TrainConfig& train_cfg, int iteration);
    // Trainer::step_with_densify (trainer.h:26-30): the step, the statistics at
    // the updated parameters, densification and opacity reset on their
    // schedules; params is resized to the new count.
    StepResult step_with_densify(OwnedRawParams& params, const Camera& camera, const float* gt_image,
                                 const RenderConfig& cfg, const TrainConfig& train_cfg, const DensifyConfig& densify_cfg,
                                 float scene_extent, int iteration);
    const float* last_rendered() const { return last_rendered_.empty() ? nullptr : last_rendered_.data(); }

private:
    void ensure_trainer(const RawGaussianParams& params);
    HipContext& ctx_;
    gs_trainer* tr_ = nullptr;
    int n_ = -1, mc_ = 0, deg_ = -1;
    uint64_t rng_ = 12345;  // the density controller's generator, kept across trainers
    std::unique_ptr<gs3d_detail::DeviceArray> gt_;
    std::vector<float> last_rendered_;
};
