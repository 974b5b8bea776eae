// gs3d_hip.cpp -- the reference's stage interfaces implemented over the C ABI.
#include "gs3d_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gs_internal.h"
#include "gs_trainer_impl.h"

void gs_check(int status, const char* what) {
    if (status != GS_OK) throw std::runtime_error(std::string(what) + ": " + gs_last_error());
}

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace gs3d_detail {

struct DeviceArray {
    void* p = nullptr;
    size_t bytes = 0;
    ~DeviceArray() {
        if (p) (void)hipFree(p);
    }
    // always leaves an allocation (256 bytes at least), so that arrays sized
    // by an empty scene are valid, non-null device pointers
    void* ensure(size_t need) {
        if (!p || need > bytes) {
            if (p) hip_check(hipFree(p), "hipFree");
            p = nullptr;
            bytes = 0;
            hip_check(hipMalloc(&p, std::max<size_t>(need, 256)), "hipMalloc");
            bytes = std::max<size_t>(need, 256);
        }
        return p;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Device side of a PreprocessOutput (what device_data points at).
struct PreprocessDevice {
    DeviceArray means2D, depths, conics, op2d, rgb, radii, tiles;
    DeviceArray pos, scale, rot, sh, op, filter;  // model upload
    DeviceArray g2s, cinv, moff;                   // eval_3D arrays
    bool eval3d = false;
    int n = 0;
    gs_preprocessed_3d view3d() const {
        return gs_preprocessed_3d{g2s.as<float>(), cinv.as<float>(), moff.as<float>()};
    }
    void alloc3d(int N) {
        size_t n = (size_t)N;
        g2s.ensure(n * 64);
        cinv.ensure(n * 24);
        moff.ensure(n * 12);
    }
    gs_preprocessed view() const {
        return gs_preprocessed{means2D.as<float>(), depths.as<float>(), conics.as<float>(), op2d.as<float>(),
                               rgb.as<float>(),     radii.as<int>(),    tiles.as<int>()};
    }
    void alloc(int N) {
        size_t n = (size_t)N;
        means2D.ensure(n * 8);
        depths.ensure(n * 4);
        conics.ensure(n * 12);
        op2d.ensure(n * 4);
        rgb.ensure(n * 12);
        radii.ensure(n * 4);
        tiles.ensure(n * 4);
        this->n = N;
    }
};

// Device side of a BinningOutput.
struct BinningDevice {
    DeviceArray offsets, keys, values, keys_sorted, values_sorted, ranges;
    uint32_t K = 0;
    int num_tiles = 0;
    bool sorted = false;
};

}  // namespace gs3d_detail

using namespace gs3d_detail;

static void upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes) hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync(H2D)");
}
static void download(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes) hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(D2H)");
}

static gs_camera to_gs(const Camera& c) {
    gs_camera g;
    std::memcpy(g.view_matrix, c.view_matrix, 64);
    std::memcpy(g.viewproj_matrix, c.viewproj_matrix, 64);
    std::memcpy(g.cam_pos, c.cam_pos, 12);
    g.tan_fovx = c.tan_fovx;
    g.tan_fovy = c.tan_fovy;
    g.width = c.width;
    g.height = c.height;
    return g;
}

static gs_render_cfg to_gs(const RenderConfig& c) {
    gs_render_cfg g;
    for (int i = 0; i < 3; i++) g.bg_color[i] = c.bg_color[i];
    g.scale_modifier = c.scale_modifier;
    g.antialiasing = c.antialiasing;
    g.training = c.training;
    g.sh_degree = c.sh_degree;
    g.tile_w = c.tile_w;
    g.tile_h = c.tile_h;
    return g;
}

// Device view of a PreprocessOutput: its own device_data, or an upload of the
// host arrays into `up`.
static const PreprocessDevice& pre_device(const PreprocessOutput& pre, int N, std::unique_ptr<PreprocessDevice>& up,
                                          hipStream_t s) {
    if (pre.device_data) return *static_cast<const PreprocessDevice*>(pre.device_data);
    if (!up) up.reset(new PreprocessDevice());
    up->alloc(N);
    size_t n = (size_t)N;
    upload(up->means2D.p, pre.means2D, n * 8, s);
    upload(up->depths.p, pre.depths, n * 4, s);
    upload(up->conics.p, pre.conics, n * 12, s);
    upload(up->op2d.p, pre.opacities_2d, n * 4, s);
    upload(up->rgb.p, pre.rgb, n * 12, s);
    upload(up->radii.p, pre.radii, n * 4, s);
    upload(up->tiles.p, pre.tiles_touched, n * 4, s);
    up->eval3d = pre.eval_3D && pre.cov3D_inv;
    if (up->eval3d) {
        up->alloc3d(N);
        if (pre.gauss2screen) upload(up->g2s.p, pre.gauss2screen, n * 64, s);
        upload(up->cinv.p, pre.cov3D_inv, n * 24, s);
        upload(up->moff.p, pre.mean_offset, n * 12, s);
    }
    return *up;
}

// ---------------------------------------------------------------- context --
HipContext::HipContext(int device, void* stream) { gs_check(gs_ctx_create(device, stream, &ctx_), "gs_ctx_create"); }
HipContext::~HipContext() { gs_ctx_destroy(ctx_); }
void HipContext::synchronize() { gs_check(gs_ctx_synchronize(ctx_), "gs_ctx_synchronize"); }

// ------------------------------------------------------------------- loss --
double l1_loss_double(HipContext& ctx, const float* rendered, const float* gt, int H, int W, float* d_image) {
    if (H < 0 || W < 0 || (H * W > 0 && (!rendered || !gt || !d_image)))
        throw std::runtime_error("l1_loss: bad arguments");
    hipStream_t s = ctx.get()->stream;
    const size_t n = (size_t)H * W * 3;
    DeviceArray r, g, d, loss;
    r.ensure(n * 4);
    g.ensure(n * 4);
    d.ensure(n * 4);
    loss.ensure(8);
    upload(r.p, rendered, n * 4, s);
    upload(g.p, gt, n * 4, s);
    gs_check(gs_l1_loss(ctx.get(), W, H, r.as<float>(), g.as<float>(), d.as<float>(), loss.as<double>()),
             "gs_l1_loss");
    double out = 0.0;
    download(d_image, d.p, n * 4, s);
    download(&out, loss.p, 8, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return out;
}

float l1_loss(HipContext& ctx, const float* rendered, const float* gt, int H, int W, float* d_image) {
    return (float)l1_loss_double(ctx, rendered, gt, H, W, d_image);
}

static HipContext& default_context() {
    static HipContext ctx(0);
    return ctx;
}

double l1_loss_double(const float* rendered, const float* gt, int H, int W, float* d_image) {
    return l1_loss_double(default_context(), rendered, gt, H, W, d_image);
}

float l1_loss(const float* rendered, const float* gt, int H, int W, float* d_image) {
    return l1_loss(default_context(), rendered, gt, H, W, d_image);
}

// ------------------------------------------------------------ preprocess --
PreprocessorHIP::PreprocessorHIP(HipContext& ctx) : ctx_(ctx), dev_(new PreprocessDevice()) {}
PreprocessorHIP::~PreprocessorHIP() = default;

PreprocessOutput PreprocessorHIP::process(const GaussianData& g, const Camera& cam, const RenderConfig& cfg,
                                          FrameAllocator& alloc, ForwardCache* cache) {
    (void)cache;
    hipStream_t s = ctx_.get()->stream;
    const int N = g.count;
    const size_t n = (size_t)N;
    PreprocessDevice& d = *dev_;
    d.alloc(N);
    d.pos.ensure(n * 12);
    d.scale.ensure(n * 12);
    d.rot.ensure(n * 16);
    d.sh.ensure(n * g.max_coeffs * 12);
    d.op.ensure(n * 4);
    upload(d.pos.p, g.positions, n * 12, s);
    upload(d.scale.p, g.scales, n * 12, s);
    upload(d.rot.p, g.rotations, n * 16, s);
    upload(d.sh.p, g.sh_coeffs, n * g.max_coeffs * 12, s);
    upload(d.op.p, g.opacities, n * 4, s);
    gs_gaussians gg{N, g.sh_degree, g.max_coeffs, 0, d.pos.as<float>(), d.scale.as<float>(), d.rot.as<float>(),
                    d.sh.as<float>(), d.op.as<float>()};
    gs_camera c = to_gs(cam);
    gs_render_cfg f = to_gs(cfg);
    gs_preprocessed out = d.view();
    d.eval3d = cfg.eval_3D;
    if (cfg.eval_3D) {
        // preprocessor_cpu.cpp:64-187; GaussianData::filter_3D is optional
        d.alloc3d(N);
        const float* flt = nullptr;
        if (g.filter_3D) {
            d.filter.ensure(n * 4);
            upload(d.filter.p, g.filter_3D, n * 4, s);
            flt = d.filter.as<float>();
        }
        gs_preprocessed_3d o3 = d.view3d();
        gs_check(gs_preprocess_fwd_3d(ctx_.get(), &gg, flt, &c, &f, &out, &o3), "gs_preprocess_fwd_3d");
    } else {
        gs_check(gs_preprocess_fwd(ctx_.get(), &gg, &c, &f, &out), "gs_preprocess_fwd");
    }

    PreprocessOutput o{};
    o.radii = alloc.allocate_array<int>(n);
    o.tiles_touched = alloc.allocate_array<int>(n);
    download(o.radii, d.radii.p, n * 4, s);
    download(o.tiles_touched, d.tiles.p, n * 4, s);
    if (fill_host_outputs) {
        o.means2D = alloc.allocate_array<float>(n * 2);
        o.depths = alloc.allocate_array<float>(n);
        o.conics = alloc.allocate_array<float>(n * 3);
        o.opacities_2d = alloc.allocate_array<float>(n);
        o.rgb = alloc.allocate_array<float>(n * 3);
        download(o.means2D, d.means2D.p, n * 8, s);
        download(o.depths, d.depths.p, n * 4, s);
        download(o.conics, d.conics.p, n * 12, s);
        download(o.opacities_2d, d.op2d.p, n * 4, s);
        download(o.rgb, d.rgb.p, n * 12, s);
    }
    o.gauss2screen = o.cov3D_inv = o.mean_offset = nullptr;
    o.eval_3D = cfg.eval_3D;
    if (cfg.eval_3D) {
        // types.h:77-79: allocated whenever eval_3D is on (preprocessor_cpu.cpp:25-29)
        o.gauss2screen = alloc.allocate_array<float>(n * 16);
        o.cov3D_inv = alloc.allocate_array<float>(n * 6);
        o.mean_offset = alloc.allocate_array<float>(n * 3);
        if (fill_host_outputs) {
            download(o.gauss2screen, d.g2s.p, n * 64, s);
            download(o.cov3D_inv, d.cinv.p, n * 24, s);
            download(o.mean_offset, d.moff.p, n * 12, s);
        }
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    o.device_data = dev_.get();
    return o;
}

// ---------------------------------------------------------------- binning --
TileBinnerHIP::TileBinnerHIP(HipContext& ctx) : ctx_(ctx), dev_(new BinningDevice()) {}
TileBinnerHIP::~TileBinnerHIP() = default;

BinningOutput TileBinnerHIP::bin(const PreprocessOutput& pre, int N, const Camera& cam, const RenderConfig& cfg,
                                 FrameAllocator& alloc) {
    if (cfg.tile_w != 16 || cfg.tile_h != 16) throw std::runtime_error("TileBinnerHIP: tiles must be 16x16");
    hipStream_t s = ctx_.get()->stream;
    std::unique_ptr<PreprocessDevice> up;
    const PreprocessDevice& pd = pre_device(pre, N, up, s);
    BinningDevice& d = *dev_;
    const int gx = (cam.width + 15) / 16, gy = (cam.height + 15) / 16;
    BinningOutput o{};
    o.num_tiles = gx * gy;
    o.tile_ranges = alloc.allocate_array<uint32_t>((size_t)o.num_tiles * 2);
    std::memset(o.tile_ranges, 0, (size_t)o.num_tiles * 8);
    d.num_tiles = o.num_tiles;
    d.sorted = false;
    long long K = 0;
    d.offsets.ensure((size_t)std::max(N, 1) * 4);
    gs_check(gs_scan_tiles(ctx_.get(), N, pd.tiles.as<int>(), d.offsets.as<uint32_t>(), nullptr, &K), "gs_scan_tiles");
    d.K = (uint32_t)K;
    d.keys.ensure((size_t)K * 8 + 8);
    d.values.ensure((size_t)K * 4 + 4);
    if (pre.eval_3D && pd.eval3d) {
        // tile_binner_cpu.cpp:78-89: per-tile depth keys
        gs_camera c = to_gs(cam);
        gs_preprocessed pv = pd.view();
        gs_preprocessed_3d p3 = pd.view3d();
        gs_check(gs_duplicate_keys_3d(ctx_.get(), N, &pv, &p3, d.offsets.as<uint32_t>(), &c, 16, 16,
                                      d.keys.as<uint64_t>(), d.values.as<uint32_t>(), (uint32_t)K),
                 "gs_duplicate_keys_3d");
    } else {
        gs_check(gs_duplicate_keys(ctx_.get(), N, pd.means2D.as<float>(), pd.depths.as<float>(), pd.radii.as<int>(),
                                   d.offsets.as<uint32_t>(), cam.width, cam.height, 16, 16, d.keys.as<uint64_t>(),
                                   d.values.as<uint32_t>(), (uint32_t)K),
                 "gs_duplicate_keys");
    }
    o.total_pairs = (int)K;
    o.keys_unsorted = o.keys_sorted = nullptr;
    o.values_unsorted = o.values_sorted = nullptr;
    if (fill_host_outputs && K > 0) {
        o.keys_unsorted = alloc.allocate_array<uint64_t>((size_t)K);
        o.values_unsorted = alloc.allocate_array<uint32_t>((size_t)K);
        o.keys_sorted = alloc.allocate_array<uint64_t>((size_t)K);
        o.values_sorted = alloc.allocate_array<uint32_t>((size_t)K);
        download(o.keys_unsorted, d.keys.p, (size_t)K * 8, s);
        download(o.values_unsorted, d.values.p, (size_t)K * 4, s);
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    o.device_data = dev_.get();
    return o;
}

// ------------------------------------------------------------------- sort --
SorterHIP::SorterHIP(HipContext& ctx) : ctx_(ctx) {}
SorterHIP::~SorterHIP() = default;

void SorterHIP::sort(BinningOutput& bin, FrameAllocator& alloc) {
    (void)alloc;
    hipStream_t s = ctx_.get()->stream;
    std::memset(bin.tile_ranges, 0, (size_t)bin.num_tiles * 8);
    BinningDevice* d = static_cast<BinningDevice*>(bin.device_data);
    const uint32_t K = (uint32_t)std::max(bin.total_pairs, 0);
    if (!d) {  // pairs produced on the host
        if (!own_) own_.reset(new BinningDevice());
        d = own_.get();
        d->K = K;
        d->num_tiles = bin.num_tiles;
        d->keys.ensure((size_t)K * 8 + 8);
        d->values.ensure((size_t)K * 4 + 4);
        upload(d->keys.p, bin.keys_unsorted, (size_t)K * 8, s);
        upload(d->values.p, bin.values_unsorted, (size_t)K * 4, s);
        bin.device_data = d;
    }
    d->keys_sorted.ensure((size_t)K * 8 + 8);
    d->values_sorted.ensure((size_t)K * 4 + 4);
    d->ranges.ensure((size_t)bin.num_tiles * 8);
    gs_check(gs_sort_pairs(ctx_.get(), K, d->keys.as<uint64_t>(), d->values.as<uint32_t>(), d->keys_sorted.as<uint64_t>(),
                           d->values_sorted.as<uint32_t>(), bin.num_tiles, d->ranges.as<uint32_t>()),
             "gs_sort_pairs");
    d->sorted = true;
    download(bin.tile_ranges, d->ranges.p, (size_t)bin.num_tiles * 8, s);
    if (bin.keys_sorted && K) download(bin.keys_sorted, d->keys_sorted.p, (size_t)K * 8, s);
    if (bin.values_sorted && K) download(bin.values_sorted, d->values_sorted.p, (size_t)K * 4, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// Device sorted values + ranges for a BinningOutput (uploaded if it is host-only).
static const BinningDevice& bin_device(const BinningOutput& bin, std::unique_ptr<BinningDevice>& up, hipStream_t s) {
    const BinningDevice* d = static_cast<const BinningDevice*>(bin.device_data);
    if (d && d->sorted) return *d;
    if (!up) up.reset(new BinningDevice());
    const uint32_t K = (uint32_t)std::max(bin.total_pairs, 0);
    up->K = K;
    up->num_tiles = bin.num_tiles;
    up->values_sorted.ensure((size_t)K * 4 + 4);
    up->ranges.ensure((size_t)bin.num_tiles * 8);
    if (K) {
        if (!bin.values_sorted) throw std::runtime_error("binning output has neither device data nor sorted values");
        upload(up->values_sorted.p, bin.values_sorted, (size_t)K * 4, s);
    }
    upload(up->ranges.p, bin.tile_ranges, (size_t)bin.num_tiles * 8, s);
    up->sorted = true;
    return *up;
}

// ------------------------------------------------------------------ raster --
RasterizerHIP::RasterizerHIP(HipContext& ctx)
    : ctx_(ctx), img_(new DeviceArray()), T_(new DeviceArray()), nc_(new DeviceArray()), last_(new DeviceArray()),
      depth_(new DeviceArray()) {}
RasterizerHIP::~RasterizerHIP() = default;

void RasterizerHIP::rasterize(const PreprocessOutput& pre, const BinningOutput& bin, const Camera& cam,
                              const RenderConfig& cfg, float* out_img, float* out_depth, ForwardCache* cache,
                              FrameAllocator* allocator) {
    hipStream_t s = ctx_.get()->stream;
    const size_t npix = (size_t)cam.width * cam.height;
    int N = 0;
    if (pre.device_data) N = static_cast<const PreprocessDevice*>(pre.device_data)->n;
    else if (cache && cache->pre) N = 0;
    if (!pre.device_data) {
        // the Gaussian count of a host-only input is the largest index referenced + 1
        uint32_t mx = 0;
        for (int i = 0; i < bin.total_pairs; i++) mx = std::max(mx, bin.values_sorted[i] + 1);
        N = (int)mx;
    }
    const PreprocessDevice& pd = pre_device(pre, N, pre_up_, s);
    const BinningDevice& bd = bin_device(bin, bin_up_, s);
    img_->ensure(npix * 12);
    T_->ensure(npix * 4);
    nc_->ensure(npix * 4);
    last_->ensure(npix * 4);
    gs_preprocessed pv = pd.view();
    if (pre.eval_3D && pd.eval3d) {
        // rasterizer_cpu.cpp:61-135: the k-buffer blend
        gs_camera c = to_gs(cam);
        gs_preprocessed_3d p3 = pd.view3d();
        gs_check(gs_raster_fwd_3d(ctx_.get(), &c, &pv, &p3, bd.values_sorted.as<uint32_t>(), bd.ranges.as<uint32_t>(),
                                  cfg.bg_color, img_->as<float>(), T_->as<float>(), nc_->as<int>()),
                 "gs_raster_fwd_3d");
        // rasterizer_cpu.cpp:53,197: the 3D branch accumulates no inverse depth
        if (out_depth) std::fill(out_depth, out_depth + npix, 0.0f);
    } else {
        gs_pixel_state px{T_->as<float>(), nc_->as<int>(), last_->as<uint32_t>()};
        gs_check(gs_raster_fwd(ctx_.get(), cam.width, cam.height, &pv, bd.values_sorted.as<uint32_t>(),
                               bd.ranges.as<uint32_t>(), cfg.bg_color, img_->as<float>(), &px),
                 "gs_raster_fwd");
        if (out_depth) {
            // rasterizer_cpu.cpp:170-172,197
            depth_->ensure(npix * 4);
            gs_check(gs_raster_inv_depth(ctx_.get(), cam.width, cam.height, &pv, bd.values_sorted.as<uint32_t>(),
                                         bd.ranges.as<uint32_t>(), &px, depth_->as<float>()),
                     "gs_raster_inv_depth");
            download(out_depth, depth_->p, npix * 4, s);
        }
    }
    download(out_img, img_->p, npix * 12, s);
    if (cache && allocator) {
        cache->T_final = allocator->allocate_array<float>(npix);
        cache->n_contrib = allocator->allocate_array<int>(npix);
        download(cache->T_final, T_->p, npix * 4, s);
        download(cache->n_contrib, nc_->p, npix * 4, s);
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// ---------------------------------------------------------- raster backward --
RasterizerBackwardHIP::RasterizerBackwardHIP(HipContext& ctx)
    : ctx_(ctx), dimg_(new DeviceArray()), grads_(new DeviceArray()), scratch_img_(new DeviceArray()),
      T_(new DeviceArray()), nc_(new DeviceArray()), last_(new DeviceArray()) {}
RasterizerBackwardHIP::~RasterizerBackwardHIP() = default;

void RasterizerBackwardHIP::backward(const PreprocessOutput& pre, const BinningOutput& bin, const Camera& cam,
                                     const RenderConfig& cfg, const ForwardCache& cache, const float* d_image,
                                     int N, RasterGradOutput& rgrad) {
    if (cfg.eval_3D) throw std::runtime_error("RasterizerBackwardHIP: eval_3D has no backward (the reference's 3D path is forward-only)");
    (void)cache;
    hipStream_t s = ctx_.get()->stream;
    const size_t npix = (size_t)cam.width * cam.height;
    const size_t n = (size_t)N;
    const PreprocessDevice& pd = pre_device(pre, N, pre_up_, s);
    const BinningDevice& bd = bin_device(bin, bin_up_, s);
    scratch_img_->ensure(npix * 12);
    T_->ensure(npix * 4);
    nc_->ensure(npix * 4);
    last_->ensure(npix * 4);
    dimg_->ensure(npix * 12);
    grads_->ensure(n * 36);
    gs_preprocessed pv = pd.view();
    gs_pixel_state px{T_->as<float>(), nc_->as<int>(), last_->as<uint32_t>()};
    gs_check(gs_raster_fwd(ctx_.get(), cam.width, cam.height, &pv, bd.values_sorted.as<uint32_t>(),
                           bd.ranges.as<uint32_t>(), cfg.bg_color, scratch_img_->as<float>(), &px),
             "gs_raster_fwd (replay)");
    upload(dimg_->p, d_image, npix * 12, s);
    float* g = grads_->as<float>();
    gs_raster_grads rg{g, g + n * 2, g + n * 5, g + n * 8};
    gs_check(gs_zero(ctx_.get(), g, n * 36), "gs_zero");
    gs_check(gs_raster_bwd(ctx_.get(), cam.width, cam.height, &pv, bd.values_sorted.as<uint32_t>(),
                           bd.ranges.as<uint32_t>(), cfg.bg_color, &px, dimg_->as<float>(), &rg),
             "gs_raster_bwd");
    download(rgrad.d_means2D, rg.d_means2D, n * 8, s);
    download(rgrad.d_conics, rg.d_conics, n * 12, s);
    download(rgrad.d_rgb, rg.d_rgb, n * 12, s);
    download(rgrad.d_opacities_2d, rg.d_opacities_2d, n * 4, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// ------------------------------------------------------ preprocess backward --
PreprocessorBackwardHIP::PreprocessorBackwardHIP(HipContext& ctx) : ctx_(ctx), buf_(new DeviceArray()) {}
PreprocessorBackwardHIP::~PreprocessorBackwardHIP() = default;

void PreprocessorBackwardHIP::backward(const GaussianData& g, const Camera& cam, const RenderConfig& cfg,
                                       const ForwardCache& cache, const RasterGradOutput& rgrad,
                                       const RawGaussianParams& raw, GradientOutput& grads) {
    if (cfg.eval_3D) throw std::runtime_error("PreprocessorBackwardHIP: eval_3D has no backward (the reference's 3D path is forward-only)");
    if (!cache.pre || !cache.pre->radii) throw std::runtime_error("PreprocessorBackwardHIP: cache.pre->radii needed");
    hipStream_t s = ctx_.get()->stream;
    const size_t n = (size_t)g.count;
    const size_t mc3 = (size_t)raw.max_coeffs * 3;
    // layout: raw params (59 floats at SH3), radii, raster grads (9), outputs (59)
    const size_t per = 3 + 3 + 4 + mc3 + 1;
    buf_->ensure(n * (2 * per + 9 + 1) * 4 + 1024);
    float* base = buf_->as<float>();
    auto aligned = [&](float* p) { return reinterpret_cast<float*>((reinterpret_cast<uintptr_t>(p) + 15) & ~uintptr_t(15)); };
    float* pos = base;
    float* sc = aligned(pos + n * 3);
    float* rot = aligned(sc + n * 3);
    float* sh = aligned(rot + n * 4);
    float* op = aligned(sh + n * mc3);
    int* radii = reinterpret_cast<int*>(aligned(op + n));
    float* dm = aligned(reinterpret_cast<float*>(radii + n));
    float* dc = aligned(dm + n * 2);
    float* drgb = aligned(dc + n * 3);
    float* dop = aligned(drgb + n * 3);
    float* gpos = aligned(dop + n);
    float* gsc = aligned(gpos + n * 3);
    float* grot = aligned(gsc + n * 3);
    float* gsh = aligned(grot + n * 4);
    float* gop = aligned(gsh + n * mc3);
    upload(pos, raw.raw_positions, n * 12, s);
    upload(sc, raw.raw_scales, n * 12, s);
    upload(rot, raw.raw_rotations, n * 16, s);
    upload(sh, raw.raw_sh_coeffs, n * mc3 * 4, s);
    upload(op, raw.raw_opacities, n * 4, s);
    upload(radii, cache.pre->radii, n * 4, s);
    upload(dm, rgrad.d_means2D, n * 8, s);
    upload(dc, rgrad.d_conics, n * 12, s);
    upload(drgb, rgrad.d_rgb, n * 12, s);
    upload(dop, rgrad.d_opacities_2d, n * 4, s);
    gs_gaussians gg{g.count, raw.sh_degree, raw.max_coeffs, 1, pos, sc, rot, sh, op};
    gs_camera c = to_gs(cam);
    gs_render_cfg f = to_gs(cfg);
    gs_raster_grads rg{dm, dc, drgb, dop};
    gs_param_grads pg{gpos, gsc, grot, gsh, gop};
    gs_check(gs_preprocess_bwd(ctx_.get(), &gg, &c, &f, radii, &rg, &pg, 0), "gs_preprocess_bwd");
    download(grads.d_raw_positions, gpos, n * 12, s);
    download(grads.d_raw_scales, gsc, n * 12, s);
    download(grads.d_raw_rotations, grot, n * 16, s);
    download(grads.d_raw_sh_coeffs, gsh, n * mc3 * 4, s);
    download(grads.d_raw_opacities, gop, n * 4, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// -------------------------------------------------------------------- Adam --
AdamOptimizerHIP::AdamOptimizerHIP(HipContext& ctx) : ctx_(ctx), state_(new DeviceArray()) {}
AdamOptimizerHIP::~AdamOptimizerHIP() = default;

void AdamOptimizerHIP::reset(int N, int max_coeffs) {
    const size_t per = 3 + 3 + 4 + (size_t)max_coeffs * 3 + 1;
    // params, grads, m, v for the five tensors, each tensor 16-byte aligned
    state_->ensure((size_t)N * per * 16 + 4096);
    gs_check(gs_zero(ctx_.get(), state_->p, state_->bytes), "gs_zero");
    step_count_ = 0;
    last_N_ = N;
    last_mc_ = max_coeffs;
}

void AdamOptimizerHIP::step(RawGaussianParams& params, const GradientOutput& grads, const TrainConfig& cfg,
                            int iteration) {
    const int N = params.count;
    if (N != last_N_ || params.max_coeffs != last_mc_) reset(N, params.max_coeffs);  // optimizer.cpp:66-67
    hipStream_t s = ctx_.get()->stream;
    const size_t sizes[5] = {(size_t)N * 3, (size_t)N * 3, (size_t)N * 4, (size_t)N * params.max_coeffs * 3,
                             (size_t)N};
    float* host_p[5] = {params.raw_positions, params.raw_scales, params.raw_rotations, params.raw_sh_coeffs,
                        params.raw_opacities};
    const float* host_g[5] = {grads.d_raw_positions, grads.d_raw_scales, grads.d_raw_rotations, grads.d_raw_sh_coeffs,
                              grads.d_raw_opacities};
    float* P[5];
    float* G[5];
    float* M[5];
    float* V[5];
    char* w = static_cast<char*>(state_->p);
    auto take = [&](size_t elems) {
        float* r = reinterpret_cast<float*>(w);
        w += (elems * 4 + 255) & ~size_t(255);
        return r;
    };
    for (int k = 0; k < 5; k++) M[k] = take(sizes[k]);
    for (int k = 0; k < 5; k++) V[k] = take(sizes[k]);
    for (int k = 0; k < 5; k++) P[k] = take(sizes[k]);
    for (int k = 0; k < 5; k++) G[k] = take(sizes[k]);
    for (int k = 0; k < 5; k++) {
        upload(P[k], host_p[k], sizes[k] * 4, s);
        upload(G[k], host_g[k], sizes[k] * 4, s);
    }
    step_count_++;
    gs_adam_hparams hp;
    hp.lr[0] = lr_schedule(cfg.lr_position_init, cfg.lr_position_final, iteration, cfg.max_steps);
    hp.lr[1] = cfg.lr_scaling;
    hp.lr[2] = cfg.lr_rotation;
    hp.lr[3] = cfg.lr_feature;
    hp.lr[4] = cfg.lr_opacity;
    hp.beta1 = cfg.adam_beta1;
    hp.beta2 = cfg.adam_beta2;
    hp.eps = cfg.adam_eps;
    gs_adam_bias_corrections(hp.beta1, hp.beta2, step_count_, &hp.bc1, &hp.bc2);
    const float* Gc[5] = {G[0], G[1], G[2], G[3], G[4]};
    gs_check(gs_adam_step(ctx_.get(), N, params.max_coeffs, P, Gc, M, V, &hp), "gs_adam_step");
    for (int k = 0; k < 5; k++) download(host_p[k], P[k], sizes[k] * 4, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// ----------------------------------------------------------------- trainer --
// ------------------------------------------------------- density control --
DensityControllerHIP::DensityControllerHIP(HipContext& ctx) : ctx_(ctx), up_(new DeviceArray()) {}
DensityControllerHIP::~DensityControllerHIP() {
    if (tr_) gs_trainer_destroy(tr_);
}

// A device state for N Gaussians of the given SH layout; statistics already
// gathered for the same N survive a change of layout.
void DensityControllerHIP::ensure_state(int N, int sh_degree, int max_coeffs) {
    if (tr_ && tr_->n == N && mc_ == max_coeffs && deg_ == sh_degree) return;
    std::vector<float> keep[2];
    std::vector<int> keep_r;
    const bool carry = tr_ && tr_->n == N && tr_->dc_n == N && N > 0;
    hipStream_t s = ctx_.get()->stream;
    if (carry) {
        keep[0].resize(N);
        keep[1].resize(N);
        keep_r.resize(N);
        download(keep[0].data(), tr_->dc_accum.p, (size_t)N * 4, s);
        download(keep[1].data(), tr_->dc_denom.p, (size_t)N * 4, s);
        download(keep_r.data(), tr_->dc_maxr.p, (size_t)N * 4, s);
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    if (tr_) gs_trainer_destroy(tr_);
    tr_ = nullptr;
    gs_check(gs_trainer_create(ctx_.get(), N, sh_degree, max_coeffs, &tr_), "gs_trainer_create");
    deg_ = sh_degree;
    mc_ = max_coeffs;
    gs_check(gs_trainer_density_reset(tr_), "gs_trainer_density_reset");
    if (carry) {
        upload(tr_->dc_accum.p, keep[0].data(), (size_t)N * 4, s);
        upload(tr_->dc_denom.p, keep[1].data(), (size_t)N * 4, s);
        upload(tr_->dc_maxr.p, keep_r.data(), (size_t)N * 4, s);
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
}

void DensityControllerHIP::reset(int N) {
    if (N < 0) throw std::runtime_error("DensityControllerHIP::reset: negative count");
    current_N_ = N;
    if (tr_ && tr_->n == N) gs_check(gs_trainer_density_reset(tr_), "gs_trainer_density_reset");
    else ensure_state(N, deg_, mc_);
}

void DensityControllerHIP::accumulate_stats(const float* d_means2D, const int* radii, int N) {
    if (N != current_N_ || !tr_ || tr_->n != N) reset(N);  // density_controller.cpp:30
    if (N == 0) return;
    hipStream_t s = ctx_.get()->stream;
    up_->ensure((size_t)N * 12);
    upload(up_->p, d_means2D, (size_t)N * 8, s);
    const int* rd = nullptr;
    if (radii) {
        upload(static_cast<char*>(up_->p) + (size_t)N * 8, radii, (size_t)N * 4, s);
        rd = reinterpret_cast<const int*>(static_cast<char*>(up_->p) + (size_t)N * 8);
    }
    gs_check(gs_trainer_density_accumulate(tr_, up_->as<float>(), rd), "gs_trainer_density_accumulate");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int DensityControllerHIP::densify_and_prune(OwnedRawParams& p, const DensifyConfig& cfg, float scene_extent) {
    const int N = p.count();
    if (N == 0) return 0;
    if (N != current_N_) reset(N);  // statistics for another count do not apply
    ensure_state(N, p.sh_degree, p.max_coeffs);
    gs_check(gs_trainer_set_params(tr_, p.positions.data(), p.scales.data(), p.rotations.data(), p.sh_coeffs.data(),
                                   p.opacities.data(), 1),
             "gs_trainer_set_params");
    gs_densify_cfg c{cfg.densify_from_iter, cfg.densify_until_iter, cfg.densify_interval, cfg.grad_threshold,
                     cfg.min_opacity, cfg.percent_dense, cfg.opacity_reset_interval, cfg.max_screen_size,
                     cfg.enabled ? 1 : 0};
    tr_->dc_rng = rng_;
    int M = 0;
    gs_check(gs_trainer_densify_and_prune(tr_, &c, scene_extent, &M), "gs_trainer_densify_and_prune");
    rng_ = tr_->dc_rng;
    p.resize(M);
    if (M > 0)
        gs_check(gs_trainer_get_params(tr_, p.positions.data(), p.scales.data(), p.rotations.data(),
                                       p.sh_coeffs.data(), p.opacities.data(), 1),
                 "gs_trainer_get_params");
    current_N_ = M;
    return M;
}

// density_controller.cpp:199-204 (host arrays, as the reference's)
void DensityControllerHIP::reset_opacity(OwnedRawParams& p) {
    const float v = std::log(0.01f / 0.99f);
    std::fill(p.opacities.begin(), p.opacities.end(), v);
}

TrainerHIP::TrainerHIP(HipContext& ctx) : ctx_(ctx), gt_(new DeviceArray()) {}
TrainerHIP::~TrainerHIP() {
    if (tr_) gs_trainer_destroy(tr_);
}

TrainerHIP::StepResult TrainerHIP::step(RawGaussianParams& params, const Camera& camera, const float* gt_image,
                                        const RenderConfig& cfg, const TrainConfig& tc, int iteration) {
    if (cfg.eval_3D) throw std::runtime_error("TrainerHIP: eval_3D has no backward (the reference's 3D path is forward-only)");
    hipStream_t s = ctx_.get()->stream;
    ensure_trainer(params);
    gs_check(gs_trainer_set_params(tr_, params.raw_positions, params.raw_scales, params.raw_rotations,
                                   params.raw_sh_coeffs, params.raw_opacities, 1),
             "gs_trainer_set_params");
    const size_t npix = (size_t)camera.width * camera.height;
    gt_->ensure(npix * 12);
    upload(gt_->p, gt_image, npix * 12, s);
    gs_camera c = to_gs(camera);
    gs_render_cfg f = to_gs(cfg);
    gs_train_cfg t{tc.lr_position_init, tc.lr_position_final, tc.lr_feature, tc.lr_opacity, tc.lr_scaling,
                   tc.lr_rotation, tc.max_steps, tc.use_adam ? 1 : 0, tc.adam_beta1, tc.adam_beta2, tc.adam_eps};
    gs_step_result r;
    gs_check(gs_trainer_step(tr_, &c, &f, gt_->as<float>(), &t, iteration, &r), "gs_trainer_step");
    gs_check(gs_trainer_get_params(tr_, params.raw_positions, params.raw_scales, params.raw_rotations,
                                   params.raw_sh_coeffs, params.raw_opacities, 1),
             "gs_trainer_get_params");
    last_rendered_.resize(npix * 3);
    download(last_rendered_.data(), gs_trainer_buffer(tr_, "image"), npix * 12, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return StepResult{(float)r.loss, iteration, params.count};
}

// A device trainer for this count and SH layout (optimizer state restarts
// when they change, as AdamOptimizer::step does); the density generator's
// state carries over.
void TrainerHIP::ensure_trainer(const RawGaussianParams& params) {
    if (tr_ && params.count == n_ && params.max_coeffs == mc_ && params.sh_degree == deg_) return;
    if (tr_) {
        rng_ = tr_->dc_rng;
        gs_trainer_destroy(tr_);
    }
    tr_ = nullptr;
    gs_check(gs_trainer_create(ctx_.get(), params.count, params.sh_degree, params.max_coeffs, &tr_),
             "gs_trainer_create");
    tr_->dc_rng = rng_;
    n_ = params.count;
    mc_ = params.max_coeffs;
    deg_ = params.sh_degree;
}

TrainerHIP::StepResult TrainerHIP::step_with_densify(OwnedRawParams& owned, const Camera& camera,
                                                     const float* gt_image, const RenderConfig& cfg,
                                                     const TrainConfig& tc, const DensifyConfig& dcfg,
                                                     float scene_extent, int iteration) {
    if (cfg.eval_3D) throw std::runtime_error("TrainerHIP: eval_3D has no backward (the reference's 3D path is forward-only)");
    hipStream_t s = ctx_.get()->stream;
    RawGaussianParams params = owned.as_raw();
    ensure_trainer(params);
    gs_check(gs_trainer_set_params(tr_, params.raw_positions, params.raw_scales, params.raw_rotations,
                                   params.raw_sh_coeffs, params.raw_opacities, 1),
             "gs_trainer_set_params");
    const size_t npix = (size_t)camera.width * camera.height;
    gt_->ensure(npix * 12);
    upload(gt_->p, gt_image, npix * 12, s);
    gs_camera c = to_gs(camera);
    gs_render_cfg f = to_gs(cfg);
    gs_train_cfg t{tc.lr_position_init, tc.lr_position_final, tc.lr_feature, tc.lr_opacity, tc.lr_scaling,
                   tc.lr_rotation, tc.max_steps, tc.use_adam ? 1 : 0, tc.adam_beta1, tc.adam_beta2, tc.adam_eps};
    gs_densify_cfg d{dcfg.densify_from_iter, dcfg.densify_until_iter, dcfg.densify_interval, dcfg.grad_threshold,
                     dcfg.min_opacity, dcfg.percent_dense, dcfg.opacity_reset_interval, dcfg.max_screen_size,
                     dcfg.enabled ? 1 : 0};
    gs_step_result r;
    gs_check(gs_trainer_step_densify(tr_, &c, &f, gt_->as<float>(), &t, &d, scene_extent, iteration, &r),
             "gs_trainer_step_densify");
    // the rendering of the step's forward is not kept: the statistics pass
    // re-rendered the updated parameters into the same buffer
    owned.resize(r.n_gaussians);
    if (r.n_gaussians > 0)
        gs_check(gs_trainer_get_params(tr_, owned.positions.data(), owned.scales.data(), owned.rotations.data(),
                                       owned.sh_coeffs.data(), owned.opacities.data(), 1),
                 "gs_trainer_get_params");
    n_ = r.n_gaussians;
    return StepResult{(float)r.loss, iteration, r.n_gaussians};
}
