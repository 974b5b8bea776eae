// gs_batch.cpp -- batched multi-view training step.
//
// One step over V views of uneven resolution: the Gaussians' parameters are
// read once per step by a multi-view preprocess, all views are binned
// together (depth sort on view << 32 | depth bits, then direct placement of
// each Gaussian index into its tiles' lists, k_bin.hip; views with more than
// GS_BIN_MAX_TILES_XY tiles per side use a tile sort of view << 48 | tile << 32
// | index words instead), the forward and
// backward raster each run as one launch over the global (view, tile) list
// (optionally a persistent grid pulling tiles from an atomic counter), and the
// multi-view preprocess backward sums the views' gradients in registers and
// applies Adam in place.  The whole step can be captured once into a hipGraph
// and replayed: pair buffers are pre-sized from a measured step, every kernel
// reads the live pair count from device memory, and per-step Adam
// hyper-parameters are copied to the device before each replay.
//
// The multi-view loss is the sum of the views' mean L1 losses (each view's
// image gradient is sign(r - g) / n_view), so the gradients are the sum of the
// single-view gradients Trainer::step would compute per view
// (harmonyos_3dgs/src/trainer.cpp:7-70).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gs_device.h"
#include "gs_internal.h"
#include "gs_trainer_impl.h"

using gs::DevBuf;

gsd::CamArgs make_cam_args(const gs_camera* cam, const gs_render_cfg* cfg);

namespace {

constexpr int SCAN_BLOCK = 4096;  // k_sort.hip scan block size

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

struct gs_batch {
    gs_trainer* tr = nullptr;
    int V = 0, N = 0;
    std::vector<gs_camera> cams;
    gs_render_cfg cfg{};
    std::vector<int4> views;  // W, H, gx, first global tile
    std::vector<int> pix_off; // V + 1
    int total_tiles = 0;
    long long total_pix = 0;
    int tile_bits = 0, view_bits = 0;
    int radix_bits = 8;
    int raster_batch = 256, raster_bwd_batch = 256;
    int persistent = 1, fused_adam = 1, use_graph = 1;
    int persistent_groups = 0;
    uint32_t capacity = 0;
    bool direct = true;            // direct-placement binning (else the word sort)
    int bin_G = 0, bin_C = 0;      // Gaussians per group, groups per view
    size_t bin_L = 0;              // (view, tile, group) table entries + 1
    int bin_blocks = 0, bin_tpl = 1;
    int bin_diff_len = 0;  // largest view's (tiles per row + 1) * (tile rows + 1)
    int num_cus = 0;

    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    gs::GraphSlot graph;
    bool graph_ready = false;

    // device buffers
    DevBuf cams_dev, views_dev, pixoff_dev, hp_dev, counters;
    DevBuf means2D, depths, conics, op2d, rgb, radii, tiles;
    DevBuf dkeys, didx, dkeys_s, didx_s, rtiles, roff, scan_partial, dsort_ws;
    DevBuf words, words_tmp, tsort_ws, ranges;
    DevBuf values, bviews_dev, bvblock_dev, bin_table, bin_scanned, bin_partial, bin_rects;
    DevBuf rect_idx;  // direct binning: packed tile rectangles by (view, Gaussian), from the preprocess
    DevBuf image, d_image, gt, T_final, n_contrib, last;
    DevBuf rg;  // screen-space gradients, one block of V*N: means2D (8 B), conics (12), rgb (12), opacities (4)
    DevBuf losses, l1_partial, small;  // small: total32 @0, total64 @8, overflow @16, skipped @20
    int l1_nb = 0;

    gs_preprocessed pre() {
        return gs_preprocessed{means2D.f(), depths.f(), conics.f(), op2d.f(), rgb.f(), radii.i(), tiles.i()};
    }
    gs_raster_grads rgrads() {
        const size_t vn = (size_t)V * N;
        float* base = rg.f();
        return gs_raster_grads{base, base + 2 * vn, base + 5 * vn, base + 8 * vn};
    }
    uint32_t* total32() { return reinterpret_cast<uint32_t*>(small.p); }
    unsigned long long* total64() { return reinterpret_cast<unsigned long long*>(static_cast<char*>(small.p) + 8); }
    uint32_t* overflow() { return reinterpret_cast<uint32_t*>(static_cast<char*>(small.p) + 16); }
    // steps whose update the overflow flag suppressed since the last check
    uint32_t* skipped() { return reinterpret_cast<uint32_t*>(static_cast<char*>(small.p) + 20); }
    ~gs_batch();
};

gs_batch::~gs_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    graph.release();
    DevBuf* all[] = {&cams_dev, &views_dev, &pixoff_dev, &hp_dev, &counters, &means2D, &depths, &conics, &op2d, &rgb,
                     &radii, &tiles, &dkeys, &didx, &dkeys_s, &didx_s, &rtiles, &roff, &scan_partial, &dsort_ws,
                     &words, &words_tmp, &tsort_ws, &ranges, &values, &bviews_dev, &bvblock_dev, &bin_table, &bin_scanned,
                     &bin_partial, &bin_rects, &image, &d_image, &gt, &T_final, &n_contrib, &last,
                     &rg, &losses, &l1_partial, &small};
    for (DevBuf* b : all) b->release();
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (stream) (void)hipStreamDestroy(stream);
}

#define B_TRY(x)                      \
    do {                              \
        int st_ = (x);                \
        if (st_ != GS_OK) return st_; \
    } while (0)
#define B_HIP(expr, what)                                     \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return gs_hip_status(e_, what); \
    } while (0)

static int bits_for(long long n) {
    int b = 0;
    while ((1ll << b) < n) b++;
    return b;
}

// Pair-count-sized buffers for `cap` pairs.
static int batch_size_pairs(gs_batch* b, uint32_t cap) {
    if (cap <= b->capacity && (b->words.p || b->values.p)) return GS_OK;
    b->graph.release();
    b->graph_ready = false;
    if (b->direct) {
        B_TRY(b->values.ensure((size_t)cap * 4));
    } else {
        B_TRY(b->words.ensure((size_t)cap * 8));
        B_TRY(b->words_tmp.ensure((size_t)cap * 8));
        B_TRY(b->tsort_ws.ensure(radix_workspace_bytes(cap, b->radix_bits)));
    }
    b->capacity = cap;
    return GS_OK;
}

// Direct-placement binning layout for the current group size / bin shape:
// groups per view, the (view, tile, group) table and each view's first block.
static int batch_setup_binning(gs_batch* b, int group, int tpl) {
    b->bin_G = group;
    b->bin_tpl = tpl;
    b->bin_C = (b->N + b->bin_G - 1) / b->bin_G;
    b->bin_L = (size_t)b->bin_C * (size_t)b->total_tiles + 1;
    if (b->bin_L >= 0xFFFFFFFFull) return gs_set_error(GS_ERR_INVALID, "binning table too large");
    const int V = b->V;
    std::vector<int4> bviews(V);
    std::vector<int> vblock0(V + 1, 0);
    size_t tab = 0;
    b->bin_diff_len = 0;
    for (int v = 0; v < V; v++) {
        const int gx = b->views[v].z, gy = (b->views[v].y + 15) / 16;
        bviews[v] = int4{gx, gy, (int)(uint32_t)tab, b->views[v].w};
        tab += (size_t)gx * gy * b->bin_C;
        vblock0[v + 1] = vblock0[v] + bin_blocks(b->bin_C, gx, gy, b->bin_tpl);
        b->bin_diff_len = std::max(b->bin_diff_len, (gx + 1) * (gy + 1));
    }
    b->bin_blocks = vblock0[V];
    const size_t VN = (size_t)V * b->N;
    B_TRY(b->bviews_dev.ensure(sizeof(int4) * V));
    B_TRY(b->bvblock_dev.ensure(sizeof(int) * (V + 1)));
    B_TRY(b->bin_table.ensure(b->bin_L * 4));
    B_TRY(b->bin_scanned.ensure(b->bin_L * 4));
    B_TRY(b->bin_rects.ensure(VN * 8));
    B_TRY(b->bin_partial.ensure(((b->bin_L + SCAN_BLOCK - 1) / SCAN_BLOCK + 1) * 8));
    B_HIP(hipMemcpyAsync(b->bviews_dev.p, bviews.data(), sizeof(int4) * V, hipMemcpyHostToDevice, b->stream),
          "hipMemcpyAsync(views)");
    B_HIP(hipMemcpyAsync(b->bvblock_dev.p, vblock0.data(), sizeof(int) * (V + 1), hipMemcpyHostToDevice, b->stream),
          "hipMemcpyAsync(blocks)");
    B_HIP(hipStreamSynchronize(b->stream), "hipStreamSynchronize");  // the host vectors die here
    b->graph.release();
    b->graph_ready = false;
    return GS_OK;
}

extern "C" int gs_batch_create(gs_trainer* tr, int n_views, const gs_camera* cams, const gs_render_cfg* cfg,
                               gs_batch** out) {
    if (!tr || !cams || !cfg || !out || n_views <= 0)
        return gs_set_error(GS_ERR_INVALID, "gs_batch_create: bad argument");
    if (n_views > 65535) return gs_set_error(GS_ERR_INVALID, "gs_batch_create: at most 65535 views");
    for (int v = 0; v < n_views; v++) B_TRY(check_cam_cfg(&cams[v], cfg));
    gs_batch* b = new gs_batch();
    b->tr = tr;
    b->V = n_views;
    b->N = tr->n;
    b->cams.assign(cams, cams + n_views);
    b->cfg = *cfg;
    b->radix_bits = tr->ctx->radix_bits;
    b->raster_batch = tr->ctx->raster_batch;
    b->raster_bwd_batch = tr->ctx->raster_bwd_batch;
    int max_tiles = 0, max_side = 0;
    b->pix_off.push_back(0);
    for (int v = 0; v < n_views; v++) {
        int W = cams[v].width, H = cams[v].height;
        int gx = (W + 15) / 16, gy = (H + 15) / 16;
        if (gx * gy > 65536) {
            delete b;
            return gs_set_error(GS_ERR_INVALID, "gs_batch_create: view %d has more than 65536 tiles", v);
        }
        b->views.push_back(int4{W, H, gx, b->total_tiles});
        b->total_tiles += gx * gy;
        max_tiles = std::max(max_tiles, gx * gy);
        max_side = std::max(max_side, std::max(gx, gy));
        b->pix_off.push_back(b->pix_off.back() + W * H);
    }
    b->total_pix = b->pix_off.back();
    b->tile_bits = std::max(1, bits_for(max_tiles));
    b->view_bits = bits_for(n_views);
    b->direct = max_side <= GS_BIN_MAX_TILES_XY;
    int st = GS_OK;
    auto fail = [&](int s) {
        delete b;
        return s;
    };
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, tr->ctx->device) != hipSuccess) return fail(gs_set_error(GS_ERR_HIP, "props"));
    b->num_cus = prop.multiProcessorCount;
    b->persistent_groups = b->num_cus * std::max(1, tr->ctx->persistent_per_cu);
    if (tr->ctx->persistent_per_cu == 0) b->persistent = 0;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess)
        return fail(gs_set_error(GS_ERR_HIP, "gs_batch_create: stream/event creation failed"));

    const size_t VN = (size_t)n_views * b->N;
    const size_t P = (size_t)b->total_pix;
    std::vector<gsd::CamArgs> ca(n_views);
    for (int v = 0; v < n_views; v++) ca[v] = make_cam_args(&cams[v], cfg);
    if ((st = b->cams_dev.ensure(sizeof(gsd::CamArgs) * n_views)) || (st = b->views_dev.ensure(sizeof(int4) * n_views)) ||
        (st = b->pixoff_dev.ensure(sizeof(int) * (n_views + 1))) || (st = b->hp_dev.ensure(sizeof(gso::AdamDev))) ||
        (st = b->counters.ensure(256)) || (st = b->means2D.ensure(VN * 8)) || (st = b->depths.ensure(VN * 4)) ||
        (st = b->conics.ensure(VN * 12)) || (st = b->op2d.ensure(VN * 4)) || (st = b->rgb.ensure(VN * 12)) ||
        (st = b->radii.ensure(VN * 4)) || (st = b->tiles.ensure(VN * 4)) || (st = b->dkeys.ensure(VN * 8)) ||
        (st = b->didx.ensure(VN * 4)) || (st = b->dkeys_s.ensure(VN * 8)) || (st = b->didx_s.ensure(VN * 4)) ||
        (b->direct && (st = b->rect_idx.ensure(VN * 4))) ||
        (st = b->scan_partial.ensure(((VN + SCAN_BLOCK - 1) / SCAN_BLOCK + 1) * 8)) ||
        (st = b->dsort_ws.ensure(radix_workspace_bytes((uint32_t)VN, b->radix_bits))) ||
        (st = b->ranges.ensure((size_t)b->total_tiles * 8)) || (st = b->image.ensure(P * 12)) ||
        (st = b->d_image.ensure(P * 12)) || (st = b->gt.ensure(P * 12)) || (st = b->T_final.ensure(P * 4)) ||
        (st = b->n_contrib.ensure(P * 4)) || (st = b->last.ensure(P * 4)) || (st = b->rg.ensure(VN * 36)) ||
        (st = b->losses.ensure(8 * n_views)) || (st = b->small.ensure(256)))
        return fail(st);
    int max_pix = 0;
    for (int v = 0; v < n_views; v++) max_pix = std::max(max_pix, b->pix_off[v + 1] - b->pix_off[v]);
    b->l1_nb = std::max(1, std::min(512, (max_pix * 3 + 255) / 256 / 4));
    if ((st = b->l1_partial.ensure((size_t)n_views * b->l1_nb * 8))) return fail(st);
    if (b->direct) {
        if ((st = batch_setup_binning(b, tr->ctx->bin_chunk, tr->ctx->bin_tpl))) return fail(st);
    } else if ((st = b->rtiles.ensure(VN * 4)) || (st = b->roff.ensure(VN * 4))) {
        return fail(st);
    }
    hipStream_t s = b->stream;
    if (hipMemcpyAsync(b->cams_dev.p, ca.data(), sizeof(gsd::CamArgs) * n_views, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b->views_dev.p, b->views.data(), sizeof(int4) * n_views, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b->pixoff_dev.p, b->pix_off.data(), sizeof(int) * (n_views + 1), hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        launch_zero(b->small.p, 256, s) != hipSuccess || launch_zero(b->gt.p, P * 12, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(gs_set_error(GS_ERR_HIP, "gs_batch_create: upload failed"));
    *out = b;
    return GS_OK;
}

extern "C" int gs_batch_destroy(gs_batch* b) {
    delete b;
    return GS_OK;
}

extern "C" int gs_batch_set_options(gs_batch* b, int persistent, int fused_adam, int use_graph) {
    if (!b) return gs_set_error(GS_ERR_INVALID, "null batch");
    if (b->persistent != persistent || b->fused_adam != fused_adam) {
        b->graph.release();
        b->graph_ready = false;
    }
    b->persistent = persistent;
    b->fused_adam = fused_adam;
    b->use_graph = use_graph;
    return GS_OK;
}

extern "C" int gs_batch_set_targets(gs_batch* b, const float* const* gts, int from_host) {
    if (!b || !gts) return gs_set_error(GS_ERR_INVALID, "null argument");
    for (int v = 0; v < b->V; v++) {
        size_t off = (size_t)b->pix_off[v] * 3, n = (size_t)(b->pix_off[v + 1] - b->pix_off[v]) * 3;
        B_HIP(hipMemcpyAsync(b->gt.f() + off, gts[v], n * 4, from_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                             b->stream),
              "hipMemcpyAsync(targets)");
    }
    B_HIP(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    return GS_OK;
}

// Everything from preprocess to the tile ranges (the forward's binning).
static int enqueue_binning(gs_batch* b, hipStream_t s) {
    gs_trainer* t = b->tr;
    const uint32_t VN = (uint32_t)((size_t)b->V * b->N);
    gs_gaussians g = t->gaussians();
    gs_preprocessed pre = b->pre();
    const int deg = std::min(b->cfg.sh_degree, t->sh_degree);
    B_HIP(launch_preprocess_fwd_multi(g, static_cast<const gsd::CamArgs*>(b->cams_dev.p), b->V, pre,
                                      static_cast<uint64_t*>(b->dkeys.p), b->didx.u(), deg, s,
                                      b->direct ? b->rect_idx.u() : nullptr),
          "multi-view preprocess");
    B_HIP(launch_radix_sort_bits(static_cast<const uint64_t*>(b->dkeys.p), b->didx.u(),
                                 static_cast<uint64_t*>(b->dkeys_s.p), b->didx_s.u(), VN, 0, 32 + b->view_bits,
                                 b->radix_bits, b->dsort_ws.p, s, nullptr),
          "depth sort");
    if (b->direct) {
        gsb::BinArgs a{};
        a.V = b->V;
        a.N = b->N;
        a.C = b->bin_C;
        a.G = b->bin_G;
        a.S = 0;
        a.G1 = b->bin_G;
        a.sidx = b->didx_s.u();
        a.sidx_stride = 1;
        a.means2D = reinterpret_cast<const float2*>(pre.means2D);
        a.radii = pre.radii;
        a.views = static_cast<const int4*>(b->bviews_dev.p);
        a.vblock0 = b->bvblock_dev.i();
        a.table = b->bin_table.u();
        a.table_len = b->bin_L;
        a.rects = static_cast<uint2*>(b->bin_rects.p);
        a.out = b->values.u();
        a.capacity = b->capacity;
        a.overflow = b->overflow();
        a.diff_len = b->bin_diff_len;
        a.rect_idx = b->rect_idx.u();
        B_HIP(launch_bin_count(a, b->bin_blocks, b->bin_tpl, s), "binning count");
        B_HIP(launch_scan_u32_exclusive(b->bin_table.u(), (uint32_t)b->bin_L, b->bin_scanned.u(),
                                        static_cast<unsigned long long*>(b->bin_partial.p), s, b->total32(),
                                        b->total64()),
              "binning scan");
        a.table = b->bin_scanned.u();
        B_HIP(launch_bin_place(a, b->bin_blocks, b->bin_tpl, s), "binning place");
        B_HIP(launch_bin_ranges(a, b->bin_scanned.u(), (uint32_t)b->total_tiles, b->ranges.u(), s), "binning ranges");
        return GS_OK;
    }
    B_HIP(launch_gather_tiles_multi(VN, b->N, static_cast<const uint64_t*>(b->dkeys_s.p), b->didx_s.u(), pre.tiles_touched,
                                    b->rtiles.i(), s),
          "gather tiles");
    B_HIP(launch_scan_tiles(b->rtiles.i(), VN, b->roff.u(), static_cast<unsigned long long*>(b->scan_partial.p),
                            b->total32(), b->total64(), s),
          "scan");
    B_HIP(launch_duplicate_multi(VN, b->N, static_cast<const uint64_t*>(b->dkeys_s.p), b->didx_s.u(), pre.means2D,
                                 pre.radii, b->roff.u(), static_cast<const int4*>(b->views_dev.p),
                                 static_cast<uint64_t*>(b->words_tmp.p), b->capacity, b->overflow(), s),
          "duplicate");
    B_HIP(launch_radix_sort_bits(static_cast<const uint64_t*>(b->words_tmp.p), nullptr,
                                 static_cast<uint64_t*>(b->words.p), nullptr, b->capacity, 32, 32 + b->tile_bits,
                                 b->radix_bits, b->tsort_ws.p, s, b->total32()),
          "tile sort");
    B_HIP(launch_zero(b->ranges.p, (size_t)b->total_tiles * 8, s), "zero ranges");
    B_HIP(launch_tile_ranges_multi(static_cast<const uint64_t*>(b->words.p), b->capacity, b->total32(),
                                   static_cast<const int4*>(b->views_dev.p), b->ranges.u(), s),
          "tile ranges");
    return GS_OK;
}

static gsr::MultiRaster multi_raster(gs_batch* b, unsigned* counter) {
    gsr::MultiRaster m;
    m.V = b->V;
    m.N = b->N;
    m.total_tiles = b->total_tiles;
    m.views = static_cast<const int4*>(b->views_dev.p);
    m.pix_off = b->pixoff_dev.i();
    m.means2D = reinterpret_cast<const float2*>(b->means2D.p);
    m.conics = b->conics.f();
    m.opac = b->op2d.f();
    m.rgb = b->rgb.f();
    m.values = b->direct ? b->values.u() : b->words.u();
    m.vstride = b->direct ? 1 : 2;
    m.ranges = b->ranges.u();
    m.bg0 = b->cfg.bg_color[0];
    m.bg1 = b->cfg.bg_color[1];
    m.bg2 = b->cfg.bg_color[2];
    m.counter = counter;
    return m;
}

static int enqueue_raster_fwd(gs_batch* b, hipStream_t s) {
    unsigned* ctr = b->persistent ? static_cast<unsigned*>(b->counters.p) : nullptr;
    if (ctr) B_HIP(launch_zero(ctr, 32, s), "zero queue counters");
    gsr::MultiRaster m = multi_raster(b, ctr);
    gsr::MultiPix px{b->image.f(), b->T_final.f(), b->n_contrib.i(), b->last.u(), b->d_image.f()};
    B_HIP(launch_raster_fwd_multi(m, px, b->raster_batch, b->persistent_groups, s), "multi-view raster forward");
    return GS_OK;
}

static int enqueue_forward(gs_batch* b, hipStream_t s) {
    B_TRY(enqueue_binning(b, s));
    return enqueue_raster_fwd(b, s);
}

static int enqueue_l1(gs_batch* b, hipStream_t s) {
    B_HIP(launch_l1_multi(b->image.f(), b->gt.f(), b->d_image.f(), b->pixoff_dev.i(), b->V, b->l1_nb,
                          static_cast<double*>(b->l1_partial.p), static_cast<double*>(b->losses.p), s),
          "multi-view l1");
    return GS_OK;
}

// Zeroes the screen-space gradients and runs the multi-view backward raster.
static int enqueue_raster_bwd(gs_batch* b, hipStream_t s) {
    const size_t VN = (size_t)b->V * b->N;
    B_HIP(launch_zero(b->rg.p, VN * 36, s), "zero");
    unsigned* ctr = b->persistent ? static_cast<unsigned*>(b->counters.p) + 16 : nullptr;
    if (ctr) B_HIP(launch_zero(ctr, 32, s), "zero queue counters");
    gsr::MultiRaster m = multi_raster(b, ctr);
    gsr::MultiPix px{b->image.f(), b->T_final.f(), b->n_contrib.i(), b->last.u(), b->d_image.f()};
    gs_raster_grads rg = b->rgrads();
    gsr::MultiGrads mg{reinterpret_cast<float2*>(rg.d_means2D), rg.d_conics, rg.d_rgb, rg.d_opacities_2d};
    B_HIP(launch_raster_bwd_multi(m, px, mg, b->raster_bwd_batch, b->persistent_groups, s),
          "multi-view raster backward");
    return GS_OK;
}

static int enqueue_step(gs_batch* b, hipStream_t s) {
    gs_trainer* t = b->tr;
    B_TRY(enqueue_forward(b, s));
    B_TRY(enqueue_l1(b, s));
    B_TRY(enqueue_raster_bwd(b, s));
    gs_raster_grads rg = b->rgrads();
    gs_gaussians g = t->gaussians();
    gs_param_grads pg{t->grad[0].f(), t->grad[1].f(), t->grad[2].f(), t->grad[3].f(), t->grad[4].f()};
    float* P[5];
    float* M[5];
    float* Vv[5];
    for (int k = 0; k < 5; k++) {
        P[k] = t->param[k].f();
        M[k] = t->adam_m[k].f();
        Vv[k] = t->adam_v[k].f();
    }
    const gso::AdamDev* hp = static_cast<const gso::AdamDev*>(b->hp_dev.p);
    // A step whose pair lists overflowed the buffers rendered truncated lists:
    // its update is suppressed on the device (parameters and Adam moments are
    // left as they were) and counted, so the caller can redo it after the
    // regrow that batch_result performs.
    if (b->fused_adam) {
        B_HIP(launch_preprocess_bwd_multi(g, static_cast<const gsd::CamArgs*>(b->cams_dev.p), nullptr, b->V, b->radii.i(),
                                          rg, pg, P, M, Vv, hp, nullptr, t->sh_degree, s, b->overflow()),
              "multi-view preprocess backward + adam");
    } else {
        B_HIP(launch_preprocess_bwd_multi(g, static_cast<const gsd::CamArgs*>(b->cams_dev.p), nullptr, b->V, b->radii.i(),
                                          rg, pg, nullptr, nullptr, nullptr, nullptr, nullptr, t->sh_degree, s),
              "multi-view preprocess backward");
        gso::AdamArgs a{};
        const uint32_t sizes[5] = {(uint32_t)t->n * 3u, (uint32_t)t->n * 3u, (uint32_t)t->n * 4u,
                                   (uint32_t)t->n * (uint32_t)t->max_coeffs * 3u, (uint32_t)t->n};
        for (int k = 0; k < 5; k++) a.seg[k] = gso::AdamSeg{P[k], pg.d_positions, M[k], Vv[k], sizes[k], 0.0f};
        a.seg[0].g = pg.d_positions;
        a.seg[1].g = pg.d_scales;
        a.seg[2].g = pg.d_rotations;
        a.seg[3].g = pg.d_sh_coeffs;
        a.seg[4].g = pg.d_opacities;
        a.hp = hp;
        a.skip = b->overflow();
        B_HIP(launch_adam(a, s), "adam");
    }
    B_HIP(launch_count_if(b->overflow(), b->skipped(), s), "count suppressed updates");
    return GS_OK;
}

// Sizes the pair buffers with one binning pass (eager, synchronous).
static int batch_measure(gs_batch* b, long long* pairs) {
    if (b->capacity == 0) B_TRY(batch_size_pairs(b, 1024));
    B_TRY(enqueue_binning(b, b->stream));
    unsigned long long K = 0;
    B_HIP(hipMemcpyAsync(&K, b->total64(), 8, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync");
    B_HIP(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    if (K > 0xFFFFFFFFull) return gs_set_error(GS_ERR_CAPACITY, "more than 2^32-1 pairs");
    *pairs = (long long)K;
    if (K > b->capacity) {
        unsigned long long cap = K + K / 4 + 65536;
        if (cap > 0xFFFFFFFFull) cap = 0xFFFFFFFFull;
        B_TRY(batch_size_pairs(b, (uint32_t)cap));
        B_HIP(launch_zero(b->overflow(), 4, b->stream), "zero overflow");
    }
    return GS_OK;
}

static int batch_upload_hparams(gs_batch* b, const gs_train_cfg* tc, int iteration) {
    gs_trainer* t = b->tr;
    gso::AdamDev h{};
    h.lr[0] = gs_lr_schedule(tc->lr_position_init, tc->lr_position_final, iteration, tc->max_steps);
    h.lr[1] = tc->lr_scaling;
    h.lr[2] = tc->lr_rotation;
    h.lr[3] = tc->lr_feature;
    h.lr[4] = tc->lr_opacity;
    h.b1 = tc->adam_beta1;
    h.b2 = tc->adam_beta2;
    h.eps = tc->adam_eps;
    t->adam_steps++;
    gs_adam_bias_corrections(h.b1, h.b2, t->adam_steps, &h.bc1, &h.bc2);
    // pageable source: the runtime stages it before returning, so `h` may die
    B_HIP(hipMemcpyAsync(b->hp_dev.p, &h, sizeof(h), hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync(hparams)");
    return GS_OK;
}

// Reads the losses / pair count / overflow flag of the last step (synchronous).
static int batch_result(gs_batch* b, gs_step_result* out, int iteration) {
    std::vector<double> L(b->V);
    unsigned long long K = 0;
    uint32_t flags[2] = {0, 0};  // overflow, suppressed updates
    B_HIP(hipMemcpyAsync(L.data(), b->losses.p, 8 * b->V, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync");
    B_HIP(hipMemcpyAsync(&K, b->total64(), 8, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync");
    B_HIP(hipMemcpyAsync(flags, b->overflow(), 8, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync");
    B_HIP(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    const uint32_t ovf = flags[0], skipped = flags[1];
    double s = 0.0;
    for (double x : L) s += x;
    if (out) {
        out->loss = s / b->V;
        out->iteration = iteration;
        out->n_gaussians = b->N;
        out->pairs = (long long)K;
    }
    if (ovf) {
        // some step since the last check dropped pairs: that step and every
        // later one left the parameters and Adam state untouched; their Adam
        // step numbers are handed back, and the buffers grow for the redo
        B_HIP(launch_zero(b->overflow(), 8, b->stream), "zero overflow");
        b->tr->adam_steps -= (int)skipped;
        const uint32_t old_cap = b->capacity;
        unsigned long long cap = K + K / 4 + 65536;
        if (cap > 0xFFFFFFFFull) cap = 0xFFFFFFFFull;
        B_TRY(batch_size_pairs(b, (uint32_t)cap));
        return gs_set_error(GS_ERR_CAPACITY,
                            "a batched step had %llu pairs for a capacity of %u: %u step(s) were not applied; "
                            "buffers regrown to %u, repeat them",
                            K, old_cap, skipped, b->capacity);
    }
    return GS_OK;
}

extern "C" int gs_batch_step(gs_batch* b, const gs_train_cfg* tc, int iteration, gs_step_result* out) {
    if (!b || !tc) return gs_set_error(GS_ERR_INVALID, "gs_batch_step: null argument");
    if (!tc->use_adam) return gs_set_error(GS_ERR_INVALID, "gs_batch_step: the batched step uses Adam");
    gs_ctx* c = b->tr->ctx;
    B_HIP(hipEventRecord(b->ev_in, c->stream), "hipEventRecord");
    B_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0), "hipStreamWaitEvent");
    if (b->capacity == 0) {
        long long K;
        B_TRY(batch_measure(b, &K));
    }
    B_TRY(batch_upload_hparams(b, tc, iteration));
    if (b->use_graph) {
        if (!b->graph_ready) {
            B_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            int st = enqueue_step(b, b->stream);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &graph);
            if (st != GS_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return st;
            }
            if (e != hipSuccess) return gs_hip_status(e, "hipStreamEndCapture");
            b->graph.graph = graph;
            B_HIP(hipGraphInstantiate(&b->graph.exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            b->graph_ready = true;
        }
        B_HIP(hipGraphLaunch(b->graph.exec, b->stream), "hipGraphLaunch");
    } else {
        B_TRY(enqueue_step(b, b->stream));
    }
    B_HIP(hipEventRecord(b->ev_out, b->stream), "hipEventRecord");
    B_HIP(hipStreamWaitEvent(c->stream, b->ev_out, 0), "hipStreamWaitEvent");
    if (out) return batch_result(b, out, iteration);
    return GS_OK;
}

// Forward only for all views (the batched render).
extern "C" int gs_batch_render(gs_batch* b, long long* pairs) {
    if (!b) return gs_set_error(GS_ERR_INVALID, "null batch");
    gs_ctx* c = b->tr->ctx;
    B_HIP(hipEventRecord(b->ev_in, c->stream), "hipEventRecord");
    B_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0), "hipStreamWaitEvent");
    if (b->capacity == 0) {
        long long K;
        B_TRY(batch_measure(b, &K));
    }
    B_TRY(enqueue_forward(b, b->stream));
    B_HIP(hipEventRecord(b->ev_out, b->stream), "hipEventRecord");
    B_HIP(hipStreamWaitEvent(c->stream, b->ev_out, 0), "hipStreamWaitEvent");
    if (pairs) {
        unsigned long long K = 0;
        B_HIP(hipMemcpyAsync(&K, b->total64(), 8, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync");
        B_HIP(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        *pairs = (long long)K;
    }
    return GS_OK;
}

extern "C" int gs_batch_result(gs_batch* b, gs_step_result* out, double* view_losses) {
    if (!b) return gs_set_error(GS_ERR_INVALID, "null batch");
    if (view_losses) {
        B_HIP(hipMemcpyAsync(view_losses, b->losses.p, 8 * b->V, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync");
    }
    return batch_result(b, out, 0);
}

extern "C" void* gs_batch_buffer(gs_batch* b, const char* which) {
    if (!b || !which) return nullptr;
    std::string w(which);
    if (w == "image") return b->image.p;
    if (w == "d_image") return b->d_image.p;
    if (w == "target") return b->gt.p;
    if (w == "T_final") return b->T_final.p;
    if (w == "n_contrib") return b->n_contrib.p;
    if (w == "last_contrib") return b->last.p;
    if (w == "means2D") return b->means2D.p;
    if (w == "radii") return b->radii.p;
    if (w == "ranges") return b->ranges.p;
    if (w == "pairs") return b->direct ? b->values.p : b->words.p;  // uint32 indices or 64-bit words
    if (w == "losses") return b->losses.p;
    const gs_raster_grads g = b->rgrads();
    if (w == "rg_means2D") return g.d_means2D;
    if (w == "rg_conics") return g.d_conics;
    if (w == "rg_rgb") return g.d_rgb;
    if (w == "rg_op") return g.d_opacities_2d;
    return nullptr;
}

extern "C" int gs_batch_info(gs_batch* b, int* total_tiles, long long* total_pixels, long long* capacity) {
    if (!b) return gs_set_error(GS_ERR_INVALID, "null batch");
    if (total_tiles) *total_tiles = b->total_tiles;
    if (total_pixels) *total_pixels = b->total_pix;
    if (capacity) *capacity = b->capacity;
    return GS_OK;
}

// ------------------------------------------------------------ autotuning ----
// Coordinate descent over the batch's knobs, each timed on its own stage with
// HIP events on the batch's stream (median of `reps` runs after one untimed
// run): binning bin shape and group size (binning pass), persistent groups per
// CU (forward + backward raster; 0 = one group per tile) and the LDS batch
// depths of the two rasterizers.  Parameters and optimizer state are not
// touched: only binning, forward, loss and backward raster run.
namespace {

struct BatchTimer {
    hipEvent_t a = nullptr, e = nullptr;
    ~BatchTimer() {
        if (a) (void)hipEventDestroy(a);
        if (e) (void)hipEventDestroy(e);
    }
};

template <typename F>
int batch_time_once(gs_batch* b, BatchTimer& tm, F&& f, float* ms) {
    B_HIP(hipEventRecord(tm.a, b->stream), "hipEventRecord");
    B_TRY(f());
    B_HIP(hipEventRecord(tm.e, b->stream), "hipEventRecord");
    B_HIP(hipEventSynchronize(tm.e), "hipEventSynchronize");
    B_HIP(hipEventElapsedTime(ms, tm.a, tm.e), "hipEventElapsedTime");
    return GS_OK;
}

// Candidates timed round-robin, `reps` runs each after one untimed run; the
// smallest median wins and is left set.
template <typename Set, typename Run>
int batch_pick(gs_batch* b, BatchTimer& tm, int reps, const std::vector<int>& cands, Set&& set, Run&& run, int* best,
               float* best_ms) {
    std::vector<std::vector<float>> t(cands.size());
    for (size_t i = 0; i < cands.size(); i++) {
        B_TRY(set(cands[i]));
        B_TRY(run());
    }
    for (int r = 0; r < reps; r++) {
        for (size_t i = 0; i < cands.size(); i++) {
            B_TRY(set(cands[i]));
            float ms = 0.0f;
            B_TRY(batch_time_once(b, tm, run, &ms));
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
    B_TRY(set(cands[bi]));
    *best = cands[bi];
    *best_ms = bm;
    return GS_OK;
}

}  // namespace

extern "C" int gs_batch_autotune(gs_batch* b, int reps, gs_tuning* out) {
    if (!b) return gs_set_error(GS_ERR_INVALID, "null batch");
    if (reps < 1) reps = 1;
    gs_ctx* c = b->tr->ctx;
    B_HIP(hipEventRecord(b->ev_in, c->stream), "hipEventRecord");
    B_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0), "hipStreamWaitEvent");
    if (b->capacity == 0) {
        long long K;
        B_TRY(batch_measure(b, &K));
    }
    BatchTimer tm;
    B_HIP(hipEventCreate(&tm.a), "hipEventCreate");
    B_HIP(hipEventCreate(&tm.e), "hipEventCreate");
    gs_tuning res{};
    res.radix_bits = b->radix_bits;
    res.split_groups = c->split_groups;  // a single-view knob, reported unchanged
    auto bin = [&]() { return enqueue_binning(b, b->stream); };
    if (b->direct) {
        B_TRY(batch_pick(b, tm, reps, {1, 2, 4}, [&](int x) { return batch_setup_binning(b, b->bin_G, x); }, bin,
                         &res.bin_tiles_per_lane, &res.binning_ms));
        B_TRY(batch_pick(b, tm, reps, {2048, 4096, 8192, 16384, 32768, 65536},
                         [&](int x) { return batch_setup_binning(b, x, b->bin_tpl); }, bin, &res.bin_group,
                         &res.binning_ms));
    } else {
        int unused = 0;
        B_TRY(batch_pick(b, tm, reps, {0}, [](int) { return GS_OK; }, bin, &unused, &res.binning_ms));
    }
    B_TRY(bin());
    // persistent groups per CU, on forward + loss + backward raster
    const int persistent0 = b->persistent;
    auto set_ppc = [&](int x) {
        b->persistent = x > 0 ? 1 : 0;
        b->persistent_groups = b->num_cus * (x > 0 ? x : 1);
        return GS_OK;
    };
    auto fwd_bwd = [&]() {
        B_TRY(enqueue_raster_fwd(b, b->stream));
        B_TRY(enqueue_l1(b, b->stream));
        return enqueue_raster_bwd(b, b->stream);
    };
    float ms_fb = 0.0f;
    std::vector<int> ppc = {2, 4, 6, 8};
    if (!persistent0) ppc = {0};
    else ppc.insert(ppc.begin(), 0);
    B_TRY(batch_pick(b, tm, reps, ppc, set_ppc, fwd_bwd, &res.persistent_per_cu, &ms_fb));
    auto fwd = [&]() { return enqueue_raster_fwd(b, b->stream); };
    B_TRY(batch_pick(b, tm, reps, {128, 256, 512}, [&](int x) { b->raster_batch = x; return GS_OK; }, fwd,
                     &res.raster_batch, &res.raster_fwd_ms));
    B_TRY(fwd());
    B_TRY(enqueue_l1(b, b->stream));
    auto bwd = [&]() { return enqueue_raster_bwd(b, b->stream); };
    B_TRY(batch_pick(b, tm, reps, {128, 256}, [&](int x) { b->raster_bwd_batch = x; return GS_OK; }, bwd,
                     &res.raster_bwd_batch, &res.raster_bwd_ms));
    B_HIP(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    b->graph.release();
    b->graph_ready = false;
    if (out) *out = res;
    return GS_OK;
}
