// capi.hip -- the C ABI (include/gs_hip.h): argument checks, the context's
// workspace, and kernel launches on the context stream.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "gs_device.h"
#include "gs_internal.h"

static thread_local std::string g_err;

int gs_set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int gs_hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return GS_OK;
    return gs_set_error(GS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

extern "C" const char* gs_last_error(void) { return g_err.c_str(); }

#define GS_CHECK_ARG(cond, ...)                                    \
    do {                                                           \
        if (!(cond)) return gs_set_error(GS_ERR_INVALID, __VA_ARGS__); \
    } while (0)
#define GS_HIP(expr, what)                            \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return gs_hip_status(e_, what); \
    } while (0)

static bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

// ------------------------------------------------ allocation-free scan ----
// The fixed device scratch of gs_scan_tiles_noalloc: the 64-bit pair total at
// offset 0 and the per-block partial sums (one per 4096 Gaussians, plus one)
// from offset 256, sized for GS_NOALLOC_SCAN_MAX Gaussians (33 KB).  Each
// device holds its own copy until the process ends.  In host code the name
// `scratch` is only a stand-in: its device address is taken once per device
// with hipGetSymbolAddress (gs_ctx_create, with that device current) and only
// that address is used.
constexpr size_t kNoallocPartials = (GS_NOALLOC_SCAN_MAX + 4095u) / 4096u + 1u;
constexpr size_t kNoallocBytes = 256 + ((kNoallocPartials * 8 + 255) & ~size_t(255));
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static std::atomic<void*> g_scratch_ptr[kMaxDevices];

// Resolves this device's copy of `scratch` (the device must be current).
static int scratch_init(int device) {
    GS_CHECK_ARG(device >= 0 && device < kMaxDevices, "device %d beyond the %d supported", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device].load()) return GS_OK;
    void* p = nullptr;
    GS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    GS_CHECK_ARG(p != nullptr, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[device].store(p);
    return GS_OK;
}

// ------------------------------------------------------------- context ----
extern "C" int gs_ctx_create(int device, void* stream, gs_ctx** out) {
    GS_CHECK_ARG(out != nullptr, "gs_ctx_create: out is null");
    int n = 0;
    GS_HIP(hipGetDeviceCount(&n), "hipGetDeviceCount");
    GS_CHECK_ARG(device >= 0 && device < n, "gs_ctx_create: device %d out of range (%d devices)", device, n);
    GS_HIP(hipSetDevice(device), "hipSetDevice");
    int st = scratch_init(device);
    if (st) return st;
    gs_ctx* c = new gs_ctx();
    c->device = device;
    c->stream = static_cast<hipStream_t>(stream);
    *out = c;
    return GS_OK;
}

// The split a gs_bin_tiles_impl call left pending: the placement arguments
// for the remaining depth groups and what the forward raster needs.
struct BinSplit {
    bool pending = false;
    gsb::BinArgs a{};
    int tpl = 1, bins = 0, C = 0, split = 0, gx = 0, gy = 0;
};

extern "C" int gs_ctx_destroy(gs_ctx* c) {
    if (!c) return GS_OK;
    (void)hipSetDevice(c->device);
    if (c->ws) (void)hipFree(c->ws);
    if (c->ws2) (void)hipFree(c->ws2);
    if (c->small) (void)hipFree(c->small);
    delete static_cast<BinSplit*>(c->bin_split);
    delete c;
    return GS_OK;
}

extern "C" int gs_ctx_set_stream(gs_ctx* c, void* stream) {
    GS_CHECK_ARG(c != nullptr, "null context");
    c->stream = static_cast<hipStream_t>(stream);
    return GS_OK;
}

extern "C" int gs_ctx_synchronize(gs_ctx* c) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return GS_OK;
}

extern "C" int gs_ctx_set_tuning(gs_ctx* c, int radix_bits, int raster_batch, int raster_bwd_batch) {
    GS_CHECK_ARG(c != nullptr, "null context");
    if (radix_bits) {
        GS_CHECK_ARG(radix_bits >= 4 && radix_bits <= 8, "radix_bits must be in [4, 8]");
        c->radix_bits = radix_bits;
    }
    if (raster_batch) {
        GS_CHECK_ARG(raster_batch == 64 || raster_batch == 128 || raster_batch == 256 || raster_batch == 512,
                     "raster_batch must be 64, 128, 256 or 512");
        c->raster_batch = raster_batch;
    }
    if (raster_bwd_batch) {
        GS_CHECK_ARG(raster_bwd_batch == 64 || raster_bwd_batch == 128 || raster_bwd_batch == 256,
                     "raster_bwd_batch must be 64, 128 or 256");
        c->raster_bwd_batch = raster_bwd_batch;
    }
    return GS_OK;
}

// Grows a workspace to at least `bytes`.  Not for use inside graph capture:
// callers that capture reserve the workspaces beforehand (gs_ctx_reserve).
static int grow(gs_ctx* c, void** buf, size_t* have, size_t bytes, void** out) {
    if (bytes > *have) {
        if (c->capturing)
            return gs_set_error(GS_ERR_CAPACITY, "workspace of %zu bytes needed during capture (%zu reserved)", bytes,
                                *have);
        GS_HIP(hipSetDevice(c->device), "hipSetDevice");
        if (*buf) {
            GS_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            GS_HIP(hipFree(*buf), "hipFree(workspace)");
            *buf = nullptr;
            *have = 0;
        }
        size_t nb = bytes + bytes / 4 + (1 << 20);
        GS_HIP(hipMalloc(buf, nb), "hipMalloc(workspace)");
        *have = nb;
    }
    *out = *buf;
    return GS_OK;
}

int gs_ws_get(gs_ctx* c, size_t bytes, void** out) { return grow(c, &c->ws, &c->ws_bytes, bytes, out); }
int gs_ws2_get(gs_ctx* c, size_t bytes, void** out) { return grow(c, &c->ws2, &c->ws2_bytes, bytes, out); }

// Small fixed scratch: scan partials, overflow flags, loss partials.
int gs_small_get(gs_ctx* c, void** out) {
    if (!c->small) {
        GS_HIP(hipSetDevice(c->device), "hipSetDevice");
        GS_HIP(hipMalloc(&c->small, GS_SMALL_BYTES), "hipMalloc(small scratch)");
        GS_HIP(hipMemsetAsync(c->small, 0, GS_SMALL_BYTES, c->stream), "hipMemsetAsync(small scratch)");
    }
    *out = c->small;
    return GS_OK;
}

extern "C" int gs_ctx_reserve(gs_ctx* c, size_t bytes) {
    GS_CHECK_ARG(c != nullptr, "null context");
    void* p;
    int st = gs_ws_get(c, bytes, &p);
    if (st) return st;
    return gs_small_get(c, &p);
}

extern "C" int gs_zero(gs_ctx* c, void* ptr, size_t bytes) {
    GS_CHECK_ARG(c != nullptr, "null context");
    if (bytes == 0) return GS_OK;
    GS_CHECK_ARG(ptr != nullptr, "gs_zero: null pointer");
    GS_HIP(launch_zero(ptr, bytes, c->stream), "zero kernel");
    return GS_OK;
}

// --------------------------------------------------------------- helpers ----
gsd::CamArgs make_cam_args(const gs_camera* cam, const gs_render_cfg* cfg) {
    gsd::CamArgs a;
    std::memcpy(a.view, cam->view_matrix, sizeof(a.view));
    std::memcpy(a.vp, cam->viewproj_matrix, sizeof(a.vp));
    std::memcpy(a.campos, cam->cam_pos, sizeof(a.campos));
    a.tanx = cam->tan_fovx;
    a.tany = cam->tan_fovy;
    // preprocessor_cpu.cpp:50-51 (host float arithmetic, identical on the device)
    a.fx = (float)cam->width / (2.0f * cam->tan_fovx);
    a.fy = (float)cam->height / (2.0f * cam->tan_fovy);
    a.W = cam->width;
    a.H = cam->height;
    a.gx = (cam->width + gsd::TILE - 1) / gsd::TILE;
    a.gy = (cam->height + gsd::TILE - 1) / gsd::TILE;
    a.scale_mod = cfg->scale_modifier;
    a.antialiasing = cfg->antialiasing;
    a.training = cfg->training;
    a.deg_cfg = cfg->sh_degree;
    return a;
}

int check_cam_cfg(const gs_camera* cam, const gs_render_cfg* cfg) {
    GS_CHECK_ARG(cam && cfg, "null camera or render config");
    GS_CHECK_ARG(cam->width > 0 && cam->height > 0, "camera size %dx%d", cam->width, cam->height);
    GS_CHECK_ARG(cfg->tile_w == 16 && cfg->tile_h == 16, "tiles must be 16x16 (got %dx%d)", cfg->tile_w, cfg->tile_h);
    return GS_OK;
}

static int check_gaussians(const gs_gaussians* g) {
    GS_CHECK_ARG(g != nullptr, "null gaussians");
    GS_CHECK_ARG(g->count >= 0, "negative Gaussian count");
    GS_CHECK_ARG(g->max_coeffs == 1 || g->max_coeffs == 4 || g->max_coeffs == 9 || g->max_coeffs == 16,
                 "max_coeffs must be 1, 4, 9 or 16 (got %d)", g->max_coeffs);
    GS_CHECK_ARG(g->sh_degree >= 0 && (g->sh_degree + 1) * (g->sh_degree + 1) <= g->max_coeffs,
                 "sh_degree %d does not fit max_coeffs %d", g->sh_degree, g->max_coeffs);
    if (g->count == 0) return GS_OK;
    GS_CHECK_ARG(g->positions && g->scales && g->rotations && g->sh_coeffs && g->opacities,
                 "null Gaussian parameter array");
    GS_CHECK_ARG(aligned(g->rotations, 16), "rotations must be 16-byte aligned");
    GS_CHECK_ARG(aligned(g->sh_coeffs, 16), "sh_coeffs must be 16-byte aligned");
    return GS_OK;
}

static int check_pre(const gs_preprocessed* p, bool need_all) {
    GS_CHECK_ARG(p != nullptr, "null preprocess buffers");
    GS_CHECK_ARG(p->means2D && p->conics && p->opacities && p->rgb, "null preprocess array");
    GS_CHECK_ARG(aligned(p->means2D, 8), "means2D must be 8-byte aligned");
    if (need_all) GS_CHECK_ARG(p->depths && p->radii && p->tiles_touched, "null preprocess array");
    return GS_OK;
}

// ---------------------------------------------------------- entry points ----
extern "C" int gs_activate(gs_ctx* c, const gs_gaussians* raw, float* pos, float* sc, float* rot, float* sh, float* op) {
    GS_CHECK_ARG(c != nullptr, "null context");
    int st = check_gaussians(raw);
    if (st) return st;
    if (raw->count == 0) return GS_OK;
    GS_CHECK_ARG(pos && sc && rot && sh && op, "null output");
    GS_CHECK_ARG(aligned(rot, 16), "rotations output must be 16-byte aligned");
    GS_HIP(launch_activate(*raw, pos, sc, rot, sh, op, c->stream), "activate kernel");
    return GS_OK;
}

extern "C" int gs_preprocess_fwd(gs_ctx* c, const gs_gaussians* g, const gs_camera* cam, const gs_render_cfg* cfg,
                                 const gs_preprocessed* out) {
    GS_CHECK_ARG(c != nullptr, "null context");
    int st = check_gaussians(g);
    if (st) return st;
    if ((st = check_cam_cfg(cam, cfg))) return st;
    if (g->count == 0) return GS_OK;
    if ((st = check_pre(out, true))) return st;
    gsd::CamArgs a = make_cam_args(cam, cfg);
    int deg = cfg->sh_degree < g->sh_degree ? cfg->sh_degree : g->sh_degree;  // preprocessor_cpu.cpp:55
    GS_HIP(launch_preprocess_fwd(*g, a, *out, deg, c->stream), "preprocess kernel");
    return GS_OK;
}

extern "C" int gs_scan_tiles(gs_ctx* c, int n, const int* tiles, uint32_t* offsets, uint32_t* total_dev,
                             long long* total_host) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(n >= 0, "negative count");
    void* sm;
    int st = gs_small_get(c, &sm);
    if (st) return st;
    unsigned long long* total64 = reinterpret_cast<unsigned long long*>(static_cast<char*>(sm) + GS_SMALL_TOTAL64);
    if (n == 0) {
        GS_HIP(launch_zero(total64, 8, c->stream), "zero kernel");
        if (total_dev) GS_HIP(launch_zero(total_dev, 4, c->stream), "zero kernel");
    } else {
        GS_CHECK_ARG(tiles && offsets, "null array");
        uint32_t np = scan_partials_needed((uint32_t)n);
        void* ws;
        if ((st = gs_ws_get(c, (size_t)(np + 1) * 8, &ws))) return st;
        GS_HIP(launch_scan_tiles(tiles, (uint32_t)n, offsets, static_cast<unsigned long long*>(ws), total_dev, total64,
                                 c->stream),
               "scan kernels");
    }
    if (total_host) {
        unsigned long long t = 0;
        GS_HIP(hipMemcpyAsync(&t, total64, 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(total)");
        GS_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        *total_host = (long long)t;
        GS_CHECK_ARG(t <= 0xFFFFFFFFull, "pair count %llu exceeds 2^32-1", t);
    }
    return GS_OK;
}

extern "C" int gs_scan_tiles_noalloc(gs_ctx* c, int n, const int* tiles, uint32_t* offsets, uint32_t* total_dev) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(n >= 0, "negative count");
    GS_CHECK_ARG((uint32_t)n <= GS_NOALLOC_SCAN_MAX,
                 "gs_scan_tiles_noalloc: %d Gaussians exceed its fixed scratch (at most %u; use gs_scan_tiles)", n,
                 (unsigned)GS_NOALLOC_SCAN_MAX);
    void* base = g_scratch_ptr[c->device].load();
    GS_CHECK_ARG(base != nullptr, "no scratch address for device %d", c->device);
    auto* total64 = static_cast<unsigned long long*>(base);
    auto* partial = reinterpret_cast<unsigned long long*>(static_cast<char*>(base) + 256);
    if (n == 0) {
        // fill kernels, not hipMemsetAsync: this call may be graph-captured
        GS_HIP(launch_zero(total64, 8, c->stream), "zero kernel");
        if (total_dev) GS_HIP(launch_zero(total_dev, 4, c->stream), "zero kernel");
        return GS_OK;
    }
    GS_CHECK_ARG(tiles && offsets, "null array");
    // every partial and the total are written before they are read: nothing to clear
    GS_HIP(launch_scan_tiles(tiles, (uint32_t)n, offsets, partial, total_dev, total64, c->stream), "scan kernels");
    return GS_OK;
}

extern "C" int gs_duplicate_keys(gs_ctx* c, int n, const float* means2D, const float* depths, const int* radii,
                                 const uint32_t* offsets, int width, int height, int tile_w, int tile_h,
                                 uint64_t* keys, uint32_t* values, uint32_t capacity) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(tile_w == 16 && tile_h == 16, "tiles must be 16x16");
    GS_CHECK_ARG(width > 0 && height > 0, "bad image size");
    if (n <= 0) return GS_OK;
    GS_CHECK_ARG(means2D && depths && radii && offsets && keys && values, "null array");
    GS_CHECK_ARG(aligned(means2D, 8), "means2D must be 8-byte aligned");
    void* sm;
    int st = gs_small_get(c, &sm);
    if (st) return st;
    uint32_t* ovf = reinterpret_cast<uint32_t*>(static_cast<char*>(sm) + GS_SMALL_OVERFLOW);
    int gx = (width + 15) / 16, gy = (height + 15) / 16;
    GS_HIP(launch_duplicate(n, means2D, depths, radii, offsets, gx, gy, keys, values, capacity, ovf, c->stream),
           "duplicate kernel");
    return GS_OK;
}

extern "C" int gs_check_overflow(gs_ctx* c, int* overflowed) {
    GS_CHECK_ARG(c != nullptr && overflowed != nullptr, "null argument");
    void* sm;
    int st = gs_small_get(c, &sm);
    if (st) return st;
    uint32_t v = 0;
    uint32_t* ovf = reinterpret_cast<uint32_t*>(static_cast<char*>(sm) + GS_SMALL_OVERFLOW);
    GS_HIP(hipMemcpyAsync(&v, ovf, 4, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(overflow)");
    GS_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    *overflowed = v != 0;
    if (v) GS_HIP(launch_zero(ovf, 4, c->stream), "zero kernel");
    return GS_OK;
}

extern "C" int gs_radix_sort_kv(gs_ctx* c, uint32_t n, const uint64_t* ki, const uint32_t* vi, uint64_t* ko,
                                uint32_t* vo, int key_bits, int radix_bits) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(key_bits >= 0 && key_bits <= 64, "key_bits must be in [0, 64]");
    if (radix_bits == 0) radix_bits = c->radix_bits;
    GS_CHECK_ARG(radix_bits >= 4 && radix_bits <= 8, "radix_bits must be in [4, 8]");
    if (n == 0) return GS_OK;
    GS_CHECK_ARG(ki && vi && ko && vo, "null array");
    GS_CHECK_ARG(ko != ki && vo != vi, "sort outputs must not alias the inputs");
    void* ws;
    int st = gs_ws_get(c, radix_workspace_bytes(n, radix_bits), &ws);
    if (st) return st;
    GS_HIP(launch_radix_sort(ki, vi, ko, vo, n, key_bits, radix_bits, ws, c->stream), "radix sort kernels");
    return GS_OK;
}

extern "C" int gs_tile_ranges(gs_ctx* c, uint32_t n, const uint64_t* keys_sorted, int num_tiles, uint32_t* ranges) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(num_tiles > 0 && ranges, "bad tile range buffer");
    GS_HIP(launch_zero(ranges, (size_t)num_tiles * 8, c->stream), "zero kernel");
    if (n == 0) return GS_OK;
    GS_CHECK_ARG(keys_sorted != nullptr, "null keys");
    GS_HIP(launch_tile_ranges(keys_sorted, n, (uint32_t)num_tiles, ranges, c->stream), "tile ranges kernel");
    return GS_OK;
}

int gs_tile_key_bits(int num_tiles) {
    int b = 0;
    while ((1ll << b) < (long long)num_tiles) b++;
    return 32 + b;
}

extern "C" int gs_sort_pairs(gs_ctx* c, uint32_t n, const uint64_t* ki, const uint32_t* vi, uint64_t* ko, uint32_t* vo,
                             int num_tiles, uint32_t* ranges) {
    int st = gs_radix_sort_kv(c, n, ki, vi, ko, vo, gs_tile_key_bits(num_tiles), 0);
    if (st) return st;
    return gs_tile_ranges(c, n, ko, num_tiles, ranges);
}

// Depth-ordered binning: depth sort of the Gaussians, pairs emitted in depth
// order as tile << 32 | index words, stable tile sort, tile ranges.
extern "C" int gs_bin_sorted_words(gs_ctx* c, int n, const gs_preprocessed* pre, int width, int height,
                                   uint64_t* words, uint32_t capacity, uint32_t* ranges, long long* total_host) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(width > 0 && height > 0, "bad image size");
    GS_CHECK_ARG(n >= 0, "negative count");
    GS_CHECK_ARG(ranges != nullptr, "null tile ranges");
    int st = check_pre(pre, true);
    if (st) return st;
    const int gx = (width + 15) / 16, gy = (height + 15) / 16;
    const int num_tiles = gx * gy;
    GS_HIP(launch_zero(ranges, (size_t)num_tiles * 8, c->stream), "zero kernel");
    if (n == 0) {
        if (total_host) *total_host = 0;
        return GS_OK;
    }
    void* sm;
    if ((st = gs_small_get(c, &sm))) return st;
    unsigned long long* total64 = reinterpret_cast<unsigned long long*>(static_cast<char*>(sm) + GS_SMALL_TOTAL64);
    uint32_t* ovf = reinterpret_cast<uint32_t*>(static_cast<char*>(sm) + GS_SMALL_OVERFLOW);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t un = (size_t)n;
    // Gaussian-sized arrays: depth keys/indices (+ sorted), ranked counts/offsets, scan partials, sort scratch
    const size_t np = scan_partials_needed((uint32_t)n) + 1;
    void* ws2;
    if ((st = gs_ws2_get(c, 2 * al(un * 8) + 4 * al(un * 4) + al(np * 8) + radix_workspace_bytes((uint32_t)n, 8),
                         &ws2)))
        return st;
    char* w = static_cast<char*>(ws2);
    uint64_t* dkeys = reinterpret_cast<uint64_t*>(w); w += al(un * 8);
    uint64_t* dkeys_s = reinterpret_cast<uint64_t*>(w); w += al(un * 8);
    uint32_t* didx = reinterpret_cast<uint32_t*>(w); w += al(un * 4);
    uint32_t* order = reinterpret_cast<uint32_t*>(w); w += al(un * 4);
    int* rtiles = reinterpret_cast<int*>(w); w += al(un * 4);
    uint32_t* roff = reinterpret_cast<uint32_t*>(w); w += al(un * 4);
    unsigned long long* partial = reinterpret_cast<unsigned long long*>(w); w += al(np * 8);
    void* dsort_ws = w;
    GS_HIP(launch_depth_keys(n, pre->depths, pre->radii, dkeys, didx, c->stream), "depth keys kernel");
    GS_HIP(launch_radix_sort_bits(dkeys, didx, dkeys_s, order, (uint32_t)n, 0, 32, 8, dsort_ws, c->stream, nullptr),
           "depth sort");
    GS_HIP(launch_gather_tiles(n, order, pre->tiles_touched, rtiles, c->stream), "gather kernel");
    GS_HIP(launch_scan_tiles(rtiles, (uint32_t)n, roff, partial, nullptr, total64, c->stream), "scan kernels");
    unsigned long long K = 0;
    GS_HIP(hipMemcpyAsync(&K, total64, 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(total)");
    GS_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (total_host) *total_host = (long long)K;
    if (K > capacity)
        return gs_set_error(GS_ERR_CAPACITY, "gs_bin_sorted_words: %llu pairs exceed capacity %u", K, capacity);
    if (K == 0) return GS_OK;
    GS_CHECK_ARG(words != nullptr, "null words buffer");
    // pair-sized arrays: the unsorted words and the tile sort's scratch
    const int tile_bits = gs_tile_key_bits(num_tiles) - 32;
    const int rb = c->radix_bits;
    void* ws;
    if ((st = gs_ws_get(c, al((size_t)K * 8) + radix_workspace_bytes((uint32_t)K, rb), &ws))) return st;
    uint64_t* unsorted = static_cast<uint64_t*>(ws);
    void* tsort_ws = static_cast<char*>(ws) + al((size_t)K * 8);
    GS_HIP(launch_duplicate_ranked(n, order, pre->means2D, pre->radii, roff, gx, gy, unsorted, (uint32_t)K, ovf,
                                   c->stream),
           "duplicate kernel");
    GS_HIP(launch_radix_sort_bits(unsorted, nullptr, words, nullptr, (uint32_t)K, 32, 32 + tile_bits, rb, tsort_ws,
                                  c->stream, nullptr),
           "tile sort");
    GS_HIP(launch_tile_ranges(words, (uint32_t)K, (uint32_t)num_tiles, ranges, c->stream), "tile ranges kernel");
    return GS_OK;
}

// Direct-placement binning knobs (0 keeps a value): Gaussians per group
// (larger groups shrink the (tile, group) table and lengthen each block's walk
// and its contiguous output runs) and tiles per lane (bins of 8 x 8*tpl tiles;
// larger bins mean fewer walks over the records, more work per visit).
extern "C" int gs_ctx_set_binning(gs_ctx* c, int gaussians_per_group, int tiles_per_lane) {
    GS_CHECK_ARG(c != nullptr, "null context");
    if (gaussians_per_group) {
        GS_CHECK_ARG(gaussians_per_group >= 256 && gaussians_per_group <= 65536 &&
                         (gaussians_per_group & (gaussians_per_group - 1)) == 0,
                     "group size must be a power of two in [256, 65536]");
        c->bin_chunk = gaussians_per_group;
    }
    if (tiles_per_lane) {
        GS_CHECK_ARG(tiles_per_lane == 1 || tiles_per_lane == 2 || tiles_per_lane == 4,
                     "tiles per lane must be 1, 2 or 4");
        c->bin_tpl = tiles_per_lane;
    }
    return GS_OK;
}

extern "C" int gs_ctx_get_tuning(gs_ctx* c, gs_tuning* out) {
    GS_CHECK_ARG(c && out, "null argument");
    *out = gs_tuning{};
    out->radix_bits = c->radix_bits;
    out->raster_batch = c->raster_batch;
    out->raster_bwd_batch = c->raster_bwd_batch;
    out->bin_group = c->bin_chunk;
    out->bin_tiles_per_lane = c->bin_tpl;
    out->persistent_per_cu = c->persistent_per_cu;
    out->split_groups = c->split_groups;
    return GS_OK;
}

extern "C" int gs_ctx_apply_tuning(gs_ctx* c, const gs_tuning* t) {
    GS_CHECK_ARG(c && t, "null argument");
    GS_CHECK_ARG(t->persistent_per_cu >= 0 && t->persistent_per_cu <= 64, "persistent_per_cu must be in [0, 64]");
    GS_CHECK_ARG(t->split_groups >= 0, "split_groups must be >= 0");
    int st = gs_ctx_set_tuning(c, t->radix_bits, t->raster_batch, t->raster_bwd_batch);
    if (st) return st;
    if ((st = gs_ctx_set_binning(c, t->bin_group, t->bin_tiles_per_lane))) return st;
    c->persistent_per_cu = t->persistent_per_cu;
    c->split_groups = t->split_groups;
    return GS_OK;
}

// gs_bin_tiles, and its form for drivers that keep the GPU busy: with
// total_async (the device address of mapped, coherent pinned host memory) the
// scan writes the pair count there without a host wait (`ev_total` is
// recorded after the scan), the placement runs with
// the given capacity whatever the count, pairs past it are dropped (overflow
// flag) and the ranges are clamped to it; the caller compares the count with
// the capacity before using the results.
int gs_bin_tiles_impl(gs_ctx* c, int n, const gs_preprocessed* pre, int width, int height, uint32_t* values,
                      uint32_t capacity, uint32_t* ranges, long long* total_host, unsigned long long* total_async,
                      hipEvent_t ev_total, int split) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(width > 0 && height > 0, "bad image size");
    GS_CHECK_ARG(n >= 0, "negative count");
    GS_CHECK_ARG(ranges != nullptr, "null tile ranges");
    int st = check_pre(pre, true);
    if (st) return st;
    const int gx = (width + 15) / 16, gy = (height + 15) / 16;
    const int num_tiles = gx * gy;
    if (gx > GS_BIN_MAX_TILES_XY || gy > GS_BIN_MAX_TILES_XY)
        return gs_set_error(GS_ERR_UNSUPPORTED, "gs_bin_tiles: %d x %d tiles exceed %d per side (use gs_bin_sorted_words)",
                            gx, gy, GS_BIN_MAX_TILES_XY);
    GS_CHECK_ARG(aligned(values, 16) || values == nullptr, "values must be 16-byte aligned");
    if (n == 0) {
        GS_HIP(launch_zero(ranges, (size_t)num_tiles * 8, c->stream), "zero kernel");
        if (total_host) *total_host = 0;
        return GS_OK;
    }
    void* sm;
    if ((st = gs_small_get(c, &sm))) return st;
    uint32_t* ovf = reinterpret_cast<uint32_t*>(static_cast<char*>(sm) + GS_SMALL_OVERFLOW);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t un = (size_t)n;
    // with a split, the first `split` groups are small (the nearest
    // Gaussians, placed before the first raster round: small groups give that
    // placement more blocks), the rest have the configured size
    const int G = c->bin_chunk;
    const int G1 = G < GS_SPLIT_GROUP ? G : GS_SPLIT_GROUP;
    const int S = (split > 0 && (long long)split * G1 < n) ? split : 0;
    const int C = S + (n - S * G1 + G - 1) / G;
    const size_t L = (size_t)C * num_tiles + 1;
    GS_CHECK_ARG(L < 0xFFFFFFFFull, "binning table too large");
    const size_t np = scan_partials_needed((uint32_t)L) + 1;
    void* ws2;
    if ((st = gs_ws2_get(c, 2 * al(un * 8) + al(un * 8) + al(un * 4) + radix_workspace_bytes((uint32_t)n, 8) +
                                2 * al(L * 4) + al(np * 8),
                         &ws2)))
        return st;
    char* w = static_cast<char*>(ws2);
    uint64_t* dkeys = reinterpret_cast<uint64_t*>(w); w += al(un * 8);    // index << 32 | depth key
    uint64_t* dkeys_s = reinterpret_cast<uint64_t*>(w); w += al(un * 8);  // the same, in depth order
    uint2* rects = reinterpret_cast<uint2*>(w); w += al(un * 8);
    uint32_t* rect_idx = reinterpret_cast<uint32_t*>(w); w += al(un * 4);  // packed rectangles, index order
    uint32_t* table = reinterpret_cast<uint32_t*>(w); w += al(L * 4);
    uint32_t* scanned = reinterpret_cast<uint32_t*>(w); w += al(L * 4);
    unsigned long long* partial = reinterpret_cast<unsigned long long*>(w); w += al(np * 8);
    void* dsort_ws = w;
    GS_HIP(launch_depth_words(n, pre->depths, pre->radii, dkeys, c->stream, pre->means2D, gx, gy, rect_idx),
           "depth keys kernel");
    GS_HIP(launch_radix_sort_bits(dkeys, nullptr, dkeys_s, nullptr, (uint32_t)n, 0, 32, c->radix_bits, dsort_ws,
                                  c->stream, nullptr),
           "depth sort");
    gsb::BinArgs a{};
    a.V = 1;
    a.N = n;
    a.C = C;
    a.G = G;
    a.S = S;
    a.G1 = G1;
    a.sidx = reinterpret_cast<const uint32_t*>(dkeys_s) + 1;  // the index words
    a.sidx_stride = 2;
    a.means2D = reinterpret_cast<const float2*>(pre->means2D);
    a.radii = pre->radii;
    a.views = nullptr;
    a.view0 = make_int4(gx, gy, 0, 0);
    a.vblock0 = nullptr;
    a.table = table;
    a.table_len = L;
    a.rects = rects;
    a.out = values;
    a.capacity = capacity;
    a.overflow = ovf;
    a.diff_len = (gx + 1) * (gy + 1);
    a.rect_idx = rect_idx;
    unsigned long long* total64 = reinterpret_cast<unsigned long long*>(static_cast<char*>(sm) + GS_SMALL_TOTAL64);
    const int blocks = bin_blocks(C, gx, gy, c->bin_tpl);
    GS_HIP(launch_bin_count(a, blocks, c->bin_tpl, c->stream), "binning count kernel");
    // with a split, only the first `split` depth groups are placed now (blocks
    // are numbered group-major: group c's blocks are c * bins .. c * bins + bins - 1)
    if (!c->bin_split) c->bin_split = new BinSplit();
    BinSplit* bs = static_cast<BinSplit*>(c->bin_split);
    bs->pending = false;
    const int bins = blocks / C;
    int place_blocks = blocks;
    if (S > 0) {
        place_blocks = S * bins;
        bs->pending = true;
        bs->tpl = c->bin_tpl;
        bs->bins = bins;
        bs->C = C;
        bs->split = S;
        bs->gx = gx;
        bs->gy = gy;
    }
    // asynchronous: the scan writes the pair count straight into the caller's
    // mapped host memory, and the event marks when it is there
    GS_HIP(launch_scan_u32_exclusive(table, (uint32_t)L, scanned, partial, c->stream, nullptr,
                                     total_async ? total_async : total64),
           "binning scan");
    a.table = scanned;
    bs->a = a;
    if (total_async) {
        GS_CHECK_ARG(values != nullptr && capacity > 0, "null values buffer");
        if (ev_total) GS_HIP(hipEventRecord(ev_total, c->stream), "hipEventRecord");
        GS_HIP(launch_bin_place(a, place_blocks, c->bin_tpl, c->stream), "binning place kernel");
        GS_HIP(launch_bin_ranges(a, scanned, (uint32_t)num_tiles, ranges, c->stream), "binning ranges kernel");
        return GS_OK;
    }
    unsigned long long K = 0;
    GS_HIP(hipMemcpyAsync(&K, total64, 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(total)");
    GS_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (total_host) *total_host = (long long)K;
    if (K > capacity) return gs_set_error(GS_ERR_CAPACITY, "gs_bin_tiles: %llu pairs exceed capacity %u", K, capacity);
    if (K > 0) {
        GS_CHECK_ARG(values != nullptr, "null values buffer");
        GS_HIP(launch_bin_place(a, place_blocks, c->bin_tpl, c->stream), "binning place kernel");
    } else {
        bs->pending = false;
    }
    GS_HIP(launch_bin_ranges(a, scanned, (uint32_t)num_tiles, ranges, c->stream), "binning ranges kernel");
    return GS_OK;
}

extern "C" int gs_bin_tiles(gs_ctx* c, int n, const gs_preprocessed* pre, int width, int height, uint32_t* values,
                            uint32_t capacity, uint32_t* ranges, long long* total_host) {
    return gs_bin_tiles_impl(c, n, pre, width, height, values, capacity, ranges, total_host, nullptr, nullptr);
}

int gs_bin_split_info(gs_ctx* c, gsr::RasterSplit* sp, int* num_bins) {
    GS_CHECK_ARG(c && sp, "null argument");
    *sp = gsr::RasterSplit{};
    const BinSplit* bs = static_cast<const BinSplit*>(c->bin_split);
    if (!bs || !bs->pending) return GS_OK;
    sp->scanned = bs->a.table;
    sp->C = bs->C;
    sp->g1 = bs->split;
    sp->round = 1;
    if (num_bins) *num_bins = bs->bins;
    return GS_OK;
}

int gs_bin_place_rest(gs_ctx* c, const uint32_t* tile_gate) {
    GS_CHECK_ARG(c != nullptr, "null context");
    BinSplit* bs = static_cast<BinSplit*>(c->bin_split);
    GS_CHECK_ARG(bs && bs->pending, "no split binning pending");
    GS_HIP(launch_bin_place(bs->a, (bs->C - bs->split) * bs->bins, bs->tpl, c->stream, bs->split * bs->bins, tile_gate),
           "binning place kernel (remaining groups)");
    bs->pending = false;
    return GS_OK;
}

static gsr::RasterArgs raster_args(int width, int height, const gs_preprocessed* pre, const uint32_t* values,
                                   const uint32_t* ranges, const float bg[3], int vstride = 1) {
    gsr::RasterArgs a;
    a.W = width;
    a.H = height;
    a.gx = (width + 15) / 16;
    a.means2D = reinterpret_cast<const float2*>(pre->means2D);
    a.conics = pre->conics;
    a.opac = pre->opacities;
    a.rgb = pre->rgb;
    a.values = values;
    a.ranges = ranges;
    a.bg0 = bg ? bg[0] : 0.0f;
    a.bg1 = bg ? bg[1] : 0.0f;
    a.bg2 = bg ? bg[2] : 0.0f;
    a.vstride = vstride;
    return a;
}

static int raster_fwd_impl(gs_ctx* c, int width, int height, const gs_preprocessed* pre, const uint32_t* values,
                           int vstride, const uint32_t* ranges, const float bg[3], float* image,
                           const gs_pixel_state* px, const gsr::RasterSplit* sp = nullptr) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(width > 0 && height > 0, "bad image size");
    int st = check_pre(pre, false);
    if (st) return st;
    GS_CHECK_ARG(ranges && image, "null ranges or image");
    int tiles = ((width + 15) / 16) * ((height + 15) / 16);
    gsr::RasterArgs a = raster_args(width, height, pre, values, ranges, bg, vstride);
    GS_HIP(launch_raster_fwd(a, tiles, c->raster_batch, image, px ? px->T_final : nullptr,
                             px ? px->n_contrib : nullptr, px ? px->last_contrib : nullptr, c->stream, sp),
           "raster forward kernel");
    return GS_OK;
}

int gs_raster_fwd_split(gs_ctx* c, int width, int height, const gs_preprocessed* pre, const uint32_t* values,
                        const uint32_t* ranges, const float bg[3], float* image, const gs_pixel_state* px,
                        const gsr::RasterSplit* sp) {
    GS_CHECK_ARG(px && px->T_final && px->n_contrib && px->last_contrib, "a split forward needs the pixel state");
    return raster_fwd_impl(c, width, height, pre, values, 1, ranges, bg, image, px, sp);
}

extern "C" int gs_raster_fwd(gs_ctx* c, int width, int height, const gs_preprocessed* pre, const uint32_t* values,
                             const uint32_t* ranges, const float bg[3], float* image, const gs_pixel_state* px) {
    return raster_fwd_impl(c, width, height, pre, values, 1, ranges, bg, image, px);
}

extern "C" int gs_raster_fwd_words(gs_ctx* c, int width, int height, const gs_preprocessed* pre,
                                   const uint64_t* words, const uint32_t* ranges, const float bg[3], float* image,
                                   const gs_pixel_state* px) {
    return raster_fwd_impl(c, width, height, pre, reinterpret_cast<const uint32_t*>(words), 2, ranges, bg, image, px);
}

extern "C" int gs_l1_loss(gs_ctx* c, int width, int height, const float* r, const float* g, float* d, double* loss) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(width > 0 && height > 0, "bad image size");
    GS_CHECK_ARG(r && g && d && loss, "null array");
    GS_CHECK_ARG(aligned(r, 16) && aligned(g, 16) && aligned(d, 16), "images must be 16-byte aligned");
    uint32_t n = (uint32_t)width * (uint32_t)height * 3u;
    void* ws;
    int st = gs_ws_get(c, (size_t)l1_partials_needed(n) * 8, &ws);
    if (st) return st;
    GS_HIP(launch_l1(r, g, d, n, static_cast<double*>(ws), loss, c->stream), "l1 kernels");
    return GS_OK;
}

static int raster_bwd_impl(gs_ctx* c, int width, int height, const gs_preprocessed* pre, const uint32_t* values,
                           int vstride, const uint32_t* ranges, const float bg[3], const gs_pixel_state* px,
                           const float* d_image, const gs_raster_grads* g) {
    GS_CHECK_ARG(c != nullptr, "null context");
    GS_CHECK_ARG(width > 0 && height > 0, "bad image size");
    int st = check_pre(pre, false);
    if (st) return st;
    GS_CHECK_ARG(ranges && px && px->T_final && px->last_contrib && d_image && g, "null array");
    GS_CHECK_ARG(g->d_means2D && g->d_conics && g->d_rgb && g->d_opacities_2d, "null gradient array");
    GS_CHECK_ARG(aligned(g->d_means2D, 8), "d_means2D must be 8-byte aligned");
    int tiles = ((width + 15) / 16) * ((height + 15) / 16);
    gsr::RasterArgs a = raster_args(width, height, pre, values, ranges, bg, vstride);
    gsr::GradArgs ga{reinterpret_cast<float2*>(g->d_means2D), g->d_conics, g->d_rgb, g->d_opacities_2d};
    GS_HIP(launch_raster_bwd(a, tiles, c->raster_bwd_batch, px->T_final, px->last_contrib, d_image, ga, c->stream),
           "raster backward kernel");
    return GS_OK;
}

extern "C" int gs_raster_bwd(gs_ctx* c, int width, int height, const gs_preprocessed* pre, const uint32_t* values,
                             const uint32_t* ranges, const float bg[3], const gs_pixel_state* px, const float* d_image,
                             const gs_raster_grads* g) {
    return raster_bwd_impl(c, width, height, pre, values, 1, ranges, bg, px, d_image, g);
}

extern "C" int gs_raster_bwd_words(gs_ctx* c, int width, int height, const gs_preprocessed* pre,
                                   const uint64_t* words, const uint32_t* ranges, const float bg[3],
                                   const gs_pixel_state* px, const float* d_image, const gs_raster_grads* g) {
    return raster_bwd_impl(c, width, height, pre, reinterpret_cast<const uint32_t*>(words), 2, ranges, bg, px,
                           d_image, g);
}

extern "C" int gs_preprocess_bwd(gs_ctx* c, const gs_gaussians* g, const gs_camera* cam, const gs_render_cfg* cfg,
                                 const int* radii, const gs_raster_grads* rg, const gs_param_grads* out,
                                 int accumulate) {
    GS_CHECK_ARG(c != nullptr, "null context");
    int st = check_gaussians(g);
    if (st) return st;
    if ((st = check_cam_cfg(cam, cfg))) return st;
    GS_CHECK_ARG(g->raw != 0, "gs_preprocess_bwd needs raw (pre-activation) parameters");
    if (g->count == 0) return GS_OK;
    GS_CHECK_ARG(radii && rg && out, "null argument");
    GS_CHECK_ARG(rg->d_means2D && rg->d_conics && rg->d_rgb && rg->d_opacities_2d, "null raster gradient");
    GS_CHECK_ARG(out->d_positions && out->d_scales && out->d_rotations && out->d_sh_coeffs && out->d_opacities,
                 "null output gradient");
    GS_CHECK_ARG(aligned(out->d_rotations, 16), "d_rotations must be 16-byte aligned");
    gsd::CamArgs a = make_cam_args(cam, cfg);
    // The reference evaluates the SH backward with the model degree
    // (preprocessor_backward_cpu.cpp:219-221), not min(cfg, model).
    GS_HIP(launch_preprocess_bwd(*g, a, radii, *rg, *out, accumulate, g->sh_degree, c->stream),
           "preprocess backward kernel");
    return GS_OK;
}

extern "C" int gs_adam_step(gs_ctx* c, int n, int max_coeffs, float* const params[5], const float* const grads[5],
                            float* const m[5], float* const v[5], const gs_adam_hparams* hp) {
    GS_CHECK_ARG(c != nullptr && hp != nullptr, "null argument");
    GS_CHECK_ARG(n >= 0, "negative count");
    if (n == 0) return GS_OK;
    const uint32_t sizes[5] = {(uint32_t)n * 3u, (uint32_t)n * 3u, (uint32_t)n * 4u,
                               (uint32_t)n * (uint32_t)max_coeffs * 3u, (uint32_t)n};
    gso::AdamArgs a;
    for (int k = 0; k < 5; k++) {
        GS_CHECK_ARG(params[k] && grads[k] && m[k] && v[k], "null Adam array (tensor %d)", k);
        GS_CHECK_ARG(aligned(params[k], 16) && aligned(grads[k], 16) && aligned(m[k], 16) && aligned(v[k], 16),
                     "Adam arrays must be 16-byte aligned (tensor %d)", k);
        a.seg[k] = gso::AdamSeg{params[k], grads[k], m[k], v[k], sizes[k], hp->lr[k]};
    }
    a.b1 = hp->beta1;
    a.b2 = hp->beta2;
    a.eps = hp->eps;
    a.bc1 = hp->bc1;
    a.bc2 = hp->bc2;
    a.hp = nullptr;
    GS_HIP(launch_adam(a, c->stream), "adam kernel");
    return GS_OK;
}

extern "C" int gs_sgd_step(gs_ctx* c, int n, int max_coeffs, float* const params[5], const float* const grads[5],
                           const float lr[5]) {
    GS_CHECK_ARG(c != nullptr && lr != nullptr, "null argument");
    GS_CHECK_ARG(n >= 0, "negative count");
    if (n == 0) return GS_OK;
    const uint32_t sizes[5] = {(uint32_t)n * 3u, (uint32_t)n * 3u, (uint32_t)n * 4u,
                               (uint32_t)n * (uint32_t)max_coeffs * 3u, (uint32_t)n};
    gso::AdamArgs a{};
    for (int k = 0; k < 5; k++) {
        GS_CHECK_ARG(params[k] && grads[k], "null SGD array (tensor %d)", k);
        a.seg[k] = gso::AdamSeg{params[k], grads[k], nullptr, nullptr, sizes[k], lr[k]};
    }
    GS_HIP(launch_sgd(a, c->stream), "sgd kernel");
    return GS_OK;
}

// train_types.cpp:64-68 and optimizer.cpp:73-76, host float arithmetic.
extern "C" float gs_lr_schedule(float lr_init, float lr_final, int step, int max_steps) {
    float t = static_cast<float>(step) / static_cast<float>(max_steps);
    t = std::max(0.0f, std::min(1.0f, t));
    return std::exp(std::log(lr_init) * (1.0f - t) + std::log(lr_final) * t);
}

extern "C" void gs_adam_bias_corrections(float b1, float b2, int step, float* bc1, float* bc2) {
    *bc1 = 1.0f / (1.0f - std::pow(b1, (float)step));
    *bc2 = 1.0f / (1.0f - std::pow(b2, (float)step));
}
