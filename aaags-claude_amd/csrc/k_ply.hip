// k_ply.hip -- loading a 3DGS PLY file straight into device arrays
// (loadPly, ply_loader.cpp:32-189): the host parses the header and reads the
// binary vertex block; one kernel unpacks it, one thread per vertex, into the
// activated SoA layout of GaussianData: exp(scale), sigmoid(opacity),
// normalised rotation, and the SH reordered from the file's channel-major
// f_rest (all R, then G, then B) to basis-interleaved sh[k][ch].
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "gs_device.h"
#include "gs_internal.h"

namespace gsp {
using namespace gsd;

constexpr int MAX_REST = 45;

struct PlyLayout {
    int props;  // floats per vertex
    int n;
    int x, y, z, dc[3], opacity, s[3], r[4], filter;
    int rest_per_channel, max_coeffs;
    int rest[MAX_REST];
};

// ply_loader.cpp:132-186, per vertex
__global__ __launch_bounds__(256) void ply_unpack_kernel(PlyLayout L, const float* __restrict__ raw,
                                                         float* __restrict__ pos, float* __restrict__ scale,
                                                         float* __restrict__ rot, float* __restrict__ sh,
                                                         float* __restrict__ op, float* __restrict__ filter) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= L.n) return;
    const float* v = raw + (size_t)i * L.props;
    pos[i * 3] = v[L.x];
    pos[i * 3 + 1] = v[L.y];
    pos[i * 3 + 2] = v[L.z];
    float* o = sh + (size_t)i * L.max_coeffs * 3;
    for (int k = 0; k < L.max_coeffs * 3; k++) o[k] = 0.0f;
    o[0] = v[L.dc[0]];
    o[1] = v[L.dc[1]];
    o[2] = v[L.dc[2]];
    const int R = L.rest_per_channel;
    for (int k = 0; k < R; k++) {
        o[3 + k * 3] = v[L.rest[k]];
        o[3 + k * 3 + 1] = v[L.rest[R + k]];
        o[3 + k * 3 + 2] = v[L.rest[2 * R + k]];
    }
    op[i] = 1.0f / (1.0f + exp_ref(-v[L.opacity]));
    for (int k = 0; k < 3; k++) scale[i * 3 + k] = exp_ref(v[L.s[k]]);
    float r0 = v[L.r[0]], r1 = v[L.r[1]], r2 = v[L.r[2]], r3 = v[L.r[3]];
    const float len = sqrtf(r0 * r0 + r1 * r1 + r2 * r2 + r3 * r3);
    if (len > 0.0f) {
        const float inv = 1.0f / len;
        r0 *= inv; r1 *= inv; r2 *= inv; r3 *= inv;
    }
    reinterpret_cast<float4*>(rot)[i] = make_float4(r0, r1, r2, r3);
    if (filter && L.filter >= 0) filter[i] = v[L.filter];
}

// train_main.cpp:108-141, per vertex: x y z, f_dc, f_rest channel-major,
// opacity, scale, rot -- the raw (pre-activation) values.
__global__ __launch_bounds__(256) void ply_pack_kernel(int n, int max_coeffs, const float* __restrict__ pos,
                                                       const float* __restrict__ scale, const float* __restrict__ rot,
                                                       const float* __restrict__ sh, const float* __restrict__ op,
                                                       float* __restrict__ rows) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int R = max_coeffs - 1;
    const int props = 3 + 3 + 3 * R + 1 + 3 + 4;
    float* v = rows + (size_t)i * props;
    const float* s = sh + (size_t)i * max_coeffs * 3;
    int k = 0;
    v[k++] = pos[i * 3]; v[k++] = pos[i * 3 + 1]; v[k++] = pos[i * 3 + 2];
    v[k++] = s[0]; v[k++] = s[1]; v[k++] = s[2];
    for (int c = 0; c < 3; c++)
        for (int b = 0; b < R; b++) v[k++] = s[3 + b * 3 + c];
    v[k++] = op[i];
    v[k++] = scale[i * 3]; v[k++] = scale[i * 3 + 1]; v[k++] = scale[i * 3 + 2];
    v[k++] = rot[i * 4]; v[k++] = rot[i * 4 + 1]; v[k++] = rot[i * 4 + 2]; v[k++] = rot[i * 4 + 3];
}

// ply_loader.cpp:32-121: the header rules (first line "ply"; "element vertex
// N"; "property float NAME" lines after it; required properties; the SH
// degree from the count of f_rest_*).  Returns an error status or GS_OK; the
// stream is left at the start of the vertex block.
static int parse_header(std::ifstream& f, const char* path, PlyLayout& L) {
    if (!f.is_open()) return gs_set_error(GS_ERR_INVALID, "failed to open PLY file: %s", path);
    std::string line;
    std::getline(f, line);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line != "ply") return gs_set_error(GS_ERR_INVALID, "not a PLY file: %s", path);
    int count = 0;
    bool vertex = false;
    std::vector<std::string> names;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == "end_header") break;
        if (line.compare(0, 15, "element vertex ") == 0) {
            count = std::atoi(line.c_str() + 15);
            vertex = true;
        }
        if (vertex && line.compare(0, 15, "property float ") == 0) names.push_back(line.substr(15));
    }
    if (!vertex || count <= 0) return gs_set_error(GS_ERR_INVALID, "PLY header: no valid vertex count");
    std::unordered_map<std::string, int> at;
    for (int k = 0; k < (int)names.size(); k++) at[names[k]] = k;
    const char* req[] = {"x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
                         "rot_0", "rot_1", "rot_2", "rot_3"};
    int idx[14];
    for (int k = 0; k < 14; k++) {
        auto it = at.find(req[k]);
        if (it == at.end()) return gs_set_error(GS_ERR_INVALID, "PLY missing required property: %s", req[k]);
        idx[k] = it->second;
    }
    std::memset(&L, 0, sizeof(L));
    L.props = (int)names.size();
    L.n = count;
    L.x = idx[0]; L.y = idx[1]; L.z = idx[2];
    L.dc[0] = idx[3]; L.dc[1] = idx[4]; L.dc[2] = idx[5];
    L.opacity = idx[6];
    L.s[0] = idx[7]; L.s[1] = idx[8]; L.s[2] = idx[9];
    L.r[0] = idx[10]; L.r[1] = idx[11]; L.r[2] = idx[12]; L.r[3] = idx[13];
    int rest = 0;
    while (rest < MAX_REST && at.count("f_rest_" + std::to_string(rest))) rest++;
    int deg = rest >= 45 ? 3 : rest >= 24 ? 2 : rest >= 9 ? 1 : 0;
    L.rest_per_channel = deg == 3 ? 15 : deg == 2 ? 8 : deg == 1 ? 3 : 0;
    L.max_coeffs = (deg + 1) * (deg + 1);
    for (int k = 0; k < rest; k++) L.rest[k] = at["f_rest_" + std::to_string(k)];
    auto fi = at.find("filter_3D");
    L.filter = fi == at.end() ? -1 : fi->second;
    return GS_OK;
}

// initRawFromGaussianData (train_main.cpp:31-63) after the activation of
// loadPly: the log of the activated scale (floored at 1e-12) and the inverse
// sigmoid of the clamped opacity, with the reference's logf (log_ref).
__global__ __launch_bounds__(256) void raw_from_activated_kernel(int n, float* __restrict__ scale,
                                                                 float* __restrict__ op) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; k++) scale[i * 3 + k] = log_ref(smax(1e-12f, scale[i * 3 + k]));
    const float y = smax(1e-6f, smin(1.0f - 1e-6f, op[i]));
    op[i] = log_ref(y / (1.0f - y));
}

}  // namespace gsp

extern "C" int gs_ply_info(const char* path, int* count, int* sh_degree, int* max_coeffs, int* has_filter_3d) {
    if (!path || !count || !sh_degree || !max_coeffs) return gs_set_error(GS_ERR_INVALID, "null argument");
    std::ifstream f(path, std::ios::binary);
    gsp::PlyLayout L;
    int st = gsp::parse_header(f, path, L);
    if (st) return st;
    *count = L.n;
    *max_coeffs = L.max_coeffs;
    *sh_degree = L.max_coeffs == 16 ? 3 : L.max_coeffs == 9 ? 2 : L.max_coeffs == 4 ? 1 : 0;
    if (has_filter_3d) *has_filter_3d = L.filter >= 0;
    return GS_OK;
}

extern "C" int gs_ply_load(gs_ctx* c, const char* path, int n, int max_coeffs, float* positions, float* scales,
                           float* rotations, float* sh_coeffs, float* opacities, float* filter_3d) {
    if (!path) return gs_set_error(GS_ERR_INVALID, "null path");
    if (!positions || !scales || !rotations || !sh_coeffs || !opacities)
        return gs_set_error(GS_ERR_INVALID, "null output array");
    if (reinterpret_cast<uintptr_t>(rotations) % 16) return gs_set_error(GS_ERR_INVALID, "rotations must be 16-byte aligned");
    std::ifstream f(path, std::ios::binary);
    gsp::PlyLayout L;
    int st = gsp::parse_header(f, path, L);
    if (st) return st;
    // the caller's arrays were sized for (n, max_coeffs, filter): a file whose
    // header says otherwise is refused before anything reaches the device
    if (L.n != n || L.max_coeffs != max_coeffs)
        return gs_set_error(GS_ERR_INVALID, "PLY %s holds %d Gaussians with %d SH coefficients; the arrays are for %d and %d",
                            path, L.n, L.max_coeffs, n, max_coeffs);
    if (filter_3d && L.filter < 0)
        return gs_set_error(GS_ERR_INVALID, "PLY %s has no filter_3D property but a filter_3d array was passed", path);
    if (!c) return gs_set_error(GS_ERR_INVALID, "null context");
    const size_t bytes = (size_t)L.n * L.props * 4;
    void* host = nullptr;
    hipError_t e = hipHostMalloc(&host, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return gs_hip_status(e, "hipHostMalloc(PLY vertices)");
    f.read(static_cast<char*>(host), (std::streamsize)bytes);
    if (!f) {
        (void)hipHostFree(host);
        return gs_set_error(GS_ERR_INVALID, "PLY: failed to read all vertex data");
    }
    void* dev;
    if ((st = gs_ws_get(c, bytes, &dev))) {
        (void)hipHostFree(host);
        return st;
    }
    e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        gsp::ply_unpack_kernel<<<(L.n + 255) / 256, 256, 0, c->stream>>>(L, static_cast<const float*>(dev), positions,
                                                                          scales, rotations, sh_coeffs, opacities,
                                                                          filter_3d);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(c->stream);  // the pinned block is freed next
    (void)hipHostFree(host);
    if (e != hipSuccess) return gs_hip_status(e, "PLY upload / unpack kernel");
    return gs_hip_status(e2, "hipStreamSynchronize");
}

extern "C" int gs_ply_load_raw(gs_ctx* c, const char* path, int n, int max_coeffs, float* positions,
                               float* log_scales, float* rotations, float* sh_coeffs, float* opacity_logits) {
    int st = gs_ply_load(c, path, n, max_coeffs, positions, log_scales, rotations, sh_coeffs, opacity_logits, nullptr);
    if (st) return st;
    gsp::raw_from_activated_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(n, log_scales, opacity_logits);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gs_hip_status(e, "raw parameter kernel");
    return gs_hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

extern "C" int gs_ply_save(gs_ctx* c, const char* path, int n, int max_coeffs, const float* positions,
                           const float* scales, const float* rotations, const float* sh_coeffs,
                           const float* opacities) {
    if (!c || !path) return gs_set_error(GS_ERR_INVALID, "null argument");
    if (n <= 0) return gs_set_error(GS_ERR_INVALID, "gs_ply_save: no Gaussians");
    if (!(max_coeffs == 1 || max_coeffs == 4 || max_coeffs == 9 || max_coeffs == 16))
        return gs_set_error(GS_ERR_INVALID, "max_coeffs must be 1, 4, 9 or 16 (got %d)", max_coeffs);
    if (!positions || !scales || !rotations || !sh_coeffs || !opacities)
        return gs_set_error(GS_ERR_INVALID, "null parameter array");
    const int R = max_coeffs - 1;
    const int props = 3 + 3 + 3 * R + 1 + 3 + 4;
    const size_t bytes = (size_t)n * props * 4;
    void* dev;
    int st = gs_ws_get(c, bytes, &dev);
    if (st) return st;
    void* host = nullptr;
    hipError_t e = hipHostMalloc(&host, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return gs_hip_status(e, "hipHostMalloc(PLY vertices)");
    gsp::ply_pack_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(n, max_coeffs, positions, scales, rotations,
                                                                   sh_coeffs, opacities, static_cast<float*>(dev));
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipHostFree(host);
        return gs_hip_status(e, "PLY pack / download");
    }
    // train_main.cpp:90-106: the header
    std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) + "\n";
    for (const char* nm : {"x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"}) h += std::string("property float ") + nm + "\n";
    for (int k = 0; k < 3 * R; k++) h += "property float f_rest_" + std::to_string(k) + "\n";
    for (const char* nm : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"})
        h += std::string("property float ") + nm + "\n";
    h += "end_header\n";
    std::ofstream f(path, std::ios::binary);
    bool ok = f.is_open();
    if (ok) {
        f.write(h.data(), (std::streamsize)h.size());
        f.write(static_cast<const char*>(host), (std::streamsize)bytes);
        ok = (bool)f;
    }
    (void)hipHostFree(host);
    if (!ok) return gs_set_error(GS_ERR_INVALID, "failed to write PLY file: %s", path);
    return GS_OK;
}
