// gs_io.cpp -- the on-disk formats around the training path and the
// multi-view training driver:
//   gs_cameras_json ......... loadCamerasJson (harmonyos_3dgs/src/train_main.cpp:262-377;
//                             the renderer's loadCameraJson, src/main.cpp:146-264, reads the same format)
//   gs_read_ppm / gs_write_ppm  readPPM / writePPM (src/image_io.cpp)
//   gs_trainer_train_views .. the training loop of train_main.cpp:633-673 (random view per
//                             iteration, checkpoints) over the device-resident trainer
#include <hip/hip_runtime.h>

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "gs_internal.h"
#include "gs_trainer_impl.h"

namespace {

// ---- a small JSON reader: values, arrays and objects of the 3DGS camera file
struct JVal {
    enum Kind { Null, Num, Str, Arr, Obj } kind = Null;
    double num = 0.0;
    std::string str;
    std::vector<JVal> items;                              // array elements
    std::vector<std::pair<std::string, JVal>> members;    // object members, in file order
    const JVal* get(const char* key) const {
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

struct JReader {
    const char* p;
    const char* end;
    bool ok = true;
    void ws() {
        while (p < end && std::isspace((unsigned char)*p)) p++;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    std::string string() {
        std::string out;
        p++;  // opening quote
        while (p < end && *p != '"') {
            if (*p == '\\' && p + 1 < end) p++;
            out.push_back(*p++);
        }
        if (p >= end) ok = false;
        else p++;
        return out;
    }
    JVal value(int depth) {
        JVal v;
        ws();
        if (p >= end || depth > 64) {
            ok = false;
            return v;
        }
        if (*p == '{') {
            v.kind = JVal::Obj;
            p++;
            ws();
            if (p < end && *p == '}') {
                p++;
                return v;
            }
            while (ok) {
                ws();
                if (p >= end || *p != '"') {
                    ok = false;
                    break;
                }
                std::string key = string();
                ws();
                if (p >= end || *p != ':') {
                    ok = false;
                    break;
                }
                p++;
                v.members.emplace_back(key, value(depth + 1));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == '}') {
                    p++;
                    break;
                }
                ok = false;
            }
        } else if (*p == '[') {
            v.kind = JVal::Arr;
            p++;
            ws();
            if (p < end && *p == ']') {
                p++;
                return v;
            }
            while (ok) {
                v.items.push_back(value(depth + 1));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == ']') {
                    p++;
                    break;
                }
                ok = false;
            }
        } else if (*p == '"') {
            v.kind = JVal::Str;
            v.str = string();
        } else if (lit("true") || lit("false") || lit("null")) {
            v.kind = JVal::Null;
        } else {
            // a number, converted as std::stod converts it
            char* stop = nullptr;
            std::string tok;
            while (p < end && (std::isdigit((unsigned char)*p) || *p == '-' || *p == '+' || *p == '.' || *p == 'e' ||
                               *p == 'E'))
                tok.push_back(*p++);
            if (tok.empty()) {
                ok = false;
                return v;
            }
            v.kind = JVal::Num;
            v.num = std::strtod(tok.c_str(), &stop);
            if (!stop || *stop != '\0') ok = false;
        }
        return v;
    }
};

// buildPerspective (train_main.cpp:181-189), column-major
void perspective(float tan_fovx, float tan_fovy, float znear, float zfar, float proj[16]) {
    for (int k = 0; k < 16; k++) proj[k] = 0.0f;
    proj[0] = 1.0f / tan_fovx;
    proj[5] = 1.0f / tan_fovy;
    proj[10] = zfar / (zfar - znear);
    proj[11] = 1.0f;
    proj[14] = -(zfar * znear) / (zfar - znear);
}

// out = A * B for column-major 4x4 (mat4Mul, train_main.cpp:191-199): each
// entry summed from zero over k in order
void mat4_mul(const float A[16], const float B[16], float out[16]) {
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc += A[k * 4 + row] * B[col * 4 + k];
            out[col * 4 + row] = acc;
        }
}

bool num_of(const JVal* v, double* out) {
    if (!v || v->kind != JVal::Num) return false;
    *out = v->num;
    return true;
}

// One camera object of the file: id, width, height, fx, fy, position [3] and
// the camera-to-world rotation [3][3]; false when it is not a camera.
bool camera_of(const JVal& o, gs_camera* cam, std::string* img_name) {
    double id, w, h, fx, fy;
    if (!num_of(o.get("id"), &id)) return false;
    const JVal* pos = o.get("position");
    const JVal* rot = o.get("rotation");
    if (!num_of(o.get("fx"), &fx) || !num_of(o.get("fy"), &fy)) return false;
    const float ffx = (float)fx, ffy = (float)fy;
    if (ffx <= 0.0f || ffy <= 0.0f) return false;  // train_main.cpp:318
    if (!pos || pos->kind != JVal::Arr || pos->items.size() < 3) return false;
    if (!rot || rot->kind != JVal::Arr || rot->items.size() < 3) return false;
    w = h = 0.0;
    num_of(o.get("width"), &w);
    num_of(o.get("height"), &h);
    gs_camera c{};
    c.width = (int)w;
    c.height = (int)h;
    c.tan_fovx = (float)c.width / (2.0f * ffx);
    c.tan_fovy = (float)c.height / (2.0f * ffy);
    for (int k = 0; k < 3; k++) {
        double v;
        if (!num_of(&pos->items[k], &v)) return false;
        c.cam_pos[k] = (float)v;
    }
    float R[3][3];  // camera to world, row-major as stored
    for (int r = 0; r < 3; r++) {
        const JVal& row = rot->items[r];
        if (row.kind != JVal::Arr || row.items.size() < 3) return false;
        for (int k = 0; k < 3; k++) {
            double v;
            if (!num_of(&row.items[k], &v)) return false;
            R[r][k] = (float)v;
        }
    }
    // world to camera: the transpose, and t = -R^T * centre (train_main.cpp:346-356)
    float t[3];
    for (int j = 0; j < 3; j++)
        t[j] = -(R[0][j] * c.cam_pos[0] + R[1][j] * c.cam_pos[1] + R[2][j] * c.cam_pos[2]);
    for (int col = 0; col < 3; col++) {
        for (int row = 0; row < 3; row++) c.view_matrix[col * 4 + row] = R[col][row];
        c.view_matrix[col * 4 + 3] = 0.0f;
    }
    for (int row = 0; row < 3; row++) c.view_matrix[12 + row] = t[row];
    c.view_matrix[15] = 1.0f;
    float proj[16];
    perspective(c.tan_fovx, c.tan_fovy, 0.01f, 100.0f, proj);
    mat4_mul(proj, c.view_matrix, c.viewproj_matrix);
    *cam = c;
    if (img_name) {
        const JVal* nm = o.get("img_name");
        *img_name = (nm && nm->kind == JVal::Str) ? nm->str : std::string();
    }
    return true;
}

// Camera objects in file order, at any depth (the file is normally one array
// of them).
void collect(const JVal& v, std::vector<std::pair<gs_camera, std::string>>& out) {
    if (v.kind == JVal::Obj) {
        gs_camera c;
        std::string nm;
        if (camera_of(v, &c, &nm)) {
            out.emplace_back(c, nm);
            return;
        }
        for (const auto& m : v.members) collect(m.second, out);
    } else if (v.kind == JVal::Arr) {
        for (const auto& it : v.items) collect(it, out);
    }
}

}  // namespace

extern "C" int gs_cameras_json(const char* path, gs_camera* cams, int max_cams, int* count, char* img_names,
                               int name_stride) {
    if (!path || !count) return gs_set_error(GS_ERR_INVALID, "gs_cameras_json: null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return gs_set_error(GS_ERR_INVALID, "cannot open %s", path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JReader rd{text.data(), text.data() + text.size()};
    JVal root = rd.value(0);
    if (!rd.ok) return gs_set_error(GS_ERR_INVALID, "%s: malformed JSON near byte %lld", path,
                                    (long long)(rd.p - text.data()));
    std::vector<std::pair<gs_camera, std::string>> found;
    collect(root, found);
    *count = (int)found.size();
    if (!cams) return GS_OK;
    if (max_cams < (int)found.size())
        return gs_set_error(GS_ERR_INVALID, "%s holds %d cameras; room for %d", path, (int)found.size(), max_cams);
    for (size_t k = 0; k < found.size(); k++) {
        cams[k] = found[k].first;
        if (img_names && name_stride > 0) {
            char* dst = img_names + k * (size_t)name_stride;
            std::snprintf(dst, (size_t)name_stride, "%s", found[k].second.c_str());
        }
    }
    return GS_OK;
}

extern "C" int gs_read_ppm(const char* path, float* image, size_t max_floats, int* width, int* height) {
    if (!path || !width || !height) return gs_set_error(GS_ERR_INVALID, "gs_read_ppm: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return gs_set_error(GS_ERR_INVALID, "cannot open %s", path);
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    // P6 header: magic, then width, height and maxval, with '#' comments
    // skipped before the sizes (image_io.cpp readPPM)
    char magic[3] = {0, 0, 0};
    if (std::fscanf(f, "%2s", magic) != 1 || std::strcmp(magic, "P6") != 0)
        return gs_set_error(GS_ERR_INVALID, "%s: not a binary PPM (P6)", path);
    for (int ch = std::fgetc(f); ch != EOF; ch = std::fgetc(f)) {
        if (ch == '#') {
            while ((ch = std::fgetc(f)) != EOF && ch != '\n') {
            }
        } else if (ch > ' ') {
            std::ungetc(ch, f);
            break;
        }
    }
    int w = 0, h = 0, maxval = 0;
    if (std::fscanf(f, "%d %d %d", &w, &h, &maxval) != 3 || w <= 0 || h <= 0 || maxval <= 0 || maxval > 255)
        return gs_set_error(GS_ERR_INVALID, "%s: bad PPM header", path);
    std::fgetc(f);  // the single whitespace after maxval
    *width = w;
    *height = h;
    if (!image) return GS_OK;
    const size_t n = (size_t)w * h * 3;
    if (max_floats < n) return gs_set_error(GS_ERR_INVALID, "%s: %dx%d does not fit %zu floats", path, w, h, max_floats);
    std::vector<unsigned char> raw(n);
    if (std::fread(raw.data(), 1, n, f) != n) return gs_set_error(GS_ERR_INVALID, "%s: truncated pixel data", path);
    const float scale = 1.0f / (float)maxval;
    for (size_t i = 0; i < n; i++) image[i] = (float)raw[i] * scale;
    return GS_OK;
}

extern "C" int gs_write_ppm(const char* path, const float* image, int width, int height) {
    if (!path || !image || width <= 0 || height <= 0) return gs_set_error(GS_ERR_INVALID, "gs_write_ppm: bad argument");
    const size_t n = (size_t)width * height * 3;
    std::vector<unsigned char> bytes(n);
    for (size_t i = 0; i < n; i++) {
        float v = image[i];
        v = v < 1.0f ? v : 1.0f;  // std::min(1.0f, v) (a NaN becomes 1)
        v = 0.0f < v ? v : 0.0f;  // std::max(0.0f, v)
        bytes[i] = (unsigned char)(v * 255.0f + 0.5f);
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return gs_set_error(GS_ERR_INVALID, "cannot write %s", path);
    std::fprintf(f, "P6\n%d %d\n255\n", width, height);
    const bool ok = std::fwrite(bytes.data(), 1, n, f) == n;
    std::fclose(f);
    return ok ? GS_OK : gs_set_error(GS_ERR_INVALID, "failed to write %s", path);
}

extern "C" int gs_trainer_train_views(gs_trainer* t, int n_views, const gs_camera* cams, const float* const* targets,
                                      const gs_render_cfg* cfg, const gs_train_cfg* tc, int first_iteration,
                                      int iterations, unsigned seed, int save_every, const char* output_path,
                                      int log_every, double* losses, int* views) {
    if (!t || n_views <= 0 || !cams || !targets || !cfg || !tc || iterations < 0)
        return gs_set_error(GS_ERR_INVALID, "gs_trainer_train_views: bad argument");
    if (save_every > 0 && !output_path)
        return gs_set_error(GS_ERR_INVALID, "gs_trainer_train_views: checkpoints need an output path");
    // train_main.cpp:613-615: one generator for the whole run, a view drawn
    // per iteration when there are several
    std::mt19937 rng(seed);
    std::uniform_int_distribution<int> pick(0, n_views - 1);
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < iterations; k++) {
        const int it = first_iteration + k;
        const int v = n_views > 1 ? pick(rng) : 0;
        gs_step_result r;
        int st = gs_trainer_step(t, &cams[v], cfg, targets[v], tc, it, &r);
        if (st) return st;
        if (losses) losses[k] = r.loss;
        if (views) views[k] = v;
        if (log_every > 0 && (k == 0 || (k + 1) % log_every == 0 || k == iterations - 1)) {
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("iter %5d/%d  loss=%.6f  view=%d  (%.1f it/s)\n", k + 1, iterations, r.loss, v,
                        (k + 1) / (sec > 0 ? sec : 1e-9));
        }
        if (save_every > 0 && (k + 1) % save_every == 0) {  // train_main.cpp:661-666
            const std::string ck = std::string(output_path) + ".iter" + std::to_string(k + 1) + ".ply";
            gs_gaussians g = t->gaussians();
            st = gs_ply_save(t->ctx, ck.c_str(), t->n, t->max_coeffs, g.positions, g.scales, g.rotations, g.sh_coeffs,
                             g.opacities);
            if (st) return st;
        }
    }
    if (output_path) {
        gs_gaussians g = t->gaussians();
        return gs_ply_save(t->ctx, output_path, t->n, t->max_coeffs, g.positions, g.scales, g.rotations, g.sh_coeffs,
                           g.opacities);
    }
    return GS_OK;
}
