// ref_train_main.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the
// product).  Compiles the reference's own src/train_main.cpp, where it lies,
// as one more translation unit of oracle/_ref/libgs_ref.so (its main renamed),
// so the tests can call its file-local camera and parameter loaders:
// loadCamerasJson (train_main.cpp:262-377) and loadPly followed by
// initRawFromGaussianData (:31-63).
#define main ref_train_main_unused
#include "train_main.cpp"
#undef main

extern "C" {

// loadCamerasJson: *count cameras; when out is non-NULL (room for max), each
// camera's view (16), viewproj (16), cam_pos (3), tan_fovx, tan_fovy, width,
// height as 39 floats.
int ref_cameras_json(const char* path, float* out, int max, int* count) {
    try {
        std::vector<CameraWithImage> cams = loadCamerasJson(path);
        *count = (int)cams.size();
        if (!out) return 0;
        if ((int)cams.size() > max) return -1;
        for (size_t k = 0; k < cams.size(); k++) {
            const Camera& c = cams[k].cam;
            float* o = out + 39 * k;
            std::memcpy(o, c.view_matrix, 64);
            std::memcpy(o + 16, c.viewproj_matrix, 64);
            std::memcpy(o + 32, c.cam_pos, 12);
            o[35] = c.tan_fovx;
            o[36] = c.tan_fovy;
            o[37] = (float)c.width;
            o[38] = (float)c.height;
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

// loadPly + initRawFromGaussianData into caller arrays sized from *count /
// *max_coeffs (call with null arrays first to get them).
int ref_ply_raw(const char* path, int* count, int* max_coeffs, float* pos, float* sc, float* rot, float* sh,
                float* op) {
    try {
        LoadedModel m = loadPly(path);
        *count = m.data.count;
        *max_coeffs = m.data.max_coeffs;
        if (pos) {
            RawGaussianParams r = initRawFromGaussianData(m.data);
            const size_t n = (size_t)r.count;
            std::memcpy(pos, r.raw_positions, n * 12);
            std::memcpy(sc, r.raw_scales, n * 12);
            std::memcpy(rot, r.raw_rotations, n * 16);
            std::memcpy(sh, r.raw_sh_coeffs, n * r.max_coeffs * 12);
            std::memcpy(op, r.raw_opacities, n * 4);
            freeRawParams(r);
        }
        m.free();
        return 0;
    } catch (...) {
        return -1;
    }
}

// readPPM / writePPM (image_io.cpp)
int ref_read_ppm(const char* path, float* out, size_t max, int* w, int* h) {
    std::vector<float> img = readPPM(path, *w, *h);
    if (img.empty()) return -1;
    if (out) {
        if (img.size() > max) return -1;
        std::memcpy(out, img.data(), img.size() * 4);
    }
    return 0;
}
int ref_write_ppm(const char* path, const float* img, int w, int h) { return writePPM(path, img, w, h) ? 0 : -1; }

// The view sequence of the training loop (train_main.cpp:613-623): its
// generator and distribution, seeded as there, drawn n times.
void ref_view_sequence(int n_views, int n, int* out) {
    std::mt19937 rng(42);
    std::uniform_int_distribution<int> view_dist(0, n_views - 1);
    for (int k = 0; k < n; k++) out[k] = (n_views > 1) ? view_dist(rng) : 0;
}

}  // extern "C"
