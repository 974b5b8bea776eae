// k_eval3d.hip -- the eval_3D forward: AAA-Gaussians 3D evaluation with the
// StopThePop per-pixel k-buffer (16 entries), and its C ABI.
//
// Stages and the reference code each follows, in the same floating-point
// operation order (compiled with -ffp-contract=off, IEEE division and sqrt):
//   preprocess3d ... cpu/preprocessor_cpu.cpp:64-187, math_utils.cpp:158-583
//   tile rays ...... math_utils.cpp:632-651 at each tile's centre
//   duplicate3d .... cpu/tile_binner_cpu.cpp:14-20,59-99 (per-tile depth key)
//   raster3d ....... cpu/rasterizer_cpu.cpp:16-21,61-135,185-193
// The pairs are sorted with the same stable radix sort as the 2D path
// (gs_sort_pairs); only the low 32 key bits differ (depth along the ray
// through the tile's centre).
//
// Layout in HBM (the reference's PreprocessOutput arrays): gauss2screen
// [N][16] row-major, cov3D_inv [N][6] (upper triangle), mean_offset [N][3].
// The rasterizer stages rows 0, 1 and 3 of gauss2screen, the inverse
// covariance, the mean offset, opacity and view depth of 256 Gaussians at a
// time in LDS (112 B each); every lane of the tile's workgroup evaluates its
// pixel against each staged entry and keeps its 16 nearest entries (by depth
// along its own ray) sorted in a ring in LDS, blending the front one whenever
// a 17th arrives.
#include <cmath>
#include <cstring>

#include "gs_device.h"
#include "gs_internal.h"
#include "gs_pre_device.h"
#include "gs_wave.h"

namespace gs3 {
using namespace gsd;

constexpr float kAlpha = 1.0f / 255.0f;
constexpr float kPi = 3.14159265358979323846f;
constexpr int KBUF = 16;

// Camera constants for the 3D kernels.  w2s = viewport * viewproj and ivp =
// inverse(viewproj) are computed once on the host by the same arithmetic the
// reference repeats per Gaussian / per frame (math_utils.cpp:560-562, 653-696).
struct Cam3 {
    float view[16], vp[16], w2s[16], ivp[16];
    float campos[3];
    float fx, fy, focal;
    float smod;
    int W, H, gx, gy;
    int deg, training, have_ivp;
};

// ---- math_utils.cpp:166-174 (out = A * B, column-major, sums from +0) ----
__host__ __device__ __forceinline__ void mat4mul(const float* A, const float* B, float* out) {
#pragma unroll
    for (int col = 0; col < 4; col++)
#pragma unroll
        for (int row = 0; row < 4; row++) {
            float sum = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; k++) sum += A[k * 4 + row] * B[col * 4 + k];
            out[col * 4 + row] = sum;
        }
}

// ---- math_utils.cpp:186-206: squared Mahalanobis distance of the ray
// where the two planes meet (the point itself is not needed here) ----
__device__ __forceinline__ float ray_contrib(const float* pa, const float* pb, float* mp) {
    const float dx = pa[1] * pb[2] - pa[2] * pb[1];
    const float dy = pa[2] * pb[0] - pa[0] * pb[2];
    const float dz = pa[0] * pb[1] - pa[1] * pb[0];
    const float mx = pa[3] * pb[0] - pa[0] * pb[3];
    const float my = pa[3] * pb[1] - pa[1] * pb[3];
    const float mz = pa[3] * pb[2] - pa[2] * pb[3];
    const float dd = dx * dx + dy * dy + dz * dz;
    const float q0 = mx / dd, q1 = my / dd, q2 = mz / dd;
    if (mp) {
        mp[0] = dy * q2 - dz * q1;
        mp[1] = dz * q0 - dx * q2;
        mp[2] = dx * q1 - dy * q0;
        mp[3] = 1.0f;
    }
    return mx * q0 + my * q1 + mz * q2;
}

// transpose(g2s) * v for row-major g2s (math_utils.cpp:214-219, 241-246)
__device__ __forceinline__ void to_screen(const float* g2s, const float* v, float* o) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++) s += g2s[j * 4 + i] * v[j];
        o[i] = s;
    }
}

// ---- math_utils.cpp:210-221 ----
__device__ __forceinline__ float plane_contrib(const float* pl, const float* g2s, float* ps) {
    const float norm = pl[3] / (pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
    const float g[4] = {-pl[0] * norm, -pl[1] * norm, -pl[2] * norm, -pl[3] * norm};
    to_screen(g2s, g, ps);
    return -g[3];
}

// ---- math_utils.cpp:225-235 ----
__device__ __forceinline__ bool in_range(const float* ps, const float* from, const float* ext, int dim, float* d) {
    *d = ps[dim] - from[dim] * ps[3];
    return *d > 0.0f && *d < ext[dim] * ps[3];
}
__device__ __forceinline__ bool in_range(const float* ps, const float* from, const float* ext, int dim) {
    float d;
    return in_range(ps, from, ext, dim, &d);
}

__device__ __forceinline__ float ray_screen(const float* pa, const float* pb, const float* g2s, float* ps) {
    float g[4];
    const float c = ray_contrib(pa, pb, g);
    to_screen(g2s, g, ps);
    return c;
}

// ---- math_utils.cpp:253-338: the least squared distance (halved) of the
// Gaussian inside the screen frustum [x0, x1] x [y0, y1] x [-1, 1] ----
__device__ float frustum_contrib(float x0, float y0, float x1, float y1, const float* g2s) {
    float best = 1e30f;
    const float from[3] = {x0, y0, -1.0f};
    const float ext[3] = {x1 - x0, y1 - y0, 1.0f - -1.0f};
    const float mean[4] = {g2s[3], g2s[7], g2s[11], g2s[15]};
    float d[3];
    const bool bx = in_range(mean, from, ext, 0, &d[0]);
    const bool by = in_range(mean, from, ext, 1, &d[1]);
    const bool bz = in_range(mean, from, ext, 2, &d[2]);
    if (bx && by && bz) {
        best = 0.0f;
    } else {
        const float sx = (d[0] - (ext[0] * 0.5f * mean[3])) >= 0 ? ext[0] * 0.5f : -ext[0] * 0.5f;
        const float sy = (d[1] - (ext[1] * 0.5f * mean[3])) >= 0 ? ext[1] * 0.5f : -ext[1] * 0.5f;
        const float cx = from[0] + ext[0] * 0.5f + sx;
        const float cy = from[1] + ext[1] * 0.5f + sy;
        float px[4], py[4], ps[4], c;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            px[i] = g2s[i] - g2s[12 + i] * cx;
            py[i] = g2s[4 + i] - g2s[12 + i] * cy;
        }
        c = plane_contrib(px, g2s, ps);
        if (c < best && in_range(ps, from, ext, 1) && in_range(ps, from, ext, 2)) best = c;
        c = plane_contrib(py, g2s, ps);
        if (c < best && in_range(ps, from, ext, 0) && in_range(ps, from, ext, 2)) best = c;
        c = ray_screen(px, py, g2s, ps);
        if (c < best && in_range(ps, from, ext, 2)) best = c;
        float oy[4], ox[4];
        const float ocy = from[1] + ext[1] * 0.5f - sy;
#pragma unroll
        for (int i = 0; i < 4; i++) oy[i] = g2s[4 + i] - g2s[12 + i] * ocy;
        c = ray_screen(px, oy, g2s, ps);
        if (c < best && in_range(ps, from, ext, 2)) best = c;
        const float ocx = from[0] + ext[0] * 0.5f - sx;
#pragma unroll
        for (int i = 0; i < 4; i++) ox[i] = g2s[i] - g2s[12 + i] * ocx;
        c = ray_screen(ox, py, g2s, ps);
        if (c < best && in_range(ps, from, ext, 2)) best = c;
    }
    return 0.5f * best;
}

// ---- math_utils.cpp:342-389: screen-space bounds (false: not usable) ----
__device__ bool aabb_screen(const float* g2s, float cutoff, float* mean2d, float* ext2d) {
    const float t[4] = {cutoff, cutoff, cutoff, -1.0f};
    const float *r0 = g2s, *r1 = g2s + 4, *r2 = g2s + 8, *r3 = g2s + 12;
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) s += t[k] * r3[k] * r3[k];
    if (s >= 0.0f) return false;
    float f[4];
#pragma unroll
    for (int k = 0; k < 4; k++) f[k] = t[k] / s;
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        p0 += f[k] * r0[k] * r3[k];
        p1 += f[k] * r1[k] * r3[k];
        p2 += f[k] * r2[k] * r3[k];
    }
    float h0 = p0 * p0, h1 = p1 * p1, h2 = p2 * p2;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        h0 -= f[k] * r0[k] * r0[k];
        h1 -= f[k] * r1[k] * r1[k];
        h2 -= f[k] * r2[k] * r2[k];
    }
    const float ez = sqrtf(smax(0.0f, h2));
    const float lo = p2 - ez, hi = p2 + ez;
    if (lo < -1.0f || hi > 1.0f) return false;
    mean2d[0] = p0;
    mean2d[1] = p1;
    ext2d[0] = sqrtf(smax(0.0f, h0));
    ext2d[1] = sqrtf(smax(0.0f, h1));
    return true;
}

// ---- math_utils.cpp:391-396 ----
__device__ __forceinline__ float norm_angle(float th) {
    th = fmodf(th, 2.0f * kPi);
    if (th > kPi) th -= 2.0f * kPi;
    if (th <= -kPi) th += 2.0f * kPi;
    return th;
}

// ---- math_utils.cpp:413-458: tangents of the bounding angles about one axis ----
__device__ void view_theta(const float* g2v, const float* vd, int axis, float cutoff, float* tlo, float* thi) {
    const float t[4] = {cutoff, cutoff, cutoff, -1.0f};
    const float mu = atan2f(axis == 0 ? vd[0] : vd[1], vd[2]);
    const float* Ta = g2v + axis * 4;
    const float* T2 = g2v + 8;
    float sa = 0.0f, sz = 0.0f, mid = 0.0f;
    for (int i = 0; i < 4; i++) {
        sa += t[i] * Ta[i] * Ta[i];
        sz += t[i] * T2[i] * T2[i];
        mid += t[i] * Ta[i] * T2[i];
    }
    const float in = mid * mid - sz * sa;
    float lo = -(kPi / 2.0f - 1e-5f);
    float hi = kPi / 2.0f - 1e-5f;
    if (in > 0.0f) {
        const float sq = sqrtf(in);
        float a0 = atan2f(-(mid + sq), -sz);
        float a1 = atan2f(-(mid - sq), -sz);
        while (a0 > mu) a0 -= kPi;
        while (a0 < (mu - kPi)) a0 += kPi;
        while (a1 < mu) a1 += kPi;
        while (a1 > (mu + kPi)) a1 -= kPi;
        const float n0 = norm_angle(a0), n1 = norm_angle(a1);
        if (mu < 0.0f && fabsf(n0) < fabsf(n1)) {
            a0 += 2.0f * kPi;
            a1 += 2.0f * kPi;
        } else if (mu > 0.0f && fabsf(n1) < fabsf(n0)) {
            a0 -= 2.0f * kPi;
            a1 -= 2.0f * kPi;
        }
        lo = smax(lo, a0);
        hi = smin(hi, a1);
    }
    *tlo = tanf(lo);
    *thi = tanf(hi);
}

// ---- math_utils.cpp:398-472: view-space bounds (the near-plane fallback) ----
__device__ void aabb_view(const float* g2v, const float* mv, float fx, float fy, float W, float H, float cutoff,
                          float* mean2d, float* ext2d) {
    const float vlen = sqrtf(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
    const float vd[3] = {mv[0] / vlen, mv[1] / vlen, mv[2] / vlen};
    float bxl, bxh, byl, byh;
    view_theta(g2v, vd, 0, cutoff, &bxl, &bxh);
    view_theta(g2v, vd, 1, cutoff, &byl, &byh);
    const float xl = W / 2.0f + fx * bxl, xh = W / 2.0f + fx * bxh;
    const float yl = H / 2.0f + fy * byl, yh = H / 2.0f + fy * byh;
    mean2d[0] = (xh + xl) / 2.0f;
    mean2d[1] = (yh + yl) / 2.0f;
    ext2d[0] = (xh - xl) / 2.0f;
    ext2d[1] = (yh - yl) / 2.0f;
}

using gsd::depth_along;

// ---- math_utils.cpp:632-651: unit world-space ray through pixel (px, py) ----
__device__ __forceinline__ void pixel_dir(float px, float py, const Cam3& c, float* d) {
    const float nx = (2.0f * px + 1.0f) / (float)c.W - 1.0f;
    const float ny = (2.0f * py + 1.0f) / (float)c.H - 1.0f;
    const float p[4] = {nx, ny, 0.0f, 1.0f};
    float w[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        w[i] = c.ivp[i] * p[0] + c.ivp[4 + i] * p[1] + c.ivp[8 + i] * p[2] + c.ivp[12 + i] * p[3];
    const float wi = 1.0f / (w[3] + 1e-10f);
    d[0] = w[0] * wi - c.campos[0];
    d[1] = w[1] * wi - c.campos[1];
    d[2] = w[2] * wi - c.campos[2];
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len > 1e-10f) {
        d[0] /= len;
        d[1] /= len;
        d[2] /= len;
    }
}

// float -> int as the reference's host build converts it (x86 cvttss2si:
// out-of-range and NaN give INT_MIN); the GPU conversion saturates instead.
__device__ __forceinline__ int to_int_x86(float v) {
    return (v < 2147483648.0f && v >= -2147483648.0f) ? (int)v : (int)0x80000000u;
}



// ------------------------------------------------------------ preprocess ----
// One thread per Gaussian: cpu/preprocessor_cpu.cpp:64-187.
template <int DEG>
__global__ __launch_bounds__(256) void preprocess3d_kernel(gs_gaussians G, const float* __restrict__ filter3d,
                                                           Cam3 c, gs_preprocessed out, gs_preprocessed_3d o3) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= G.count) return;
    Params P;
    load_params(G, i, G.raw != 0, P);
    float pv[3];
    xform43(P.p, c.view, pv);
    bool vis = !(pv[2] < 0.2f);
    float opacity = P.op, opt = 0.0f, cutoff = 0.0f;
    float g2s[16], g2v[16], sdl[3];
    float R[3][3];
    const float filt = filter3d ? filter3d[i] : 0.0f;
    if (vis) {
        // computeGauss2Screen (math_utils.cpp:474-583), kernel size 0.3
        rot_matrix(P.q[0], P.q[1], P.q[2], P.q[3], R);
        float vr[3] = {P.p[0] - c.campos[0], P.p[1] - c.campos[1], P.p[2] - c.campos[2]};
        const float vl = sqrtf(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
        vr[0] /= vl; vr[1] /= vl; vr[2] /= vl;
        float r2[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float v = R[k][0] * vr[0] + R[k][1] * vr[1] + R[k][2] * vr[2];
            r2[k] = v * v;
        }
        // mean in view space = pv (same transform)
        const float t = pv[2] / c.focal;
        float mip = (t * t) * 0.3f;
        mip = smax(filt * filt, mip);
#pragma unroll
        for (int k = 0; k < 3; k++) sdl[k] = sqrtf(P.s[k] * P.s[k] + mip);
        const float s0 = P.s[0] * P.s[0], s1 = P.s[1] * P.s[1], s2 = P.s[2] * P.s[2];
        const float d0 = sdl[0] * sdl[0], d1 = sdl[1] * sdl[1], d2 = sdl[2] * sdl[2];
        const float dm = r2[0] * (s1 * s2) + r2[1] * (s2 * s0) + r2[2] * (s0 * s1);
        const float dmd = r2[0] * (d1 * d2) + r2[1] * (d2 * d0) + r2[2] * (d0 * d1);
        const float dil = sqrtf(dm / dmd);
        const float sm = sqrtf(c.smod);
        const float sd[3] = {sdl[0] * sm, sdl[1] * sm, sdl[2] * sm};
        float g2w[16];
#pragma unroll
        for (int col = 0; col < 3; col++) {
#pragma unroll
            for (int k = 0; k < 3; k++) g2w[col * 4 + k] = sd[col] * R[col][k];
            g2w[col * 4 + 3] = 0.0f;
        }
        g2w[12] = P.p[0]; g2w[13] = P.p[1]; g2w[14] = P.p[2]; g2w[15] = 1.0f;
        float cm[16];
        mat4mul(c.w2s, g2w, cm);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) g2s[r * 4 + q] = cm[q * 4 + r];
        mat4mul(c.view, g2w, cm);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) g2v[r * 4 + q] = cm[q * 4 + r];
        opacity *= dil;
        vis = !(opacity < kAlpha);
        if (vis) {
            opt = log_ref(opacity / kAlpha);  // std::log(float), preprocessor_cpu.cpp:86
            cutoff = smin(11.11f, 2.0f * opt);
        }
    }
    float ci[6] = {0, 0, 0, 0, 0, 0}, mo[3] = {0, 0, 0};
    float m2[2] = {0.0f, 0.0f}, rgb[3] = {0.0f, 0.0f, 0.0f};
    int radius = 0, tiles = 0;
    if (vis) {
        const float sm = sqrtf(c.smod);
        const float sd[3] = {sdl[0] * sm, sdl[1] * sm, sdl[2] * sm};
        // computeCov3DInv (math_utils.cpp:590-613)
        float inv[3];
#pragma unroll
        for (int j = 0; j < 3; j++) inv[j] = 1.0f / (sd[j] * sd[j]);
        int idx = 0;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int q = r; q < 3; q++) {
                float v = 0.0f;
#pragma unroll
                for (int k = 0; k < 3; k++) v += R[k][r] * inv[k] * R[k][q];
                ci[idx++] = v;
            }
        mo[0] = P.p[0] - c.campos[0]; mo[1] = P.p[1] - c.campos[1]; mo[2] = P.p[2] - c.campos[2];
        const float df[3] = {c.campos[0] - P.p[0], c.campos[1] - P.p[1], c.campos[2] - P.p[2]};
        float dist2 = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float cg = (R[j][0] * df[0] + R[j][1] * df[1] + R[j][2] * df[2]) / sd[j];
            dist2 = j == 0 ? cg * cg : dist2 + cg * cg;
        }
        vis = !(dist2 < cutoff);
        if (vis) vis = !(frustum_contrib(0.0f, 0.0f, (float)(c.W - 1), (float)(c.H - 1), g2s) > opt);
        if (vis) {
            float ph[4];
            xform44(P.p, c.vp, ph);
            const float pwi = 1.0f / (ph[3] + 1e-7f);
            const float pxl = ndc2pix(ph[0] * pwi, c.W);
            const float pyl = ndc2pix(ph[1] * pwi, c.H);
            float mb[2], e2[2];
            if (!aabb_screen(g2s, cutoff, mb, e2))
                aabb_view(g2v, pv, c.fx, c.fy, (float)c.W, (float)c.H, cutoff, mb, e2);
            int r = to_int_x86(ceilf(smax(e2[0], e2[1])));
            r = smax(r, 1);
            int mn[2], mx[2];
            get_rect(pxl, pyl, r, c.gx, c.gy, mn, mx);
            tiles = (mx[1] - mn[1]) * (mx[0] - mn[0]);
            vis = (mx[0] - mn[0]) * (mx[1] - mn[1]) != 0;
            if (vis) {
                radius = r;
                m2[0] = pxl;
                m2[1] = pyl;
                float sh[NB * 3];
                const float* src = G.sh_coeffs + (size_t)i * G.max_coeffs * 3;
#pragma unroll
                for (int k = 0; k < NB * 3; k++) sh[k] = src[k];
                sh_color_from<DEG>(sh, sh_dir(P.p, c.campos), c.training != 0, rgb);
            } else {
                tiles = 0;
            }
        }
    }
    if (!vis) {
        opacity = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; k++) g2s[k] = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; k++) ci[k] = 0.0f;
        mo[0] = mo[1] = mo[2] = 0.0f;
    }
    reinterpret_cast<float2*>(out.means2D)[i] = make_float2(m2[0], m2[1]);
    out.depths[i] = vis ? pv[2] : 0.0f;
    out.conics[i * 3] = 0.0f; out.conics[i * 3 + 1] = 0.0f; out.conics[i * 3 + 2] = 0.0f;
    out.opacities[i] = opacity;
    out.rgb[i * 3] = rgb[0]; out.rgb[i * 3 + 1] = rgb[1]; out.rgb[i * 3 + 2] = rgb[2];
    out.radii[i] = radius;
    out.tiles_touched[i] = tiles;
    float4* g4 = reinterpret_cast<float4*>(o3.gauss2screen + (size_t)i * 16);
#pragma unroll
    for (int k = 0; k < 4; k++) g4[k] = make_float4(g2s[4 * k], g2s[4 * k + 1], g2s[4 * k + 2], g2s[4 * k + 3]);
#pragma unroll
    for (int k = 0; k < 6; k++) o3.cov3D_inv[(size_t)i * 6 + k] = ci[k];
#pragma unroll
    for (int k = 0; k < 3; k++) o3.mean_offset[(size_t)i * 3 + k] = mo[k];
}

// --------------------------------------------------------------- binning ----
// The unit ray through each tile's centre ((x + 0.5) * 16, (y + 0.5) * 16),
// the direction tile_binner_cpu.cpp:82-85 recomputes for every pair.
__global__ __launch_bounds__(256) void tile_rays_kernel(Cam3 c, float4* __restrict__ rays) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= c.gx * c.gy) return;
    const int x = t % c.gx, y = t / c.gx;
    float d[3];
    pixel_dir((x + 0.5f) * 16.0f, (y + 0.5f) * 16.0f, c, d);
    rays[t] = make_float4(d[0], d[1], d[2], 0.0f);
}

// Pairs of each Gaussian in row-major tile order at its exclusive offset
// (inclusive scan - own count), as tile_binner_cpu.cpp:64-98.  Each lane
// reads one Gaussian; the wave then walks its 64 Gaussians one at a time,
// all lanes emitting consecutive pairs of the current one, so the writes are
// coalesced and a Gaussian covering thousands of tiles does not serialise
// on one lane.
__global__ __launch_bounds__(256) void duplicate3d_kernel(int n, const float2* __restrict__ means2D,
                                                          const float* __restrict__ depths,
                                                          const int* __restrict__ radii,
                                                          const float* __restrict__ cinv,
                                                          const float* __restrict__ moff,
                                                          const uint32_t* __restrict__ offsets_incl,
                                                          const float4* __restrict__ rays, int gx, int gy,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ values,
                                                          uint32_t capacity, uint32_t* __restrict__ overflow) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int mn[2] = {0, 0}, mx[2] = {0, 0};
    uint32_t off = 0, db = 0;
    float ci[6] = {0, 0, 0, 0, 0, 0}, mo[3] = {0, 0, 0};
    const int r = i < n ? radii[i] : 0;
    if (r > 0) {
        const float2 m = means2D[i];
        get_rect(m.x, m.y, r, gx, gy, mn, mx);
        off = offsets_incl[i] - (uint32_t)((mx[1] - mn[1]) * (mx[0] - mn[0]));
        db = __float_as_uint(depths[i]);
#pragma unroll
        for (int k = 0; k < 6; k++) ci[k] = cinv[(size_t)i * 6 + k];
#pragma unroll
        for (int k = 0; k < 3; k++) mo[k] = moff[(size_t)i * 3 + k];
    }
    const int w = mx[0] - mn[0];
    const int cnt = w * (mx[1] - mn[1]);
    uint64_t busy = __ballot(cnt > 0);
    while (busy) {
        const int src = __ffsll((long long)busy) - 1;
        busy &= busy - 1;
        const int scnt = __shfl(cnt, src), sw = __shfl(w, src);
        const int sx0 = __shfl(mn[0], src), sy0 = __shfl(mn[1], src);
        const uint32_t soff = __shfl(off, src), sdb = __shfl(db, src);
        float sci[6], smo[3];
#pragma unroll
        for (int k = 0; k < 6; k++) sci[k] = __shfl(ci[k], src);
#pragma unroll
        for (int k = 0; k < 3; k++) smo[k] = __shfl(mo[k], src);
        const uint32_t gid = (uint32_t)(blockIdx.x * 256 + (threadIdx.x & ~63) + src);
        for (int q = lane; q < scnt; q += 64) {
            const int y = sy0 + q / sw, x = sx0 + q % sw;
            const uint32_t tile = (uint32_t)(y * gx + x);
            uint32_t tb = sdb;
            if (rays) {
                const float4 rv = rays[tile];
                const float vd[3] = {rv.x, rv.y, rv.z};
                const float ptd = depth_along(sci, smo, vd);
                if (ptd > 0.0f) tb = __float_as_uint(ptd);
            }
            const uint32_t o = soff + (uint32_t)q;
            if (o < capacity) {
                keys[o] = ((uint64_t)tile << 32) | tb;
                values[o] = gid;
            } else if (overflow) {
                *overflow = 1u;
            }
        }
    }
}

// ---------------------------------------------------------------- raster ----
struct Entry3 {
    float4 r0, r1, r3;  // gauss2screen rows 0, 1, 3
    float ci[6];
    float mo[3];
    float op, depth;
    uint32_t id;
    float cut;  // 1e-3 - ln(1 / (255 op)): half the squared distance beyond which alpha < 1/255
    float pad[3];
};
static_assert(sizeof(Entry3) == 112, "one LDS record per Gaussian");

// Entries staged per round.  The k-buffers take 48 KB of LDS per workgroup;
// 32 staged entries (3.5 KB) keep a workgroup at 52 KB, so three fit a CU
// (256 entries, 28 KB, allowed two): the kernel is latency-bound, and the third
// workgroup per CU hides more than the extra staging rounds cost
// (3.75 -> 2.82 ms at 1M Gaussians, 1080p).
constexpr int R3_BATCH = 32;

// One workgroup per 16x16 tile, one pixel per lane (rasterizer_cpu.cpp:61-135).
// Each lane's k-buffer is a 16-slot ring in LDS (slot-major, so the lanes of
// a wave touch consecutive words): the front entry leaves by advancing the
// head, and an entry is inserted from the back, moving only the deeper
// entries.  The pairs arrive in order of depth at the tile's centre, so the
// move count is usually zero or one.
__global__ __launch_bounds__(256) void raster3d_kernel(Cam3 c, const uint32_t* __restrict__ ranges,
                                                       const uint32_t* __restrict__ values,
                                                       const float* __restrict__ g2s, const float* __restrict__ cinv,
                                                       const float* __restrict__ moff, const float* __restrict__ op2d,
                                                       const float* __restrict__ depths, const float* __restrict__ rgb,
                                                       float bg0, float bg1, float bg2, float* __restrict__ img,
                                                       float* __restrict__ T_final, int* __restrict__ n_contrib) {
    __shared__ Entry3 sm[R3_BATCH];
    __shared__ float s_kd[KBUF][256];
    __shared__ float s_ka[KBUF][256];
    __shared__ uint32_t s_ki[KBUF][256];
    __shared__ uint64_t s_e2[32];  // exp_ref's table (read before the first barrier below)
    const int tile = gsd::xcd_remap(blockIdx.x, gridDim.x);  // neighbouring tiles share an XCD's L2
    const int tx = tile % c.gx, ty = tile / c.gx;
    const int t = threadIdx.x;
    if (t < 32) s_e2[t] = kExp2TabDev[t];
    const int px = tx * 16 + (t & 15), py = ty * 16 + (t >> 4);
    const bool inside = px < c.W && py < c.H;
    const uint32_t r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
    float rd[3];
    pixel_dir((float)px + 0.5f, (float)py + 0.5f, c, rd);
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;

    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    int cnt = 0, kc = 0, head = 0;
    bool done = !inside;

    for (uint32_t base = r0; base < r1; base += R3_BATCH) {
        // the workgroup stops once every pixel is finished
        if (__syncthreads_count(done ? 0 : 1) == 0) break;
        const uint32_t j = base + t;
        if (t < R3_BATCH && j < r1) {  // threads 0 .. R3_BATCH-1 stage one entry each
            const uint32_t id = values[j];
            const float4* g4 = reinterpret_cast<const float4*>(g2s + (size_t)id * 16);
            Entry3& e = sm[t];
            e.r0 = g4[0];
            e.r1 = g4[1];
            e.r3 = g4[3];
            const float2* c2 = reinterpret_cast<const float2*>(cinv + (size_t)id * 6);
            const float2 a = c2[0], b = c2[1], d = c2[2];
            e.ci[0] = a.x; e.ci[1] = a.y; e.ci[2] = b.x; e.ci[3] = b.y; e.ci[4] = d.x; e.ci[5] = d.y;
            e.mo[0] = moff[(size_t)id * 3];
            e.mo[1] = moff[(size_t)id * 3 + 1];
            e.mo[2] = moff[(size_t)id * 3 + 2];
            const float o = op2d[id];
            e.op = o;
            e.depth = depths[id];
            e.id = id;
            e.cut = 1e-3f + __logf(255.0f * o);
        }
        __syncthreads();
        const int m = (int)((r1 - base) < (uint32_t)R3_BATCH ? (r1 - base) : (uint32_t)R3_BATCH);
        for (int q = 0; q < m && !done; q++) {
            const Entry3& e = sm[q];
            const float plx[4] = {e.r0.x - e.r3.x * fx, e.r0.y - e.r3.y * fx, e.r0.z - e.r3.z * fx,
                                  e.r0.w - e.r3.w * fx};
            const float ply[4] = {e.r1.x - e.r3.x * fy, e.r1.y - e.r3.y * fy, e.r1.z - e.r3.z * fy,
                                  e.r1.w - e.r3.w * fy};
            {
                // Far below the 1/255 cut the exact evaluation can be skipped:
                // |m|^2 rcp(dd) is within 1e-6 (relative) of the exact squared
                // distance, so 0.5 md (1 - 1e-5) > 1e-3 + ln(255 o) gives
                // o e^power < e^-0.001 / 255 for the exact power, which the
                // exact path also skips (NaN and inf fall through or agree).
                const float dx = plx[1] * ply[2] - plx[2] * ply[1];
                const float dy = plx[2] * ply[0] - plx[0] * ply[2];
                const float dz = plx[0] * ply[1] - plx[1] * ply[0];
                const float mx = plx[3] * ply[0] - plx[0] * ply[3];
                const float my = plx[3] * ply[1] - plx[1] * ply[3];
                const float mz = plx[3] * ply[2] - plx[2] * ply[3];
                const float mm = mx * mx + my * my + mz * mz;  // must not overflow where the exact path does not
                const float md = mm * __builtin_amdgcn_rcpf(dx * dx + dy * dy + dz * dz);
                if (mm < 3.0e38f && 0.5f * md * (1.0f - 1e-5f) > e.cut) continue;
            }
            const float power = -0.5f * ray_contrib(plx, ply, nullptr);
            if (power > 0.0f) continue;
            const float alpha = smin(0.99f, e.op * exp_ref_tab(power, s_e2));
            if (alpha < kAlpha) continue;
            float pd = e.depth;
            if (c.have_ivp) {
                const float ptd = depth_along(e.ci, e.mo, rd);
                if (ptd > 0.0f) pd = ptd;
            }
            if (kc >= KBUF) {
                // blend the front entry (the buffer is full)
                const float a = s_ka[head][t];
                const float tt = T * (1.0f - a);
                if (tt < 0.0001f) {
                    done = true;
                    continue;
                }
                const uint32_t gid = s_ki[head][t];
                C0 += rgb[gid * 3] * a * T;
                C1 += rgb[gid * 3 + 1] * a * T;
                C2 += rgb[gid * 3 + 2] * a * T;
                T = tt;
                cnt++;
                head = (head + 1) & (KBUF - 1);
                kc = KBUF - 1;
            }
            // insertion after the last entry not deeper than pd
            int at = kc;
            while (at > 0) {
                const int sl = (head + at - 1) & (KBUF - 1);
                const float dprev = s_kd[sl][t];
                if (!(dprev > pd)) break;
                const int dst = (head + at) & (KBUF - 1);
                s_kd[dst][t] = dprev;
                s_ka[dst][t] = s_ka[sl][t];
                s_ki[dst][t] = s_ki[sl][t];
                at--;
            }
            const int dst = (head + at) & (KBUF - 1);
            s_kd[dst][t] = pd;
            s_ka[dst][t] = alpha;
            s_ki[dst][t] = e.id;
            kc++;
        }
        __syncthreads();
    }
    if (!inside) return;
    // flush, front to back
    if (!done) {
        for (int k = 0; k < kc; k++) {
            const int sl = (head + k) & (KBUF - 1);
            const float a = s_ka[sl][t];
            const float tt = T * (1.0f - a);
            if (tt < 0.0001f) break;
            const uint32_t gid = s_ki[sl][t];
            C0 += rgb[gid * 3] * a * T;
            C1 += rgb[gid * 3 + 1] * a * T;
            C2 += rgb[gid * 3 + 2] * a * T;
            T = tt;
            cnt++;
        }
    }
    const int pix = py * c.W + px;
    if (T_final) T_final[pix] = T;
    if (n_contrib) n_contrib[pix] = cnt;
    img[pix * 3] = C0 + T * bg0;
    img[pix * 3 + 1] = C1 + T * bg1;
    img[pix * 3 + 2] = C2 + T * bg2;
}

// ------------------------------------------------------------------ host ----
// math_utils.cpp:653-696, host float arithmetic (identical to the reference's)
static bool invert4(const float* m, float* inv) {
    const float a00 = m[0], a10 = m[1], a20 = m[2], a30 = m[3];
    const float a01 = m[4], a11 = m[5], a21 = m[6], a31 = m[7];
    const float a02 = m[8], a12 = m[9], a22 = m[10], a32 = m[11];
    const float a03 = m[12], a13 = m[13], a23 = m[14], a33 = m[15];
    const float b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10, b02 = a00 * a13 - a03 * a10;
    const float b03 = a01 * a12 - a02 * a11, b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
    const float b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30, b08 = a20 * a33 - a23 * a30;
    const float b09 = a21 * a32 - a22 * a31, b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
    const float det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
    if (std::fabs(det) < 1e-12f) return false;
    const float id = 1.0f / det;
    inv[0] = (a11 * b11 - a12 * b10 + a13 * b09) * id;
    inv[1] = (-a10 * b11 + a12 * b08 - a13 * b07) * id;
    inv[2] = (a10 * b10 - a11 * b08 + a13 * b06) * id;
    inv[3] = (-a10 * b09 + a11 * b07 - a12 * b06) * id;
    inv[4] = (-a01 * b11 + a02 * b10 - a03 * b09) * id;
    inv[5] = (a00 * b11 - a02 * b08 + a03 * b07) * id;
    inv[6] = (-a00 * b10 + a01 * b08 - a03 * b06) * id;
    inv[7] = (a00 * b09 - a01 * b07 + a02 * b06) * id;
    inv[8] = (a31 * b05 - a32 * b04 + a33 * b03) * id;
    inv[9] = (-a30 * b05 + a32 * b02 - a33 * b01) * id;
    inv[10] = (a30 * b04 - a31 * b02 + a33 * b00) * id;
    inv[11] = (-a30 * b03 + a31 * b01 - a32 * b00) * id;
    inv[12] = (-a21 * b05 + a22 * b04 - a23 * b03) * id;
    inv[13] = (a20 * b05 - a22 * b02 + a23 * b01) * id;
    inv[14] = (-a20 * b04 + a21 * b02 - a23 * b00) * id;
    inv[15] = (a20 * b03 - a21 * b01 + a22 * b00) * id;
    return true;
}

static Cam3 make_cam3(const gs_camera* cam, const gs_render_cfg* cfg, int model_deg) {
    Cam3 c;
    std::memset(&c, 0, sizeof(c));
    std::memcpy(c.view, cam->view_matrix, 64);
    std::memcpy(c.vp, cam->viewproj_matrix, 64);
    std::memcpy(c.campos, cam->cam_pos, 12);
    c.W = cam->width;
    c.H = cam->height;
    c.gx = (c.W + 15) / 16;
    c.gy = (c.H + 15) / 16;
    // preprocessor_cpu.cpp:50-51, 70
    c.fx = (float)c.W / (2.0f * cam->tan_fovx);
    c.fy = (float)c.H / (2.0f * cam->tan_fovy);
    c.focal = (c.fx < c.fy) ? c.fy : c.fx;
    if (cfg) {
        c.smod = cfg->scale_modifier;
        c.training = cfg->training;
        c.deg = cfg->sh_degree < model_deg ? cfg->sh_degree : model_deg;
    }
    // math_utils.cpp:546-562: viewport (column-major) times viewproj
    float port[16];
    std::memset(port, 0, sizeof(port));
    port[0] = c.W / 2.0f;
    port[5] = c.H / 2.0f;
    port[10] = 1.0f;
    port[15] = 1.0f;
    port[12] = c.W / 2.0f - 0.5f;
    port[13] = c.H / 2.0f - 0.5f;
    mat4mul(port, c.vp, c.w2s);
    c.have_ivp = invert4(c.vp, c.ivp) ? 1 : 0;
    return c;
}

// ------------------------------------------------- direct 3D binning ----
// The pair order the reference's duplicate + stable sort produce is, per tile
// (in tile order), the Gaussians that touch it ordered by their per-tile depth
// key, equal keys in Gaussian order.  The direct-placement binning of the 2D
// path, fed the Gaussians in index order, writes every tile's list in Gaussian
// order; each tile's list is then sorted stably by its per-tile key in LDS
// (TS_CAP pairs at most; longer lists go through the general radix sort).

constexpr int TS_THREADS = 1024, TS_ITEMS = 16, TS_WAVES = TS_THREADS / 64;
constexpr int TS_CAP = TS_THREADS * TS_ITEMS;  // 16384 pairs per tile

__global__ __launch_bounds__(256) void iota_kernel(uint32_t n, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] = i;
}

// A pair key's inputs per Gaussian in one 64-byte line: cov3D_inv (6),
// mean_offset (3), depth, so that a pair's key costs three 16-byte loads of
// one cache line instead of ten loads from three arrays.
__global__ __launch_bounds__(256) void pack_keys3d_kernel(int n, const float* __restrict__ depths,
                                                          const float* __restrict__ cinv,
                                                          const float* __restrict__ moff, float4* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* c = cinv + (size_t)i * 6;
    const float* m = moff + (size_t)i * 3;
    out[(size_t)i * 4] = make_float4(c[0], c[1], c[2], c[3]);
    out[(size_t)i * 4 + 1] = make_float4(c[4], c[5], m[0], m[1]);
    out[(size_t)i * 4 + 2] = make_float4(m[2], depths[i], 0.0f, 0.0f);
}

// The low word of a pair's key, as duplicate3d_kernel computes it (ray: the
// tile's ray, used when have_ray).
__device__ __forceinline__ uint32_t tile_depth_key(uint32_t g, const float4* __restrict__ packed, bool have_ray,
                                                   float4 ray) {
    const float4 c = packed[(size_t)g * 4 + 2];
    uint32_t tb = __float_as_uint(c.y);
    if (have_ray) {
        const float4 a = packed[(size_t)g * 4], b = packed[(size_t)g * 4 + 1];
        const float ci[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
        const float mo[3] = {b.z, b.w, c.x};
        const float vd[3] = {ray.x, ray.y, ray.z};
        const float ptd = depth_along(ci, mo, vd);
        if (ptd > 0.0f) tb = __float_as_uint(ptd);
    }
    return tb;
}

// One workgroup per tile: its list (Gaussian order) sorted stably by the
// per-tile key, four 8-bit LSD passes through LDS; values rewritten in place,
// keys (tile << 32 | key) written when keys_out is set.  A list longer than
// TS_CAP is left as it is and its length recorded in bigcnt (0 otherwise).
__global__ __launch_bounds__(TS_THREADS) void tile_sort3d_kernel(const uint32_t* __restrict__ ranges,
                                                                 uint32_t* __restrict__ values,
                                                                 uint64_t* __restrict__ keys_out,
                                                                 const float4* __restrict__ packed,
                                                                 const float4* __restrict__ rays,
                                                                 uint32_t* __restrict__ bigcnt) {
    using namespace gsw;
    __shared__ uint32_t s_k[TS_CAP];
    __shared__ uint32_t s_v[TS_CAP];
    __shared__ uint32_t s_wh[TS_WAVES][256];
    __shared__ uint32_t s_dstart[256];
    __shared__ uint32_t s_scan[TS_WAVES];
    const uint32_t tile = (uint32_t)gsd::xcd_remap(blockIdx.x, gridDim.x);  // neighbouring tiles share an L2
    const uint32_t s = ranges[2 * tile], e = ranges[2 * tile + 1];
    const uint32_t cnt = e > s ? e - s : 0u;
    if (threadIdx.x == 0) bigcnt[tile] = cnt > (uint32_t)TS_CAP ? cnt : 0u;
    if (cnt == 0u || cnt > (uint32_t)TS_CAP) return;
    const int w = threadIdx.x >> 6, l = lane_id();
    const uint32_t wbase = (uint32_t)w * (TS_ITEMS * 64);
    const float4 ray = rays ? rays[tile] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t k_r[TS_ITEMS], v_r[TS_ITEMS], rank_r[TS_ITEMS];
#pragma unroll
    for (int r = 0; r < TS_ITEMS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        v_r[r] = i < cnt ? values[s + i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < TS_ITEMS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        k_r[r] = i < cnt ? tile_depth_key(v_r[r], packed, rays != nullptr, ray) : 0u;
    }
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 8 * pass;
        for (int i = threadIdx.x; i < TS_WAVES * 256; i += TS_THREADS) (&s_wh[0][0])[i] = 0u;
        __syncthreads();
        // per-wave stable ranks (rounds in list order, lanes in list order)
#pragma unroll
        for (int r = 0; r < TS_ITEMS; r++) {
            const uint32_t i = wbase + r * 64 + l;
            const bool valid = i < cnt;
            const uint64_t active = __ballot(valid);
            rank_r[r] = 0u;
            if (active != 0) {
                const uint32_t d = (k_r[r] >> shift) & 255u;
                const uint64_t peers = match_digit<8>(d, active);
                uint32_t before = 0u;
                if (valid) before = s_wh[w][d];
                __builtin_amdgcn_wave_barrier();
                if (valid && (peers & lanemask_lt()) == 0) s_wh[w][d] = before + (uint32_t)__popcll(peers);
                __builtin_amdgcn_wave_barrier();
                rank_r[r] = before + (uint32_t)__popcll(peers & lanemask_lt());
            }
        }
        __syncthreads();
        // digit totals, per-wave offsets inside each digit, digit starts
        uint32_t tot = 0u;
        if (threadIdx.x < 256) {
            const int d = threadIdx.x;
#pragma unroll
            for (int k = 0; k < TS_WAVES; k++) {
                const uint32_t c = s_wh[k][d];
                s_wh[k][d] = tot;
                tot += c;
            }
        }
        {
            uint32_t total;
            const uint32_t ex = block_exclusive_scan<uint32_t>(tot, s_scan, &total);
            if (threadIdx.x < 256) s_dstart[threadIdx.x] = ex;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < TS_ITEMS; r++) {
            const uint32_t i = wbase + r * 64 + l;
            if (i < cnt) {
                const uint32_t d = (k_r[r] >> shift) & 255u;
                const uint32_t p = s_dstart[d] + s_wh[w][d] + rank_r[r];
                s_k[p] = k_r[r];
                s_v[p] = v_r[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < TS_ITEMS; r++) {
            const uint32_t i = wbase + r * 64 + l;
            if (i < cnt) {
                k_r[r] = s_k[i];
                v_r[r] = s_v[i];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < TS_ITEMS; r++) {
        const uint32_t i = wbase + r * 64 + l;
        if (i < cnt) {
            values[s + i] = v_r[r];
            if (keys_out) keys_out[s + i] = ((uint64_t)tile << 32) | k_r[r];
        }
    }
}

// Lists longer than TS_CAP: their (tile << 32 | key, Gaussian) pairs copied to
// [boff[tile], boff[tile] + bigcnt[tile]) of bk / bv (tiles in order), one
// workgroup per tile; tiles with bigcnt 0 return at once.
__global__ __launch_bounds__(256) void big_gather3d_kernel(const uint32_t* __restrict__ ranges,
                                                           const uint32_t* __restrict__ bigcnt,
                                                           const uint32_t* __restrict__ boff,
                                                           const uint32_t* __restrict__ values,
                                                           const float4* __restrict__ packed,
                                                           const float4* __restrict__ rays, uint64_t* __restrict__ bk,
                                                           uint32_t* __restrict__ bv) {
    const uint32_t tile = blockIdx.x;
    const uint32_t cnt = bigcnt[tile];
    if (cnt == 0u) return;
    const uint32_t s = ranges[2 * tile], o = boff[tile];
    const float4 ray = rays ? rays[tile] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t i = threadIdx.x; i < cnt; i += 256u) {
        const uint32_t g = values[s + i];
        bk[o + i] = ((uint64_t)tile << 32) | tile_depth_key(g, packed, rays != nullptr, ray);
        bv[o + i] = g;
    }
}

// The sorted long lists back into their tiles' ranges.
__global__ __launch_bounds__(256) void big_scatter3d_kernel(const uint32_t* __restrict__ ranges,
                                                            const uint32_t* __restrict__ bigcnt,
                                                            const uint32_t* __restrict__ boff,
                                                            const uint64_t* __restrict__ bk,
                                                            const uint32_t* __restrict__ bv,
                                                            uint32_t* __restrict__ values,
                                                            uint64_t* __restrict__ keys_out) {
    const uint32_t tile = blockIdx.x;
    const uint32_t cnt = bigcnt[tile];
    if (cnt == 0u) return;
    const uint32_t s = ranges[2 * tile], o = boff[tile];
    for (uint32_t i = threadIdx.x; i < cnt; i += 256u) {
        values[s + i] = bv[o + i];
        if (keys_out) keys_out[s + i] = bk[o + i];
    }
}

}  // namespace gs3

// ------------------------------------------------------------------- ABI ----
#define G3_ARG(cond, ...)                                                  \
    do {                                                                   \
        if (!(cond)) return gs_set_error(GS_ERR_INVALID, __VA_ARGS__);     \
    } while (0)
#define G3_HIP(expr, what)                                      \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) return gs_hip_status(e_, what);   \
    } while (0)

static bool al(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

extern "C" int gs_preprocess_fwd_3d(gs_ctx* c, const gs_gaussians* g, const float* filter_3d, const gs_camera* cam,
                                    const gs_render_cfg* cfg, const gs_preprocessed* out,
                                    const gs_preprocessed_3d* out3d) {
    G3_ARG(c != nullptr, "null context");
    int st = check_cam_cfg(cam, cfg);
    if (st) return st;
    G3_ARG(g != nullptr && out != nullptr && out3d != nullptr, "null argument");
    G3_ARG(g->count >= 0, "negative Gaussian count");
    G3_ARG(g->max_coeffs == 1 || g->max_coeffs == 4 || g->max_coeffs == 9 || g->max_coeffs == 16,
           "max_coeffs must be 1, 4, 9 or 16 (got %d)", g->max_coeffs);
    G3_ARG(g->sh_degree >= 0 && (g->sh_degree + 1) * (g->sh_degree + 1) <= g->max_coeffs,
           "sh_degree %d does not fit max_coeffs %d", g->sh_degree, g->max_coeffs);
    if (g->count == 0) return GS_OK;
    G3_ARG(g->positions && g->scales && g->rotations && g->sh_coeffs && g->opacities, "null Gaussian array");
    G3_ARG(al(g->rotations, 16), "rotations must be 16-byte aligned");
    G3_ARG(out->means2D && out->depths && out->conics && out->opacities && out->rgb && out->radii &&
               out->tiles_touched,
           "null preprocess output");
    G3_ARG(al(out->means2D, 8), "means2D must be 8-byte aligned");
    G3_ARG(out3d->gauss2screen && out3d->cov3D_inv && out3d->mean_offset, "null eval_3D output");
    G3_ARG(al(out3d->gauss2screen, 16), "gauss2screen must be 16-byte aligned");
    const gs3::Cam3 k = gs3::make_cam3(cam, cfg, g->sh_degree);
    const dim3 grid((g->count + 255) / 256), block(256);
    switch (k.deg <= 0 ? 0 : k.deg) {
        case 0: gs3::preprocess3d_kernel<0><<<grid, block, 0, c->stream>>>(*g, filter_3d, k, *out, *out3d); break;
        case 1: gs3::preprocess3d_kernel<1><<<grid, block, 0, c->stream>>>(*g, filter_3d, k, *out, *out3d); break;
        case 2: gs3::preprocess3d_kernel<2><<<grid, block, 0, c->stream>>>(*g, filter_3d, k, *out, *out3d); break;
        default: gs3::preprocess3d_kernel<3><<<grid, block, 0, c->stream>>>(*g, filter_3d, k, *out, *out3d); break;
    }
    G3_HIP(hipGetLastError(), "preprocess3d kernel");
    return GS_OK;
}

extern "C" int gs_duplicate_keys_3d(gs_ctx* c, int n, const gs_preprocessed* pre, const gs_preprocessed_3d* pre3d,
                                    const uint32_t* offsets, const gs_camera* cam, int tile_w, int tile_h,
                                    uint64_t* keys, uint32_t* values, uint32_t capacity) {
    G3_ARG(c != nullptr, "null context");
    G3_ARG(cam != nullptr && cam->width > 0 && cam->height > 0, "bad camera");
    G3_ARG(tile_w == 16 && tile_h == 16, "tiles must be 16x16");
    if (n <= 0) return GS_OK;
    G3_ARG(pre && pre3d && offsets && keys && values, "null argument");
    G3_ARG(pre->means2D && pre->depths && pre->radii && pre3d->cov3D_inv && pre3d->mean_offset, "null array");
    G3_ARG(al(pre->means2D, 8), "means2D must be 8-byte aligned");
    void* sm;
    int st = gs_small_get(c, &sm);
    if (st) return st;
    uint32_t* ovf = reinterpret_cast<uint32_t*>(static_cast<char*>(sm) + GS_SMALL_OVERFLOW);
    const gs3::Cam3 k = gs3::make_cam3(cam, nullptr, 0);
    float4* rays = nullptr;
    if (k.have_ivp) {
        void* ws2;
        if ((st = gs_ws2_get(c, (size_t)k.gx * k.gy * 16, &ws2))) return st;
        rays = static_cast<float4*>(ws2);
        gs3::tile_rays_kernel<<<(k.gx * k.gy + 255) / 256, 256, 0, c->stream>>>(k, rays);
        G3_HIP(hipGetLastError(), "tile rays kernel");
    }
    gs3::duplicate3d_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(
        n, reinterpret_cast<const float2*>(pre->means2D), pre->depths, pre->radii, pre3d->cov3D_inv,
        pre3d->mean_offset, offsets, rays, k.gx, k.gy, keys, values, capacity, ovf);
    G3_HIP(hipGetLastError(), "duplicate3d kernel");
    return GS_OK;
}

extern "C" int gs_bin_tiles_3d(gs_ctx* c, int n, const gs_preprocessed* pre, const gs_preprocessed_3d* pre3d,
                               const gs_camera* cam, uint32_t* values, uint64_t* keys, uint32_t capacity,
                               uint32_t* ranges, long long* total) {
    G3_ARG(c != nullptr, "null context");
    G3_ARG(cam != nullptr && cam->width > 0 && cam->height > 0, "bad camera");
    G3_ARG(n >= 0, "negative count");
    G3_ARG(ranges != nullptr, "null tile ranges");
    const gs3::Cam3 k = gs3::make_cam3(cam, nullptr, 0);
    const int num_tiles = k.gx * k.gy;
    if (k.gx > GS_BIN_MAX_TILES_XY || k.gy > GS_BIN_MAX_TILES_XY)
        return gs_set_error(GS_ERR_UNSUPPORTED,
                            "gs_bin_tiles_3d: %d x %d tiles exceed %d per side (use gs_duplicate_keys_3d + gs_sort_pairs)",
                            k.gx, k.gy, GS_BIN_MAX_TILES_XY);
    if (total) *total = 0;
    if (n == 0) {
        G3_HIP(launch_zero(ranges, (size_t)num_tiles * 8, c->stream), "zero kernel");
        return GS_OK;
    }
    G3_ARG(pre && pre3d, "null argument");
    G3_ARG(pre->means2D && pre->depths && pre->radii && pre3d->cov3D_inv && pre3d->mean_offset, "null array");
    G3_ARG(al(pre->means2D, 8), "means2D must be 8-byte aligned");
    G3_ARG(values == nullptr || al(values, 16), "values must be 16-byte aligned");
    void* sm;
    int st = gs_small_get(c, &sm);
    if (st) return st;
    uint32_t* ovf = reinterpret_cast<uint32_t*>(static_cast<char*>(sm) + GS_SMALL_OVERFLOW);
    unsigned long long* total64 = reinterpret_cast<unsigned long long*>(static_cast<char*>(sm) + GS_SMALL_TOTAL64);
    auto alg = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // Gaussian-sized scratch: index order, rectangles, the (tile, group) table
    // and its scan, the tiles' rays, the long lists' counts and offsets
    const size_t un = (size_t)n;
    const int G = c->bin_chunk;
    const int C = (n + G - 1) / G;
    const size_t L = (size_t)C * num_tiles + 1;
    G3_ARG(L < 0xFFFFFFFFull, "binning table too large");
    const size_t np = scan_partials_needed((uint32_t)std::max<size_t>(L, (size_t)num_tiles + 1)) + 1;
    void* ws2;
    if ((st = gs_ws2_get(c, alg(un * 4) + alg(un * 8) + 2 * alg(L * 4) + alg(np * 8) + alg((size_t)num_tiles * 16) +
                                2 * alg(((size_t)num_tiles + 1) * 4) + alg(un * 64),
                         &ws2)))
        return st;
    char* w = static_cast<char*>(ws2);
    uint32_t* order = reinterpret_cast<uint32_t*>(w); w += alg(un * 4);
    uint2* rects = reinterpret_cast<uint2*>(w); w += alg(un * 8);
    uint32_t* table = reinterpret_cast<uint32_t*>(w); w += alg(L * 4);
    uint32_t* scanned = reinterpret_cast<uint32_t*>(w); w += alg(L * 4);
    unsigned long long* partial = reinterpret_cast<unsigned long long*>(w); w += alg(np * 8);
    float4* rays = reinterpret_cast<float4*>(w); w += alg((size_t)num_tiles * 16);
    uint32_t* bigcnt = reinterpret_cast<uint32_t*>(w); w += alg(((size_t)num_tiles + 1) * 4);
    uint32_t* boff = reinterpret_cast<uint32_t*>(w); w += alg(((size_t)num_tiles + 1) * 4);
    float4* packed = reinterpret_cast<float4*>(w);
    if (k.have_ivp) {
        gs3::tile_rays_kernel<<<(num_tiles + 255) / 256, 256, 0, c->stream>>>(k, rays);
        G3_HIP(hipGetLastError(), "tile rays kernel");
    }
    gs3::iota_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>((uint32_t)n, order);
    G3_HIP(hipGetLastError(), "iota kernel");
    gsb::BinArgs a{};
    a.V = 1;
    a.N = n;
    a.C = C;
    a.G = G;
    a.S = 0;
    a.G1 = G;
    a.sidx = order;
    a.sidx_stride = 1;
    a.means2D = reinterpret_cast<const float2*>(pre->means2D);
    a.radii = pre->radii;
    a.views = nullptr;
    a.view0 = make_int4(k.gx, k.gy, 0, 0);
    a.vblock0 = nullptr;
    a.table = table;
    a.table_len = L;
    a.rects = rects;
    a.out = values;
    a.capacity = capacity;
    a.overflow = ovf;
    a.diff_len = (k.gx + 1) * (k.gy + 1);
    const int blocks = bin_blocks(C, k.gx, k.gy, c->bin_tpl);
    G3_HIP(launch_bin_count(a, blocks, c->bin_tpl, c->stream), "binning count kernel");
    G3_HIP(launch_scan_u32_exclusive(table, (uint32_t)L, scanned, partial, c->stream, nullptr, total64),
           "binning scan");
    a.table = scanned;
    unsigned long long K = 0;
    G3_HIP(hipMemcpyAsync(&K, total64, 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(total)");
    G3_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (total) *total = (long long)K;
    if (K > capacity)
        return gs_set_error(GS_ERR_CAPACITY, "gs_bin_tiles_3d: %llu pairs exceed capacity %u", K, capacity);
    G3_HIP(launch_bin_ranges(a, scanned, (uint32_t)num_tiles, ranges, c->stream), "binning ranges kernel");
    if (K == 0) return GS_OK;
    G3_ARG(values != nullptr, "null values buffer");
    G3_HIP(launch_bin_place(a, blocks, c->bin_tpl, c->stream), "binning place kernel");
    const float4* rays_arg = k.have_ivp ? rays : nullptr;
    gs3::pack_keys3d_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(n, pre->depths, pre3d->cov3D_inv,
                                                                     pre3d->mean_offset, packed);
    G3_HIP(hipGetLastError(), "key pack kernel");
    gs3::tile_sort3d_kernel<<<num_tiles, gs3::TS_THREADS, 0, c->stream>>>(ranges, values, keys, packed, rays_arg,
                                                                          bigcnt);
    G3_HIP(hipGetLastError(), "tile sort kernel");
    // lists longer than one workgroup's LDS: the general radix sort over all of them
    uint32_t B = 0;
    uint32_t* B_dev = reinterpret_cast<uint32_t*>(static_cast<char*>(sm) + GS_SMALL_TOTAL64);
    G3_HIP(launch_scan_u32_exclusive(bigcnt, (uint32_t)num_tiles, boff, partial, c->stream, B_dev, nullptr),
           "long-list scan");
    G3_HIP(hipMemcpyAsync(&B, B_dev, 4, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(long lists)");
    G3_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (B == 0) return GS_OK;
    const size_t ub = (size_t)B;
    const int key_bits = gs_tile_key_bits(num_tiles);
    void* ws;
    if ((st = gs_ws_get(c, 2 * alg(ub * 8) + 2 * alg(ub * 4) + radix_workspace_bytes(B, c->radix_bits), &ws)))
        return st;
    char* q = static_cast<char*>(ws);
    uint64_t* bk = reinterpret_cast<uint64_t*>(q); q += alg(ub * 8);
    uint64_t* bk2 = reinterpret_cast<uint64_t*>(q); q += alg(ub * 8);
    uint32_t* bv = reinterpret_cast<uint32_t*>(q); q += alg(ub * 4);
    uint32_t* bv2 = reinterpret_cast<uint32_t*>(q); q += alg(ub * 4);
    gs3::big_gather3d_kernel<<<num_tiles, 256, 0, c->stream>>>(ranges, bigcnt, boff, values, packed, rays_arg, bk,
                                                               bv);
    G3_HIP(hipGetLastError(), "long-list gather kernel");
    G3_HIP(launch_radix_sort_bits(bk, bv, bk2, bv2, B, 0, key_bits, c->radix_bits, q, c->stream, nullptr),
           "long-list sort");
    gs3::big_scatter3d_kernel<<<num_tiles, 256, 0, c->stream>>>(ranges, bigcnt, boff, bk2, bv2, values, keys);
    G3_HIP(hipGetLastError(), "long-list scatter kernel");
    return GS_OK;
}

extern "C" int gs_raster_fwd_3d(gs_ctx* c, const gs_camera* cam, const gs_preprocessed* pre,
                                const gs_preprocessed_3d* pre3d, const uint32_t* values_sorted,
                                const uint32_t* ranges, const float bg[3], float* image, float* T_final,
                                int* n_contrib) {
    G3_ARG(c != nullptr, "null context");
    G3_ARG(cam != nullptr && cam->width > 0 && cam->height > 0, "bad camera");
    G3_ARG(pre && pre3d && ranges && bg && image, "null argument");
    G3_ARG(pre->depths && pre->opacities && pre->rgb && pre3d->gauss2screen && pre3d->cov3D_inv &&
               pre3d->mean_offset,
           "null array");
    G3_ARG(al(pre3d->gauss2screen, 16), "gauss2screen must be 16-byte aligned");
    G3_ARG(al(pre3d->cov3D_inv, 8), "cov3D_inv must be 8-byte aligned");
    // rasterizer_cpu.cpp:16-21 reads the inverse uninitialised when the
    // view-projection matrix is singular; there the view depth orders the
    // k-buffer, as the binner's keys do (tile_binner_cpu.cpp:80).
    const gs3::Cam3 k = gs3::make_cam3(cam, nullptr, 0);
    gs3::raster3d_kernel<<<k.gx * k.gy, 256, 0, c->stream>>>(k, ranges, values_sorted, pre3d->gauss2screen,
                                                              pre3d->cov3D_inv, pre3d->mean_offset, pre->opacities,
                                                              pre->depths, pre->rgb, bg[0], bg[1], bg[2], image,
                                                              T_final, n_contrib);
    G3_HIP(hipGetLastError(), "raster3d kernel");
    return GS_OK;
}
