// gs_device.h -- device-side math of the 3DGS 2D path, written for gfx950.
//
// Every function keeps the floating-point operation order of the reference's
// CPU path (harmonyos_3dgs/src/math_utils.cpp, sh_eval.cpp,
// cpu/preprocessor_cpu.cpp) so that, with -ffp-contract=off and IEEE division
// and square root, the GPU reproduces the CPU results bit for bit.  The
// exponential and logarithm the CPU path calls (glibc's expf and logf) are
// restated as exp_ref and log_ref.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_hip.h"

namespace gsd {

// expf exactly as the reference's CPU path evaluates it: glibc's expf, the
// build with fused multiply-add that x86-64 hosts with FMA select (2.35,
// sysdeps/ieee754/flt-32/e_expf.c; the published algorithm of Szabolcs
// Nagy's optimized-routines).  exp(x) = 2^(k/32) * 2^(r/32): with
// z = x * 32/ln2 in double, k = round(z) (via the 1.5 * 2^52 shift), r = z - k,
// 2^(k/32) from a table of 2^(i/32) and the exponent bits of k, and 2^(r/32)
// from a cubic; the double result is rounded to float once.  The table is
// tab[i] = bits(2^(i/32) rounded to double) - (i << 47).  Equal to glibc's
// expf on all 2^32 float inputs (checked exhaustively on the CPU).
//
// The tables below are defined once (GS_EXP2_TAB, GS_LOGF_TAB) and
// instantiated for the device (__constant__) and for the host, so the same
// functions compile for both; a host build of them is checked against the
// system's glibc on every float input.
#define GS_EXP2_TAB                                                                    \
    {0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull, \
     0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, \
     0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull, \
     0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull, \
     0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull, \
     0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull, \
     0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull, \
     0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull}
static __constant__ uint64_t kExp2TabDev[32] = GS_EXP2_TAB;
static const uint64_t kExp2TabHost[32] = GS_EXP2_TAB;

// exp_ref with the 2^(i/32) table at `tab` (a copy of GS_EXP2_TAB, e.g. in LDS:
// the table index differs from lane to lane, so a copy in LDS saves a global
// load's latency per call in loops that evaluate many exponentials).
__host__ __device__ __forceinline__ float exp_ref_tab(float x, const uint64_t* tab) {
    if (!(x <= 0x1.62e42ep6f)) return x != x ? x + x : __builtin_inff();  // NaN, overflow
    if (x < -0x1.9fe368p6f) return 0.0f;                                   // below 2^-150
    constexpr double kInvLn2N = 0x1.71547652b82fep+0 * 32.0, kShift = 0x1.8p+52;
    constexpr double kC0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, kC1 = 0x1.ebfce50fac4f3p-3 / 32 / 32,
                     kC2 = 0x1.62e42ff0c52d6p-1 / 32;
    const double xd = (double)x;
    double kd = __builtin_fma(kInvLn2N, xd, kShift);
    uint64_t ki;
    __builtin_memcpy(&ki, &kd, 8);
    kd -= kShift;
    const double r = __builtin_fma(kInvLn2N, xd, -kd);
    const uint64_t sb = tab[ki & 31u] + (ki << 47);
    double sc;
    __builtin_memcpy(&sc, &sb, 8);
    const double z = __builtin_fma(kC0, r, kC1);
    const double r2 = r * r;
    double y = __builtin_fma(kC2, r, 1.0);
    y = __builtin_fma(z, r2, y);
    return (float)(y * sc);
}

__host__ __device__ __forceinline__ float exp_ref(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return exp_ref_tab(x, kExp2TabDev);
#else
    return exp_ref_tab(x, kExp2TabHost);
#endif
}

// logf exactly as the reference's CPU path evaluates it: glibc's logf (2.35,
// sysdeps/ieee754/flt-32/e_logf.c, the published optimized-routines
// algorithm, fused multiply-adds as in the FMA build).  With x = 2^k z,
// z in [0x1.66p-1, 0x1.66p0) split into 16 intervals by its top mantissa
// bits: log(x) = k ln2 + log(c) + log1p(z/c - 1), where invc = 1/c and
// logc = log(c) (double) are tabulated per interval and log1p(r) is a
// quartic in r evaluated in double; one rounding to float.  Equal to glibc's
// logf on all 2^32 float inputs (NaN payloads aside).
#define GS_LOGF_TAB                                                               \
    {{0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2}, {0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2}, \
     {0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2},  {0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3}, \
     {0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3}, {0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3},   \
     {0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4}, {0x1.12358f08ae5bap+0, -0x1.1973c5a611cccp-4}, \
     {0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5}, {0x1p+0, 0x0p+0},                            \
     {0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5},  {0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4},  \
     {0x1.b2036576afce6p-1, 0x1.526e57720db08p-3},  {0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3},  \
     {0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2},  {0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2}}
struct LogfEntry {
    double invc, logc;
};
static __constant__ LogfEntry kLogfTabDev[16] = GS_LOGF_TAB;
static const LogfEntry kLogfTabHost[16] = GS_LOGF_TAB;

__host__ __device__ __forceinline__ float log_ref(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const LogfEntry* tab = kLogfTabDev;
#else
    const LogfEntry* tab = kLogfTabHost;
#endif
    constexpr double kLn2 = 0x1.62e42fefa39efp-1;
    constexpr double kA0 = -0x1.00ea348b88334p-2, kA1 = 0x1.5575b0be00b6ap-2, kA2 = -0x1.ffffef20a4123p-2;
    constexpr uint32_t kOff = 0x3f330000u;
    uint32_t ix;
    __builtin_memcpy(&ix, &x, 4);
    if (ix == 0x3f800000u) return 0.0f;
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
        // zero, subnormal, negative, infinity or NaN
        if (ix * 2u == 0u) return -__builtin_inff();
        if (ix == 0x7f800000u) return x;
        if ((ix & 0x80000000u) || ix * 2u >= 0xff000000u) return (x - x) / (x - x);
        const float xs = x * 0x1p23f;  // subnormal: normalise
        __builtin_memcpy(&ix, &xs, 4);
        ix -= 23u << 23;
    }
    const uint32_t tmp = ix - kOff;
    const int i = (int)((tmp >> (23 - 4)) % 16u);
    const int k = (int32_t)tmp >> 23;
    const uint32_t iz = ix - (tmp & 0xff800000u);
    float zf;
    __builtin_memcpy(&zf, &iz, 4);
    const double z = (double)zf;
    const double r = __builtin_fma(z, tab[i].invc, -1.0);
    const double y0 = __builtin_fma((double)k, kLn2, tab[i].logc);
    const double r2 = r * r;
    double y = __builtin_fma(kA1, r, kA2);
    y = __builtin_fma(kA0, r2, y);
    y = __builtin_fma(y, r2, y0 + r);
    return (float)y;
}

// std::min / std::max semantics: the first argument wins unless strictly beaten.
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ int smin(int a, int b) { return (b < a) ? b : a; }
__device__ __forceinline__ int smax(int a, int b) { return (a < b) ? b : a; }

// SH basis constants (sh_eval.cpp:5-22).
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f, SH_C2_1 = -1.0925484305920792f,
                SH_C2_2 = 0.31539156525252005f, SH_C2_3 = -1.0925484305920792f,
                SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f, SH_C3_1 = 2.890611442640554f,
                SH_C3_2 = -0.4570457994644658f, SH_C3_3 = 0.3731763325901154f,
                SH_C3_4 = -0.4570457994644658f, SH_C3_5 = 1.445305721320277f,
                SH_C3_6 = -0.5900435899266435f;

constexpr int TILE = 16;

// ---- math_utils.cpp:8-19 ----
__device__ __forceinline__ void xform43(const float p[3], const float* m, float o[3]) {
    o[0] = m[0] * p[0] + m[4] * p[1] + m[8] * p[2] + m[12];
    o[1] = m[1] * p[0] + m[5] * p[1] + m[9] * p[2] + m[13];
    o[2] = m[2] * p[0] + m[6] * p[1] + m[10] * p[2] + m[14];
}
__device__ __forceinline__ void xform44(const float p[3], const float* m, float o[4]) {
    o[0] = m[0] * p[0] + m[4] * p[1] + m[8] * p[2] + m[12];
    o[1] = m[1] * p[0] + m[5] * p[1] + m[9] * p[2] + m[13];
    o[2] = m[2] * p[0] + m[6] * p[1] + m[10] * p[2] + m[14];
    o[3] = m[3] * p[0] + m[7] * p[1] + m[11] * p[2] + m[15];
}

// ---- math_utils.h:12-14 ----
__device__ __forceinline__ float ndc2pix(float v, int S) { return ((v + 1.0f) * (float)S - 1.0f) * 0.5f; }

// ---- math_utils.cpp:21-28 (truncate toward zero, then clamp) ----
__device__ __forceinline__ void get_rect(float px, float py, int r, int gx, int gy, int mn[2], int mx[2]) {
    const float tw = (float)TILE;
    mn[0] = smin(gx, smax(0, (int)((px - (float)r) / tw)));
    mn[1] = smin(gy, smax(0, (int)((py - (float)r) / tw)));
    mx[0] = smin(gx, smax(0, (int)((px + (float)r + tw - 1.0f) / tw)));
    mx[1] = smin(gy, smax(0, (int)((py + (float)r + tw - 1.0f) / tw)));
}

// ---- math_utils.cpp:35-45 (R[row][col]) ----
__device__ __forceinline__ void rot_matrix(float r, float x, float y, float z, float R[3][3]) {
    R[0][0] = 1.f - 2.f * (y * y + z * z);
    R[0][1] = 2.f * (x * y + r * z);
    R[0][2] = 2.f * (x * z - r * y);
    R[1][0] = 2.f * (x * y - r * z);
    R[1][1] = 1.f - 2.f * (x * x + z * z);
    R[1][2] = 2.f * (y * z + r * x);
    R[2][0] = 2.f * (x * z + r * y);
    R[2][1] = 2.f * (y * z - r * x);
    R[2][2] = 1.f - 2.f * (x * x + y * y);
}

// ---- math_utils.cpp:30-62 ----
__device__ __forceinline__ void cov3d_of(const float s_in[3], float mod, const float q[4], float cov[6]) {
    float s[3] = {mod * s_in[0], mod * s_in[1], mod * s_in[2]};
    float R[3][3], M[3][3];
    rot_matrix(q[0], q[1], q[2], q[3], R);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) M[i][j] = s[i] * R[i][j];
    cov[0] = M[0][0] * M[0][0] + M[1][0] * M[1][0] + M[2][0] * M[2][0];
    cov[1] = M[0][0] * M[0][1] + M[1][0] * M[1][1] + M[2][0] * M[2][1];
    cov[2] = M[0][0] * M[0][2] + M[1][0] * M[1][2] + M[2][0] * M[2][2];
    cov[3] = M[0][1] * M[0][1] + M[1][1] * M[1][1] + M[2][1] * M[2][1];
    cov[4] = M[0][1] * M[0][2] + M[1][1] * M[1][2] + M[2][1] * M[2][2];
    cov[5] = M[0][2] * M[0][2] + M[1][2] * M[1][2] + M[2][2] * M[2][2];
}

// Clamped view point, J, W and T = W*J, all stored [col][row]
// (math_utils.cpp:69-109, reused by preprocessor_backward_cpu.cpp:316-351).
struct ProjJac {
    float t[3], txtz, tytz, limx, limy;
    float J00, J02, J11, J12;  // the non-zero entries of J[col][row]
    float T[3][3];
};

__device__ __forceinline__ void proj_jacobian(const float tv[3], const float* vm, float tanx, float tany,
                                              float fx, float fy, ProjJac& o) {
    o.t[0] = tv[0]; o.t[1] = tv[1]; o.t[2] = tv[2];
    o.limx = 1.3f * tanx;
    o.limy = 1.3f * tany;
    o.txtz = o.t[0] / o.t[2];
    o.tytz = o.t[1] / o.t[2];
    o.t[0] = smin(o.limx, smax(-o.limx, o.txtz)) * o.t[2];
    o.t[1] = smin(o.limy, smax(-o.limy, o.tytz)) * o.t[2];
    const float* t = o.t;
    o.J00 = fx / t[2];
    o.J02 = -(fx * t[0]) / (t[2] * t[2]);
    o.J11 = fy / t[2];
    o.J12 = -(fy * t[1]) / (t[2] * t[2]);
    // J[col][row]: col0 = (J00, 0, J02), col1 = (0, J11, J12), col2 = 0.
    // T[col][row] = sum_k W[k][row] * J[col][k] accumulated from 0 in k order;
    // the zero terms are kept so the additions round exactly as the reference.
    const float J[3][3] = {{o.J00, 0.0f, o.J02}, {0.0f, o.J11, o.J12}, {0.0f, 0.0f, 0.0f}};
    const float W[3][3] = {{vm[0], vm[4], vm[8]}, {vm[1], vm[5], vm[9]}, {vm[2], vm[6], vm[10]}};
#pragma unroll
    for (int col = 0; col < 3; col++)
#pragma unroll
        for (int row = 0; row < 3; row++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) acc += W[k][row] * J[col][k];
            o.T[col][row] = acc;
        }
}

// ---- math_utils.cpp:64-144: cov2D entries (a, b, c) before the +0.3 ----
__device__ __forceinline__ void cov2d_from(const ProjJac& pj, const float c3[6], float out[3]) {
    const float V[3][3] = {{c3[0], c3[1], c3[2]}, {c3[1], c3[3], c3[4]}, {c3[2], c3[4], c3[5]}};
    float tmp[2][3];
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
        for (int row = 0; row < 3; row++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) acc += V[row][k] * pj.T[col][k];
            tmp[col][row] = acc;
        }
    // res[col][row] = sum_k T[row][k] * tmp[col][k]
    float r00 = 0.0f, r10 = 0.0f, r11 = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; k++) r00 += pj.T[0][k] * tmp[0][k];
#pragma unroll
    for (int k = 0; k < 3; k++) r10 += pj.T[0][k] * tmp[1][k];
#pragma unroll
    for (int k = 0; k < 3; k++) r11 += pj.T[1][k] * tmp[1][k];
    out[0] = r00;
    out[1] = r10;
    out[2] = r11;
}

// SH direction terms: dir = normalize(pos - cam), and the basis values the
// forward and backward share (sh_eval.cpp:30-37).
struct ShDir {
    float x, y, z, len;
};
__device__ __forceinline__ ShDir sh_dir(const float p[3], const float cam[3]) {
    float dx = p[0] - cam[0], dy = p[1] - cam[1], dz = p[2] - cam[2];
    float len = sqrtf(dx * dx + dy * dy + dz * dz);
    dx /= len; dy /= len; dz /= len;
    return ShDir{dx, dy, dz, len};
}

// ---- sh_eval.cpp:39-77: the colour before +0.5 and clamping, one channel.
// sh points at coefficient 0 of this channel; stride 3 floats per basis. ----
template <int DEG>
__device__ __forceinline__ float sh_raw_channel(const ShDir& d, const float* sh) {
    const float x = d.x, y = d.y, z = d.z;
    float v = SH_C0 * sh[0];
    if (DEG > 0) {
        v += -SH_C1 * y * sh[1 * 3] + SH_C1 * z * sh[2 * 3] + -SH_C1 * x * sh[3 * 3];
        if (DEG > 1) {
            float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            v += SH_C2_0 * xy * sh[4 * 3] + SH_C2_1 * yz * sh[5 * 3] + SH_C2_2 * (2.0f * zz - xx - yy) * sh[6 * 3] +
                 SH_C2_3 * xz * sh[7 * 3] + SH_C2_4 * (xx - yy) * sh[8 * 3];
            if (DEG > 2) {
                v += SH_C3_0 * y * (3.0f * xx - yy) * sh[9 * 3] + SH_C3_1 * xy * z * sh[10 * 3] +
                     SH_C3_2 * y * (4.0f * zz - xx - yy) * sh[11 * 3] +
                     SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * sh[12 * 3] +
                     SH_C3_4 * x * (4.0f * zz - xx - yy) * sh[13 * 3] + SH_C3_5 * z * (xx - yy) * sh[14 * 3] +
                     SH_C3_6 * x * (xx - 3.0f * yy) * sh[15 * 3];
            }
        }
    }
    return v;
}

__device__ __forceinline__ float sh_raw_channel_rt(int deg, const ShDir& d, const float* sh) {
    switch (deg) {
        case 0: return sh_raw_channel<0>(d, sh);
        case 1: return sh_raw_channel<1>(d, sh);
        case 2: return sh_raw_channel<2>(d, sh);
        default: return sh_raw_channel<3>(d, sh);
    }
}

// Activation of one Gaussian's raw parameters (train_types.cpp:6-36).
__device__ __forceinline__ void activate_scale(const float rs[3], float s[3]) {
    s[0] = exp_ref(rs[0]); s[1] = exp_ref(rs[1]); s[2] = exp_ref(rs[2]);
}
__device__ __forceinline__ void activate_rot(const float q[4], float o[4]) {
    float len = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (len < 1e-12f) len = 1e-12f;
    float inv = 1.0f / len;
    o[0] = q[0] * inv; o[1] = q[1] * inv; o[2] = q[2] * inv; o[3] = q[3] * inv;
}
__device__ __forceinline__ float activate_opacity(float l) { return 1.0f / (1.0f + exp_ref(-l)); }

// e^x for the rasterizers: the hardware exp2 of x log2(e).  The product's
// rounding adds at most |x| ln(2) 2^-24 to the relative error.  The
// rasterizers use the value only where o e^x >= 1/255, i.e. x > -ln(255)
// > -5.6, where the result is within 2^-21 (relative) of e^x; alpha_fix
// recomputes exactly wherever that could decide the 1/255 cut.
__device__ __forceinline__ float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

// alpha = min(0.99, o * exp(power)) of rasterizer_cpu.cpp:164-166 and the
// 1/255 cut that follows, in two halves, for code where all 64 lanes run
// together: alpha_fast gives the alpha from exp_fast and whether it lies
// within 2^-18 (relative) of the cut, where exp_fast's error could decide it;
// where some lane of the wave is that near (a uniform branch), alpha_fix
// recomputes with exp_ref, every lane computing and the near ones taking it.
// So every include/skip decision is the CPU path's, and the common path
// carries no exec-mask bookkeeping.
__device__ __forceinline__ float alpha_fast(float o, float power, float* ep, bool* near) {
    const float e = exp_fast(power);
    const float a = smin(0.99f, o * e);
    constexpr float kCut = 1.0f / 255.0f;
    *near = fabsf(a - kCut) <= kCut * 0x1p-18f;
    *ep = e;
    return a;
}

__device__ __forceinline__ void alpha_fix(float o, float power, bool near, float* alpha, float* ep) {
    const float er = exp_ref(power);
    *ep = near ? er : *ep;
    *alpha = smin(0.99f, o * *ep);
}

__device__ __forceinline__ float raster_alpha_wave(float o, float power, float* ep) {
    bool near;
    float a = alpha_fast(o, power, ep, &near);
    if (__ballot(near) != 0ull) alpha_fix(o, power, near, &a, ep);
    return a;
}

// Camera constants shared by the per-Gaussian kernels, passed by value.
struct CamArgs {
    float view[16];
    float vp[16];
    float campos[3];
    float tanx, tany, fx, fy;
    int W, H, gx, gy;
    float scale_mod;
    int antialiasing, training, deg_cfg;
};

}  // namespace gsd

namespace gsd {
// The depth of a Gaussian along a unit ray (math_utils.cpp:616-630): ci its
// inverse 3D covariance (upper triangle), mo its mean relative to the camera.
__device__ __forceinline__ float depth_along(const float* ci, const float* mo, const float* vd) {
    const float s0 = ci[0] * vd[0] + ci[1] * vd[1] + ci[2] * vd[2];
    const float s1 = ci[1] * vd[0] + ci[3] * vd[1] + ci[4] * vd[2];
    const float s2 = ci[2] * vd[0] + ci[4] * vd[1] + ci[5] * vd[2];
    const float num = mo[0] * s0 + mo[1] * s1 + mo[2] * s2;
    const float den = vd[0] * s0 + vd[1] * s1 + vd[2] * s2;
    if (fabsf(den) < 1e-10f) return sqrtf(mo[0] * mo[0] + mo[1] * mo[1] + mo[2] * mo[2]);
    return num / den;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8) and each XCD has its own L2.  This bijection on [0, nb) gives
// every XCD one contiguous run of work items, so items that share data
// (neighbouring tiles, neighbouring chunks) share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    constexpr int XCD = 8;
    const int k = nb / XCD, r = nb % XCD;
    const int x = b % XCD, q = b / XCD;
    return x * k + (x < r ? x : r) + q;
}
}  // namespace gsd

// Kernel argument blocks shared by the launchers and the C ABI layer.
namespace gsr {
struct RasterArgs {
    int W, H, gx;
    const float2* means2D;
    const float* conics;
    const float* opac;
    const float* rgb;
    const uint32_t* values;  // Gaussian index of sorted pair j at values[j * vstride]
    const uint32_t* ranges;
    float bg0, bg1, bg2;
    int vstride;  // 1: uint32 value array; 2: low words of tile<<32|index pairs
};
// Forward raster in two rounds over a pair list placed in two parts (one
// view): the binning placed only the first g1 depth groups of every tile's
// list; round 1 renders each tile up to where group g1 begins, writes every
// tile's flag (1: not finished there) and saves the pixels' state of the
// unfinished ones; the rest of the list is then placed for bins with a
// flagged tile only, and round 2 resumes the flagged tiles.  round 0: one
// pass (no split).
struct RasterSplit {
    const uint32_t* scanned;  // the scanned binning table: slot of (tile, group)'s first pair
    int C, g1;                // groups per view; groups placed before round 1 (g1 < C)
    uint32_t* tile_flag;      // per tile, written by round 1
    float* rc;                // per pixel: accumulated colour (3 floats) of an unfinished tile
    int round;                // 0, 1 or 2
};

struct GradArgs {
    float2* d_means2D;
    float* d_conics;
    float* d_rgb;
    float* d_op;
};
}  // namespace gsr

namespace gsr {
// A batch of views rasterized in one launch over a global tile list.
struct MultiRaster {
    int V, N, total_tiles;
    const int4* views;     // per view: (W, H, tiles per row, first global tile)
    const int* pix_off;    // per view: first pixel in the concatenated images (V + 1 entries)
    const float2* means2D; // view-major [V][N]
    const float* conics;   // [V][N*3]
    const float* opac;     // [V][N]
    const float* rgb;      // [V][N*3]
    const uint32_t* values;
    int vstride;
    const uint32_t* ranges;  // [total_tiles * 2]
    float bg0, bg1, bg2;
    unsigned int* counter;  // persistent work queue (nullptr: one group per tile)
};
struct MultiPix {
    float* img;
    float* T_final;
    int* n_contrib;
    uint32_t* last_contrib;
    const float* d_image;
};
struct MultiGrads {  // view-major [V][N]
    float2* d_means2D;
    float* d_conics;
    float* d_rgb;
    float* d_op;
};
}  // namespace gsr

namespace gsb {
// Direct-placement binning over V views of N Gaussians each (k_bin.hip).
struct BinArgs {
    int V, N;
    int C, G;               // groups per view; Gaussians per group after the first S
    int S, G1;              // the first S groups hold G1 Gaussians each (0: none)
    const uint32_t* sidx;   // [V*N]: each view's Gaussian indices in depth order, sidx_stride apart
    int sidx_stride;        // 1: an index array; 2: the high words of sorted index << 32 | key words
    const float2* means2D;  // view-major [V][N]
    const int* radii;       // [V][N]
    const int4* views;      // per view: (tiles per row, tile rows, table offset, first global tile)
    int4 view0;             // used instead when views == nullptr (one view)
    const int* vblock0;     // per view: first block (V + 1 entries); nullptr for one view
    uint32_t* table;        // count: per-(view, tile, group) counts; place: their exclusive scan
    size_t table_len;       // sum over views of tiles * C, plus one trailing entry
    uint2* rects;           // [V*N] scratch: depth-ordered (x0|x1<<8|y0<<16|y1<<24, index)
    uint32_t* out;          // place: pair values (Gaussian index); 16-byte aligned
    uint32_t capacity;
    uint32_t* overflow;
    int diff_len;           // max over views of (tiles per row + 1) * (tile rows + 1); 0: no LDS count
    const uint32_t* rect_idx;  // optional [V][N]: packed rectangles by Gaussian index (else from means2D, radii)
};
}  // namespace gsb

namespace gso {
// Adam hyper-parameters read from device memory (updated per step outside a
// captured graph).
struct AdamDev {
    float lr[5];
    float b1, b2, eps, bc1, bc2;
};
struct AdamSeg {
    float* p;
    const float* g;
    float* m;
    float* v;
    uint32_t n;
    float lr;
};
struct AdamArgs {
    AdamSeg seg[5];
    float b1, b2, eps, bc1, bc2;
    const AdamDev* hp;  // when set, learning rates and the rest come from device memory
    const uint32_t* skip = nullptr;  // when set and nonzero, the launch updates nothing
};
}  // namespace gso
