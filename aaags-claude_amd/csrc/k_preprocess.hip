// k_preprocess.hip -- per-Gaussian kernels: activation, preprocess forward and
// preprocess backward (the full adjoint of the 2D projection, SH and
// activation chain).
//
// One thread per Gaussian, 256-thread workgroups.  The kernels are HBM-bound:
// each Gaussian's raw parameters (56 B + 12 B per SH basis) are read once; the
// SH block of a visible Gaussian is read with 16-byte loads when the model has
// 4 or 16 bases.  Forward quantities the backward needs (cov3D, view point,
// clip w, filtered cov2D, det) are recomputed from the parameters rather than
// stored, which trades ~100 FLOPs for 56 B/Gaussian of HBM traffic each way.
#include "gs_device.h"

namespace gsd {

struct Geom {
    float pv[3];  // view-space position
    float phw;    // clip-space w
    float px, py; // pixel position
    float c3[6];  // 3D covariance
    float c2[3];  // 2D covariance after the +0.3 filter
    float det;
    float conic[3];
    float hconv;
    int radius;
    int tiles;
};

// 2D-branch geometry of preprocessor_cpu.cpp:188-236.  Returns false when the
// Gaussian is culled.
__device__ __forceinline__ bool preprocess_geom(const float p[3], const float s[3], const float q[4],
                                                const CamArgs& c, Geom& g) {
    xform43(p, c.view, g.pv);
    if (g.pv[2] <= 0.2f) return false;
    float ph[4];
    xform44(p, c.vp, ph);
    g.phw = ph[3];
    float pw = 1.0f / (ph[3] + 1e-7f);
    float ndc0 = ph[0] * pw, ndc1 = ph[1] * pw;
    g.px = ndc2pix(ndc0, c.W);
    g.py = ndc2pix(ndc1, c.H);
    cov3d_of(s, c.scale_mod, q, g.c3);
    ProjJac pj;
    float tv[3];
    xform43(p, c.view, tv);
    proj_jacobian(tv, c.view, c.tanx, c.tany, c.fx, c.fy, pj);
    cov2d_from(pj, g.c3, g.c2);
    float det_cov = g.c2[0] * g.c2[2] - g.c2[1] * g.c2[1];
    g.c2[0] += 0.3f;
    g.c2[2] += 0.3f;
    float det_h = g.c2[0] * g.c2[2] - g.c2[1] * g.c2[1];
    g.hconv = 1.0f;
    if (c.antialiasing) g.hconv = sqrtf(smax(0.000025f, det_cov / det_h));
    g.det = det_h;
    if (g.det == 0.0f) return false;
    float dinv = 1.0f / g.det;
    g.conic[0] = g.c2[2] * dinv;
    g.conic[1] = -g.c2[1] * dinv;
    g.conic[2] = g.c2[0] * dinv;
    float mid = 0.5f * (g.c2[0] + g.c2[2]);
    float l1 = mid + sqrtf(smax(0.1f, mid * mid - g.det));
    float l2 = mid - sqrtf(smax(0.1f, mid * mid - g.det));
    g.radius = (int)ceilf(3.0f * sqrtf(smax(l1, l2)));
    if (g.radius > smax(c.W, c.H)) return false;
    int mn[2], mx[2];
    get_rect(g.px, g.py, g.radius, c.gx, c.gy, mn, mx);
    g.tiles = (mx[1] - mn[1]) * (mx[0] - mn[0]);
    if ((mx[0] - mn[0]) * (mx[1] - mn[1]) == 0) return false;
    return true;
}

struct Params {
    float p[3], s[3], q[4], op;
    float qraw[4];
};

__device__ __forceinline__ void load_params(const gs_gaussians& G, int i, bool raw, Params& P) {
    P.p[0] = G.positions[i * 3]; P.p[1] = G.positions[i * 3 + 1]; P.p[2] = G.positions[i * 3 + 2];
    float sc[3] = {G.scales[i * 3], G.scales[i * 3 + 1], G.scales[i * 3 + 2]};
    float4 qv = reinterpret_cast<const float4*>(G.rotations)[i];
    P.qraw[0] = qv.x; P.qraw[1] = qv.y; P.qraw[2] = qv.z; P.qraw[3] = qv.w;
    float o = G.opacities[i];
    if (raw) {
        activate_scale(sc, P.s);
        activate_rot(P.qraw, P.q);
        P.op = activate_opacity(o);
    } else {
        P.s[0] = sc[0]; P.s[1] = sc[1]; P.s[2] = sc[2];
        P.q[0] = P.qraw[0]; P.q[1] = P.qraw[1]; P.q[2] = P.qraw[2]; P.q[3] = P.qraw[3];
        P.op = o;
    }
}

// Loads the first NB bases (3 floats each) of a Gaussian's interleaved SH
// block into registers, with 16-byte loads when the block is 16-byte aligned.
template <int MAXC, int NB>
__device__ __forceinline__ void load_sh(const float* sh_all, int i, float* out) {
    const float* src = sh_all + (size_t)i * (MAXC * 3);
    if constexpr ((MAXC * 3) % 4 == 0 && (NB * 3) % 4 == 0) {
        const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
        for (int k = 0; k < NB * 3 / 4; k++) {
            float4 v = s4[k];
            out[4 * k] = v.x; out[4 * k + 1] = v.y; out[4 * k + 2] = v.z; out[4 * k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < NB * 3; k++) out[k] = src[k];
    }
}

template <int MAXC, int DEG>
__device__ __forceinline__ void sh_color(const float* sh_all, int i, const float p[3], const float campos[3],
                                         bool training, float rgb[3]) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    float sh[NB * 3];
    load_sh<MAXC, NB>(sh_all, i, sh);
    ShDir d = sh_dir(p, campos);
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float v = sh_raw_channel<DEG>(d, sh + ch);
        v = smax(0.0f, v + 0.5f);
        if (!training) v = smin(1.0f, v);
        rgb[ch] = v;
    }
}

template <int MAXC>
__device__ __forceinline__ void sh_color_rt(int deg, const float* sh_all, int i, const float p[3],
                                            const float campos[3], bool training, float rgb[3]) {
    if (deg <= 0) sh_color<MAXC, 0>(sh_all, i, p, campos, training, rgb);
    else if (deg == 1) { if constexpr (MAXC >= 4) sh_color<MAXC, 1>(sh_all, i, p, campos, training, rgb); }
    else if (deg == 2) { if constexpr (MAXC >= 9) sh_color<MAXC, 2>(sh_all, i, p, campos, training, rgb); }
    else { if constexpr (MAXC >= 16) sh_color<MAXC, 3>(sh_all, i, p, campos, training, rgb); }
}

template <int MAXC>
__global__ __launch_bounds__(256) void preprocess_fwd_kernel(gs_gaussians G, CamArgs c, gs_preprocessed out,
                                                             int deg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.count) return;
    Params P;
    load_params(G, i, G.raw != 0, P);
    Geom g;
    bool vis = preprocess_geom(P.p, P.s, P.q, c, g);
    float2 m2 = make_float2(0.0f, 0.0f);
    float con[3] = {0.0f, 0.0f, 0.0f}, rgb[3] = {0.0f, 0.0f, 0.0f};
    float depth = 0.0f, op2d = 0.0f;
    int radius = 0, tiles = 0;
    if (vis) {
        sh_color_rt<MAXC>(deg, G.sh_coeffs, i, P.p, c.campos, c.training != 0, rgb);
        m2 = make_float2(g.px, g.py);
        con[0] = g.conic[0]; con[1] = g.conic[1]; con[2] = g.conic[2];
        depth = g.pv[2];
        op2d = P.op * g.hconv;
        radius = g.radius;
        tiles = g.tiles;
    }
    reinterpret_cast<float2*>(out.means2D)[i] = m2;
    out.depths[i] = depth;
    out.conics[i * 3] = con[0]; out.conics[i * 3 + 1] = con[1]; out.conics[i * 3 + 2] = con[2];
    out.opacities[i] = op2d;
    out.rgb[i * 3] = rgb[0]; out.rgb[i * 3 + 1] = rgb[1]; out.rgb[i * 3 + 2] = rgb[2];
    out.radii[i] = radius;
    out.tiles_touched[i] = tiles;
}

__global__ __launch_bounds__(256) void activate_kernel(gs_gaussians G, float* pos, float* sc, float* rot, float* sh,
                                                       float* op) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.count) return;
    Params P;
    load_params(G, i, true, P);
    pos[i * 3] = P.p[0]; pos[i * 3 + 1] = P.p[1]; pos[i * 3 + 2] = P.p[2];
    sc[i * 3] = P.s[0]; sc[i * 3 + 1] = P.s[1]; sc[i * 3 + 2] = P.s[2];
    reinterpret_cast<float4*>(rot)[i] = make_float4(P.q[0], P.q[1], P.q[2], P.q[3]);
    op[i] = P.op;
    int nsh = G.max_coeffs * 3;
    for (int k = 0; k < nsh; k++) sh[(size_t)i * nsh + k] = G.sh_coeffs[(size_t)i * nsh + k];
}

// ---- preprocessor_backward_cpu.cpp:8-203: d_sh (written) and the SH term
// of d_pos (accumulated into dpos). ----
template <int MAXC, int DEG>
__device__ __forceinline__ void sh_backward(const float* sh_all, int i, const float p[3], const float campos[3],
                                            const float drgb[3], float* gsh_out, bool accumulate,
                                            float dpos[3]) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    float sh[NB * 3];
    load_sh<MAXC, NB>(sh_all, i, sh);
    ShDir d = sh_dir(p, campos);
    const float x = d.x, y = d.y, z = d.z;
    float de[3];
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float raw = sh_raw_channel<DEG>(d, sh + ch);
        float mask = (raw + 0.5f >= 0.0f) ? 1.0f : 0.0f;
        de[ch] = drgb[ch] * mask;
    }
    float b[16];
    b[0] = SH_C0;
    if (DEG > 0) {
        b[1] = -SH_C1 * y; b[2] = SH_C1 * z; b[3] = -SH_C1 * x;
        if (DEG > 1) {
            float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            b[4] = SH_C2_0 * xy; b[5] = SH_C2_1 * yz; b[6] = SH_C2_2 * (2.0f * zz - xx - yy);
            b[7] = SH_C2_3 * xz; b[8] = SH_C2_4 * (xx - yy);
            if (DEG > 2) {
                b[9] = SH_C3_0 * y * (3.0f * xx - yy);
                b[10] = SH_C3_1 * xy * z;
                b[11] = SH_C3_2 * y * (4.0f * zz - xx - yy);
                b[12] = SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
                b[13] = SH_C3_4 * x * (4.0f * zz - xx - yy);
                b[14] = SH_C3_5 * z * (xx - yy);
                b[15] = SH_C3_6 * x * (xx - 3.0f * yy);
            }
        }
    }
    float* dst = gsh_out + (size_t)i * (MAXC * 3);
    // Bases above DEG get a zero gradient (the reference zeroes the whole block).
#pragma unroll
    for (int k = 0; k < MAXC; k++)
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float v = (k < NB) ? b[k < NB ? k : 0] * de[ch] : 0.0f;
            if (accumulate) dst[k * 3 + ch] += v;
            else dst[k * 3 + ch] = v;
        }
    if (DEG > 0) {
        float dd[3][3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            dd[0][ch] = 0.0f + -SH_C1 * sh[3 * 3 + ch];
            dd[1][ch] = 0.0f + -SH_C1 * sh[1 * 3 + ch];
            dd[2][ch] = 0.0f + SH_C1 * sh[2 * 3 + ch];
        }
        if (DEG > 1) {
            float xx = x * x, yy = y * y, zz = z * z;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                float s4 = sh[4 * 3 + ch], s5 = sh[5 * 3 + ch], s6 = sh[6 * 3 + ch], s7 = sh[7 * 3 + ch],
                      s8 = sh[8 * 3 + ch];
                dd[0][ch] += SH_C2_0 * y * s4 - 2.0f * SH_C2_2 * x * s6 + SH_C2_3 * z * s7 + 2.0f * SH_C2_4 * x * s8;
                dd[1][ch] += SH_C2_0 * x * s4 + SH_C2_1 * z * s5 - 2.0f * SH_C2_2 * y * s6 - 2.0f * SH_C2_4 * y * s8;
                dd[2][ch] += SH_C2_1 * y * s5 + 4.0f * SH_C2_2 * z * s6 + SH_C2_3 * x * s7;
            }
            if (DEG > 2) {
#pragma unroll
                for (int ch = 0; ch < 3; ch++) {
                    float s9 = sh[9 * 3 + ch], s10 = sh[10 * 3 + ch], s11 = sh[11 * 3 + ch], s12 = sh[12 * 3 + ch],
                          s13 = sh[13 * 3 + ch], s14 = sh[14 * 3 + ch], s15 = sh[15 * 3 + ch];
                    dd[0][ch] += SH_C3_0 * 6.0f * x * y * s9 + SH_C3_1 * y * z * s10 +
                                 SH_C3_2 * (-2.0f * x * y) * s11 + SH_C3_3 * (-6.0f * x * z) * s12 +
                                 SH_C3_4 * (4.0f * zz - 3.0f * xx - yy) * s13 + SH_C3_5 * 2.0f * x * z * s14 +
                                 SH_C3_6 * (3.0f * xx - 3.0f * yy) * s15;
                    dd[1][ch] += SH_C3_0 * (3.0f * xx - 3.0f * yy) * s9 + SH_C3_1 * x * z * s10 +
                                 SH_C3_2 * (4.0f * zz - xx - 3.0f * yy) * s11 + SH_C3_3 * (-6.0f * y * z) * s12 +
                                 SH_C3_4 * (-2.0f * x * y) * s13 - SH_C3_5 * 2.0f * y * z * s14 +
                                 SH_C3_6 * (-6.0f * x * y) * s15;
                    dd[2][ch] += SH_C3_1 * x * y * s10 + SH_C3_2 * 8.0f * y * z * s11 +
                                 SH_C3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy) * s12 + SH_C3_4 * 8.0f * x * z * s13 +
                                 SH_C3_5 * (xx - yy) * s14;
                }
            }
        }
        float ddir[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float acc = 0.0f;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) acc += de[ch] * dd[k][ch];
            ddir[k] = acc;
        }
        float inv_len = 1.0f / d.len;
        float dot = ddir[0] * x + ddir[1] * y + ddir[2] * z;
        dpos[0] += (ddir[0] - x * dot) * inv_len;
        dpos[1] += (ddir[1] - y * dot) * inv_len;
        dpos[2] += (ddir[2] - z * dot) * inv_len;
    }
}

template <int MAXC>
__device__ __forceinline__ void sh_backward_rt(int deg, const float* sh_all, int i, const float p[3],
                                               const float campos[3], const float drgb[3], float* gsh,
                                               bool accumulate, float dpos[3]) {
    if (deg <= 0) sh_backward<MAXC, 0>(sh_all, i, p, campos, drgb, gsh, accumulate, dpos);
    else if (deg == 1) { if constexpr (MAXC >= 4) sh_backward<MAXC, 1>(sh_all, i, p, campos, drgb, gsh, accumulate, dpos); }
    else if (deg == 2) { if constexpr (MAXC >= 9) sh_backward<MAXC, 2>(sh_all, i, p, campos, drgb, gsh, accumulate, dpos); }
    else { if constexpr (MAXC >= 16) sh_backward<MAXC, 3>(sh_all, i, p, campos, drgb, gsh, accumulate, dpos); }
}

// ---- preprocessor_backward_cpu.cpp:205-645 ----
template <int MAXC>
__global__ __launch_bounds__(256) void preprocess_bwd_kernel(gs_gaussians G, CamArgs c, const int* __restrict__ radii,
                                                             gs_raster_grads rg, gs_param_grads out, int accumulate,
                                                             int sh_deg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.count) return;
    const bool acc = accumulate != 0;
    if (radii[i] <= 0) {
        if (!acc) {
            out.d_positions[i * 3] = 0.0f; out.d_positions[i * 3 + 1] = 0.0f; out.d_positions[i * 3 + 2] = 0.0f;
            out.d_scales[i * 3] = 0.0f; out.d_scales[i * 3 + 1] = 0.0f; out.d_scales[i * 3 + 2] = 0.0f;
            reinterpret_cast<float4*>(out.d_rotations)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (int k = 0; k < MAXC * 3; k++) out.d_sh_coeffs[(size_t)i * MAXC * 3 + k] = 0.0f;
            out.d_opacities[i] = 0.0f;
        }
        return;
    }
    Params P;
    load_params(G, i, true, P);
    Geom g;
    preprocess_geom(P.p, P.s, P.q, c, g);

    float gpos[3] = {0.0f, 0.0f, 0.0f};
    float drgb[3] = {rg.d_rgb[i * 3], rg.d_rgb[i * 3 + 1], rg.d_rgb[i * 3 + 2]};
    sh_backward_rt<MAXC>(sh_deg, G.sh_coeffs, i, P.p, c.campos, drgb, out.d_sh_coeffs, acc, gpos);

    float sig = P.op;
    float gop = rg.d_opacities_2d[i] * sig * (1.0f - sig);

    // conic -> cov2D
    const float a = g.c2[0], b = g.c2[1], cc = g.c2[2];
    const float inv_det = 1.0f / g.det;
    const float dc0 = rg.d_conics[i * 3], dc1 = rg.d_conics[i * 3 + 1], dc2 = rg.d_conics[i * 3 + 2];
    const float id2 = inv_det * inv_det;
    const float d_a = dc0 * (-cc * cc * id2) + dc1 * (b * cc * id2) + dc2 * (-b * b * id2);
    const float d_b = dc0 * (2.0f * b * cc * id2) + dc1 * (-(a * cc + b * b) * id2) + dc2 * (2.0f * a * b * id2);
    const float d_c = dc0 * (-b * b * id2) + dc1 * (a * b * id2) + dc2 * (-a * a * id2);

    // cov2D -> cov3D
    ProjJac pj;
    proj_jacobian(g.pv, c.view, c.tanx, c.tany, c.fx, c.fy, pj);
    float dV[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            dV[k][j] = d_a * pj.T[0][k] * pj.T[0][j] + d_b * pj.T[0][k] * pj.T[1][j] + d_c * pj.T[1][k] * pj.T[1][j];
    const float dc3[6] = {dV[0][0], dV[0][1] + dV[1][0], dV[0][2] + dV[2][0],
                          dV[1][1], dV[1][2] + dV[2][1], dV[2][2]};

    // cov3D -> M -> scale, rotation
    const float mod = c.scale_mod;
    const float s[3] = {mod * P.s[0], mod * P.s[1], mod * P.s[2]};
    const float rq = P.q[0], xq = P.q[1], yq = P.q[2], zq = P.q[3];
    float R[3][3], M[3][3];
    rot_matrix(rq, xq, yq, zq, R);
#pragma unroll
    for (int ii = 0; ii < 3; ii++)
#pragma unroll
        for (int jj = 0; jj < 3; jj++) M[ii][jj] = s[ii] * R[ii][jj];
    float dM[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        dM[k][0] = 2.0f * dc3[0] * M[k][0] + dc3[1] * M[k][1] + dc3[2] * M[k][2];
        dM[k][1] = dc3[1] * M[k][0] + 2.0f * dc3[3] * M[k][1] + dc3[4] * M[k][2];
        dM[k][2] = dc3[2] * M[k][0] + dc3[4] * M[k][1] + 2.0f * dc3[5] * M[k][2];
    }
    float dscale[3], dR[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float acc_s = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            acc_s += dM[k][j] * R[k][j];
            dR[k][j] = dM[k][j] * s[k];
        }
        dscale[k] = acc_s;
    }
    float gsc[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gsc[k] = dscale[k] * mod * P.s[k];

    float dq0 = 0.0f, dq1 = 0.0f, dq2 = 0.0f, dq3 = 0.0f;
    dq2 += dR[0][0] * (-4.0f * yq);
    dq3 += dR[0][0] * (-4.0f * zq);
    dq0 += dR[0][1] * 2.0f * zq;
    dq1 += dR[0][1] * 2.0f * yq;
    dq2 += dR[0][1] * 2.0f * xq;
    dq3 += dR[0][1] * 2.0f * rq;
    dq0 += dR[0][2] * (-2.0f * yq);
    dq1 += dR[0][2] * 2.0f * zq;
    dq2 += dR[0][2] * (-2.0f * rq);
    dq3 += dR[0][2] * 2.0f * xq;
    dq0 += dR[1][0] * (-2.0f * zq);
    dq1 += dR[1][0] * 2.0f * yq;
    dq2 += dR[1][0] * 2.0f * xq;
    dq3 += dR[1][0] * (-2.0f * rq);
    dq1 += dR[1][1] * (-4.0f * xq);
    dq3 += dR[1][1] * (-4.0f * zq);
    dq0 += dR[1][2] * 2.0f * xq;
    dq1 += dR[1][2] * 2.0f * rq;
    dq2 += dR[1][2] * 2.0f * zq;
    dq3 += dR[1][2] * 2.0f * yq;
    dq0 += dR[2][0] * 2.0f * yq;
    dq1 += dR[2][0] * 2.0f * zq;
    dq2 += dR[2][0] * 2.0f * rq;
    dq3 += dR[2][0] * 2.0f * xq;
    dq0 += dR[2][1] * (-2.0f * xq);
    dq1 += dR[2][1] * (-2.0f * rq);
    dq2 += dR[2][1] * 2.0f * zq;
    dq3 += dR[2][1] * 2.0f * yq;
    dq1 += dR[2][2] * (-4.0f * xq);
    dq2 += dR[2][2] * (-4.0f * yq);

    const float* qr = P.qraw;
    float qlen = sqrtf(qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    if (qlen < 1e-12f) qlen = 1e-12f;
    const float il = 1.0f / qlen;
    const float dotq = dq0 * rq + dq1 * xq + dq2 * yq + dq3 * zq;
    float4 grot = make_float4((dq0 - rq * dotq) * il, (dq1 - xq * dotq) * il, (dq2 - yq * dotq) * il,
                              (dq3 - zq * dotq) * il);

    // Path A: cov2D -> T -> J -> clamped view point -> position
    const float* c3 = g.c3;
    const float V[3][3] = {{c3[0], c3[1], c3[2]}, {c3[1], c3[3], c3[4]}, {c3[2], c3[4], c3[5]}};
    float VT[2][3];
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
        for (int row = 0; row < 3; row++) {
            float acc_v = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) acc_v += V[row][k] * pj.T[col][k];
            VT[col][row] = acc_v;
        }
    float dT[2][3];
#pragma unroll
    for (int qq = 0; qq < 3; qq++) {
        dT[0][qq] = 2.0f * d_a * VT[0][qq] + d_b * VT[1][qq];
        dT[1][qq] = d_b * VT[0][qq] + 2.0f * d_c * VT[1][qq];
    }
    const float* vm = c.view;
    const float W[3][3] = {{vm[0], vm[4], vm[8]}, {vm[1], vm[5], vm[9]}, {vm[2], vm[6], vm[10]}};
    float dJ[2][3];
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float acc_j = 0.0f;
#pragma unroll
            for (int row = 0; row < 3; row++) acc_j += dT[col][row] * W[k][row];
            dJ[col][k] = acc_j;
        }
    const float* t = pj.t;
    const float tz2 = t[2] * t[2];
    const float tz3 = tz2 * t[2];
    float dt0 = dJ[0][2] * (-c.fx / tz2);
    float dt1 = dJ[1][2] * (-c.fy / tz2);
    float dt2 = dJ[0][0] * (-c.fx / tz2) + dJ[1][1] * (-c.fy / tz2) + dJ[0][2] * (2.0f * c.fx * t[0] / tz3) +
                dJ[1][2] * (2.0f * c.fy * t[1] / tz3);
    if (fabsf(pj.txtz) >= pj.limx) dt0 = 0.0f;
    if (fabsf(pj.tytz) >= pj.limy) dt1 = 0.0f;
    const float dt[3] = {dt0, dt1, dt2};
#pragma unroll
    for (int col = 0; col < 3; col++)
#pragma unroll
        for (int row = 0; row < 3; row++) gpos[col] += dt[row] * vm[col * 4 + row];

    // means2D -> position through the projection
    const float dpx = rg.d_means2D[i * 2], dpy = rg.d_means2D[i * 2 + 1];
    const float dnx = dpx * (float)c.W * 0.5f;
    const float dny = dpy * (float)c.H * 0.5f;
    const float wh = g.phw + 1e-7f;
    const float iw = 1.0f / wh;
    const float nx = (2.0f * g.px + 1.0f) / (float)c.W - 1.0f;
    const float ny = (2.0f * g.py + 1.0f) / (float)c.H - 1.0f;
    const float dph0 = dnx * iw, dph1 = dny * iw, dph2 = 0.0f;
    const float dph3 = -(nx * dnx + ny * dny) * iw;
    const float* VP = c.vp;
#pragma unroll
    for (int dim = 0; dim < 3; dim++)
        gpos[dim] += VP[dim * 4 + 0] * dph0 + VP[dim * 4 + 1] * dph1 + VP[dim * 4 + 2] * dph2 + VP[dim * 4 + 3] * dph3;

    if (acc) {
        out.d_positions[i * 3] += gpos[0]; out.d_positions[i * 3 + 1] += gpos[1]; out.d_positions[i * 3 + 2] += gpos[2];
        out.d_scales[i * 3] += gsc[0]; out.d_scales[i * 3 + 1] += gsc[1]; out.d_scales[i * 3 + 2] += gsc[2];
        float4 r0 = reinterpret_cast<float4*>(out.d_rotations)[i];
        reinterpret_cast<float4*>(out.d_rotations)[i] =
            make_float4(r0.x + grot.x, r0.y + grot.y, r0.z + grot.z, r0.w + grot.w);
        out.d_opacities[i] += gop;
    } else {
        out.d_positions[i * 3] = gpos[0]; out.d_positions[i * 3 + 1] = gpos[1]; out.d_positions[i * 3 + 2] = gpos[2];
        out.d_scales[i * 3] = gsc[0]; out.d_scales[i * 3 + 1] = gsc[1]; out.d_scales[i * 3 + 2] = gsc[2];
        reinterpret_cast<float4*>(out.d_rotations)[i] = grot;
        out.d_opacities[i] = gop;
    }
}

}  // namespace gsd

using namespace gsd;

// ---- launchers (called from capi.cpp) ----
hipError_t launch_preprocess_fwd(const gs_gaussians& G, const CamArgs& c, const gs_preprocessed& out, int deg,
                                 hipStream_t s) {
    dim3 grid((G.count + 255) / 256), block(256);
    switch (G.max_coeffs) {
        case 1: preprocess_fwd_kernel<1><<<grid, block, 0, s>>>(G, c, out, deg); break;
        case 4: preprocess_fwd_kernel<4><<<grid, block, 0, s>>>(G, c, out, deg); break;
        case 9: preprocess_fwd_kernel<9><<<grid, block, 0, s>>>(G, c, out, deg); break;
        case 16: preprocess_fwd_kernel<16><<<grid, block, 0, s>>>(G, c, out, deg); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_activate(const gs_gaussians& G, float* pos, float* sc, float* rot, float* sh, float* op,
                           hipStream_t s) {
    activate_kernel<<<(G.count + 255) / 256, 256, 0, s>>>(G, pos, sc, rot, sh, op);
    return hipGetLastError();
}

hipError_t launch_preprocess_bwd(const gs_gaussians& G, const CamArgs& c, const int* radii, const gs_raster_grads& rg,
                                 const gs_param_grads& out, int accumulate, int sh_deg, hipStream_t s) {
    dim3 grid((G.count + 255) / 256), block(256);
    switch (G.max_coeffs) {
        case 1: preprocess_bwd_kernel<1><<<grid, block, 0, s>>>(G, c, radii, rg, out, accumulate, sh_deg); break;
        case 4: preprocess_bwd_kernel<4><<<grid, block, 0, s>>>(G, c, radii, rg, out, accumulate, sh_deg); break;
        case 9: preprocess_bwd_kernel<9><<<grid, block, 0, s>>>(G, c, radii, rg, out, accumulate, sh_deg); break;
        case 16: preprocess_bwd_kernel<16><<<grid, block, 0, s>>>(G, c, radii, rg, out, accumulate, sh_deg); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
