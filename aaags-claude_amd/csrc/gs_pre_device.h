// gs_pre_device.h -- per-Gaussian device math shared by the single-view and
// multi-view kernels: parameter loading/activation, the 2D projection
// (preprocessor_cpu.cpp:188-236), SH colour (sh_eval.cpp:24-89) and the full
// adjoint of both (preprocessor_backward_cpu.cpp:8-645), each in the
// reference's operation order.
#pragma once
#include "gs_device.h"

namespace gsd {

struct Geom {
    float pv[3];   // view-space position
    float phw;     // clip-space w
    float px, py;  // pixel position
    float c3[6];   // 3D covariance
    float c2[3];   // 2D covariance after the +0.3 filter
    float det;
    float conic[3];
    float hconv;
    float det_cov, c2a, c2c;  // determinant and diagonal before the filter (antialiasing backward)
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
    proj_jacobian(g.pv, c.view, c.tanx, c.tany, c.fx, c.fy, pj);
    cov2d_from(pj, g.c3, g.c2);
    float det_cov = g.c2[0] * g.c2[2] - g.c2[1] * g.c2[1];
    g.det_cov = det_cov;
    g.c2a = g.c2[0];
    g.c2c = g.c2[2];
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

// First NB bases (3 floats each) of Gaussian i's interleaved SH block, with
// 16-byte loads when the block is 16-byte aligned.
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

// sh_eval.cpp:79-88: +0.5, clamp at 0, and at 1 unless training.
template <int DEG>
__device__ __forceinline__ void sh_color_from(const float* sh, const ShDir& d, bool training, float rgb[3]) {
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float v = sh_raw_channel<DEG>(d, sh + ch);
        v = smax(0.0f, v + 0.5f);
        if (!training) v = smin(1.0f, v);
        rgb[ch] = v;
    }
}

// preprocessor_backward_cpu.cpp:8-203: gsh[0 .. MAXC*3) is written (bases
// above DEG get 0); the SH term of the position gradient is added to dpos.
template <int MAXC, int DEG>
__device__ __forceinline__ void sh_backward_from(const float* sh, const ShDir& d, const float drgb[3], float* gsh,
                                                 float dpos[3]) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
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
#pragma unroll
    for (int k = 0; k < MAXC; k++)
#pragma unroll
        for (int ch = 0; ch < 3; ch++) gsh[k * 3 + ch] = (k < NB) ? b[k < NB ? k : 0] * de[ch] : 0.0f;
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

// Screen-space gradients of one Gaussian in one view.
struct RGrad {
    float dm0, dm1, dc0, dc1, dc2, dop;
};

// preprocessor_backward_cpu.cpp:230-644 after the SH chain: opacity (sigmoid),
// conic -> cov2D -> cov3D -> scale, rotation (with the normalisation), the
// cov2D -> J -> view point path and the projection path of the position.
// gpos is added to; gsc, grot and gop are written.
__device__ __forceinline__ void geom_backward(const Params& P, const Geom& g, const CamArgs& c, const RGrad& r,
                                              float gpos[3], float gsc[3], float grot[4], float& gop) {
    const float sig = P.op;
    // antialiasing (preprocessor_cpu.cpp:213-216,250): op2d = op * h with
    // h = sqrt(max(2.5e-5, det_cov / det)); the reference's backward has no
    // term for h, this one carries it into the opacity and, through det_cov
    // and det, into the 2D covariance (h = 1 without antialiasing)
    gop = r.dop * g.hconv * sig * (1.0f - sig);

    const float a = g.c2[0], b = g.c2[1], cc = g.c2[2];
    const float inv_det = 1.0f / g.det;
    const float id2 = inv_det * inv_det;
    float d_a = r.dc0 * (-cc * cc * id2) + r.dc1 * (b * cc * id2) + r.dc2 * (-b * b * id2);
    float d_b = r.dc0 * (2.0f * b * cc * id2) + r.dc1 * (-(a * cc + b * b) * id2) + r.dc2 * (2.0f * a * b * id2);
    float d_c = r.dc0 * (-b * b * id2) + r.dc1 * (a * b * id2) + r.dc2 * (-a * a * id2);
    if (c.antialiasing) {
        const float ratio = g.det_cov / g.det;
        if (ratio > 0.000025f) {
            const float dL_dratio = r.dop * sig * 0.5f / g.hconv;
            const float dL_dcovdet = dL_dratio / g.det;
            const float dL_ddet = -dL_dratio * ratio / g.det;
            d_a += dL_dcovdet * g.c2c + dL_ddet * cc;
            d_c += dL_dcovdet * g.c2a + dL_ddet * a;
            d_b += (dL_dcovdet + dL_ddet) * (-2.0f * b);
        }
    }

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
    grot[0] = (dq0 - rq * dotq) * il;
    grot[1] = (dq1 - xq * dotq) * il;
    grot[2] = (dq2 - yq * dotq) * il;
    grot[3] = (dq3 - zq * dotq) * il;

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
    const float dnx = r.dm0 * (float)c.W * 0.5f;
    const float dny = r.dm1 * (float)c.H * 0.5f;
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
}

}  // namespace gsd
