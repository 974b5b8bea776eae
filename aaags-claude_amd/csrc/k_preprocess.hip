// k_preprocess.hip -- per-Gaussian kernels: activation, preprocess forward and
// preprocess backward (the full adjoint of the projection, SH and activation
// chain), single-view and multi-view.
//
// One thread per Gaussian, 256-thread workgroups.  The kernels are HBM-bound:
// each Gaussian's raw parameters (56 B + 12 B per SH basis) are read once; the
// SH block is read with 16-byte loads when the model has 4 or 16 bases.
// Forward quantities the backward needs (cov3D, view point, clip w, filtered
// cov2D, det) are recomputed from the parameters rather than stored, which
// trades ~100 FLOPs for 56 B/Gaussian of HBM traffic each way.  The
// multi-view kernels read a Gaussian's parameters once and loop over the
// views; the multi-view backward sums the views' gradients in registers and
// can apply Adam in place, so gradients never reach HBM.
#include "gs_pre_device.h"

namespace gsd {

// Each thread loads its own SH row (load_sh: 16-byte loads when the row is
// whole float4s) before its geometry, so the loads are in flight while it is
// computed.  Without LDS staging the kernel holds 78 VGPRs and 6 waves per
// SIMD; staging the workgroup's rows through LDS as one coalesced run was
// measured slower (0.108 against 0.105 ms with cold caches, 0.061 against
// 0.055 ms warm, at 1M Gaussians), its 50 KB of LDS leaving 3 waves per SIMD.
template <int MAXC, int DEG>
__global__ __launch_bounds__(256) void preprocess_fwd_kernel(gs_gaussians G, CamArgs c, gs_preprocessed out) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.count) return;
    float sh[NB * 3];
    load_sh<MAXC, NB>(G.sh_coeffs, i, sh);
    Params P;
    Geom g;
    load_params(G, i, G.raw != 0, P);
    const bool vis = preprocess_geom(P.p, P.s, P.q, c, g);
    float2 m2 = make_float2(0.0f, 0.0f);
    float con[3] = {0.0f, 0.0f, 0.0f}, rgb[3] = {0.0f, 0.0f, 0.0f};
    float depth = 0.0f, op2d = 0.0f;
    int radius = 0, tiles = 0;
    if (vis) {
        sh_color_from<DEG>(sh, sh_dir(P.p, c.campos), c.training != 0, rgb);
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

// Stores (or adds) a Gaussian's MAXC*3 SH gradient values.
template <int MAXC>
__device__ __forceinline__ void store_sh(float* dst_all, int i, const float* v, bool accumulate) {
    float* dst = dst_all + (size_t)i * (MAXC * 3);
    if constexpr ((MAXC * 3) % 4 == 0) {
        float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
        for (int k = 0; k < MAXC * 3 / 4; k++) {
            float4 x = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
            if (accumulate) {
                float4 o = d4[k];
                x = make_float4(o.x + x.x, o.y + x.y, o.z + x.z, o.w + x.w);
            }
            d4[k] = x;
        }
    } else {
#pragma unroll
        for (int k = 0; k < MAXC * 3; k++) dst[k] = accumulate ? dst[k] + v[k] : v[k];
    }
}

// ---- preprocessor_backward_cpu.cpp:205-645, one view ----
template <int MAXC, int DEG>
__global__ __launch_bounds__(256) void preprocess_bwd_kernel(gs_gaussians G, CamArgs c, const int* __restrict__ radii,
                                                             gs_raster_grads rg, gs_param_grads out, int accumulate) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.count) return;
    const bool acc = accumulate != 0;
    if (radii[i] <= 0) {
        if (!acc) {
            out.d_positions[i * 3] = 0.0f; out.d_positions[i * 3 + 1] = 0.0f; out.d_positions[i * 3 + 2] = 0.0f;
            out.d_scales[i * 3] = 0.0f; out.d_scales[i * 3 + 1] = 0.0f; out.d_scales[i * 3 + 2] = 0.0f;
            reinterpret_cast<float4*>(out.d_rotations)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float z[MAXC * 3];
#pragma unroll
            for (int k = 0; k < MAXC * 3; k++) z[k] = 0.0f;
            store_sh<MAXC>(out.d_sh_coeffs, i, z, false);
            out.d_opacities[i] = 0.0f;
        }
        return;
    }
    Params P;
    load_params(G, i, true, P);
    Geom g;
    preprocess_geom(P.p, P.s, P.q, c, g);
    float sh[NB * 3];
    load_sh<MAXC, NB>(G.sh_coeffs, i, sh);
    float gpos[3] = {0.0f, 0.0f, 0.0f};
    float gsh[MAXC * 3];
    const float drgb[3] = {rg.d_rgb[i * 3], rg.d_rgb[i * 3 + 1], rg.d_rgb[i * 3 + 2]};
    sh_backward_from<MAXC, DEG>(sh, sh_dir(P.p, c.campos), drgb, gsh, gpos);
    RGrad r{rg.d_means2D[i * 2], rg.d_means2D[i * 2 + 1], rg.d_conics[i * 3], rg.d_conics[i * 3 + 1],
            rg.d_conics[i * 3 + 2], rg.d_opacities_2d[i]};
    float gsc[3], grot[4], gop;
    geom_backward(P, g, c, r, gpos, gsc, grot, gop);
    store_sh<MAXC>(out.d_sh_coeffs, i, gsh, acc);
    if (acc) {
        out.d_positions[i * 3] += gpos[0]; out.d_positions[i * 3 + 1] += gpos[1]; out.d_positions[i * 3 + 2] += gpos[2];
        out.d_scales[i * 3] += gsc[0]; out.d_scales[i * 3 + 1] += gsc[1]; out.d_scales[i * 3 + 2] += gsc[2];
        float4 r0 = reinterpret_cast<float4*>(out.d_rotations)[i];
        reinterpret_cast<float4*>(out.d_rotations)[i] =
            make_float4(r0.x + grot[0], r0.y + grot[1], r0.z + grot[2], r0.w + grot[3]);
        out.d_opacities[i] += gop;
    } else {
        out.d_positions[i * 3] = gpos[0]; out.d_positions[i * 3 + 1] = gpos[1]; out.d_positions[i * 3 + 2] = gpos[2];
        out.d_scales[i * 3] = gsc[0]; out.d_scales[i * 3 + 1] = gsc[1]; out.d_scales[i * 3 + 2] = gsc[2];
        reinterpret_cast<float4*>(out.d_rotations)[i] = make_float4(grot[0], grot[1], grot[2], grot[3]);
        out.d_opacities[i] = gop;
    }
}

// ------------------------------------------------------------ multi-view ----
// Forward over V views: parameters and SH read once, per-view outputs written
// view-major ([v][N]), plus the depth-sort keys view << 32 | depth bits
// (culled: view << 32 | 0xFFFFFFFF) and Gaussian indices; rects (optional,
// views of at most 255 x 255 tiles): each Gaussian's packed tile rectangle
// x0 | x1 << 8 | y0 << 16 | y1 << 24 (0 when culled), for the binning.
template <int MAXC, int DEG>
__global__ __launch_bounds__(256) void preprocess_fwd_multi_kernel(gs_gaussians G, const CamArgs* __restrict__ cams,
                                                                   int V, gs_preprocessed out,
                                                                   uint64_t* __restrict__ dkeys,
                                                                   uint32_t* __restrict__ didx,
                                                                   uint32_t* __restrict__ rects) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = G.count;
    if (i >= N) return;
    Params P;
    load_params(G, i, G.raw != 0, P);
    float sh[NB * 3];
    load_sh<MAXC, NB>(G.sh_coeffs, i, sh);
    for (int v = 0; v < V; v++) {
        const CamArgs& c = cams[v];
        const size_t o = (size_t)v * N + i;
        Geom g;
        bool vis = preprocess_geom(P.p, P.s, P.q, c, g);
        float2 m2 = make_float2(0.0f, 0.0f);
        float con[3] = {0.0f, 0.0f, 0.0f}, rgb[3] = {0.0f, 0.0f, 0.0f};
        float depth = 0.0f, op2d = 0.0f;
        int radius = 0, tiles = 0;
        if (vis) {
            sh_color_from<DEG>(sh, sh_dir(P.p, c.campos), c.training != 0, rgb);
            m2 = make_float2(g.px, g.py);
            con[0] = g.conic[0]; con[1] = g.conic[1]; con[2] = g.conic[2];
            depth = g.pv[2];
            op2d = P.op * g.hconv;
            radius = g.radius;
            tiles = g.tiles;
        }
        reinterpret_cast<float2*>(out.means2D)[o] = m2;
        out.depths[o] = depth;
        out.conics[o * 3] = con[0]; out.conics[o * 3 + 1] = con[1]; out.conics[o * 3 + 2] = con[2];
        out.opacities[o] = op2d;
        out.rgb[o * 3] = rgb[0]; out.rgb[o * 3 + 1] = rgb[1]; out.rgb[o * 3 + 2] = rgb[2];
        out.radii[o] = radius;
        out.tiles_touched[o] = tiles;
        dkeys[o] = ((uint64_t)v << 32) | (vis ? (uint64_t)__float_as_uint(depth) : 0xFFFFFFFFull);
        didx[o] = (uint32_t)i;
        if (rects) {
            uint32_t rec = 0u;
            if (radius > 0) {
                int mn[2], mx[2];
                get_rect(m2.x, m2.y, radius, c.gx, c.gy, mn, mx);
                rec = (uint32_t)mn[0] | ((uint32_t)mx[0] << 8) | ((uint32_t)mn[1] << 16) | ((uint32_t)mx[1] << 24);
            }
            rects[o] = rec;
        }
    }
}

// Adam state for the fused update.
struct AdamState {
    float* p[5];
    float* m[5];
    float* v[5];
    const gso::AdamDev* hp;  // hyper-parameters in device memory (graph replays), or
    gso::AdamDev h0;         // by value when hp is null
    const uint32_t* skip;    // when set and nonzero (a capacity overflow earlier in the
                             // step), the launch returns without touching any state
};

// One Adam element (AdamOptimizer::adam_update, optimizer.cpp:60-101) on
// values already loaded; returns the new parameter and updates *m, *v.
__device__ __forceinline__ float adam_math(float p, float* m, float* v, float g, float lr, const gso::AdamDev& h) {
    float mm = h.b1 * *m + (1.0f - h.b1) * g;
    float vv = h.b2 * *v + (1.0f - h.b2) * g * g;
    *m = mm;
    *v = vv;
    float mh = mm * h.bc1;
    float vh = vv * h.bc2;
    return p - lr * mh / (sqrtf(vh) + h.eps);
}

// Adam on K consecutive elements starting at k0: every load is issued before
// any store (the arrays are distinct), so the K updates overlap in memory.
template <int K>
__device__ __forceinline__ void adam_run(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                         size_t k0, const float* g, float lr, const gso::AdamDev& h) {
    float pp[K], mm[K], vv[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        pp[k] = p[k0 + k];
        mm[k] = __builtin_nontemporal_load(m + k0 + k);
        vv[k] = __builtin_nontemporal_load(v + k0 + k);
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        pp[k] = adam_math(pp[k], &mm[k], &vv[k], g[k], lr, h);
        p[k0 + k] = pp[k];
        __builtin_nontemporal_store(mm[k], m + k0 + k);
        __builtin_nontemporal_store(vv[k], v + k0 + k);
    }
}

// Backward over V views, the views' gradients summed in view order in
// registers.  FUSED_ADAM applies the optimizer in place (parameters, m, v);
// otherwise the summed gradients are written (overwriting `out`).
//
// With FUSED_ADAM the SH part of the update (MAXC*3 floats per Gaussian, most
// of the parameter bytes) runs cooperatively: each wave stores its 64
// Gaussians' SH gradients in LDS in the global layout (rows padded by one
// float so that the per-thread stores hit distinct banks), then sweeps the
// wave's contiguous SH block with one float per lane per step, so every load
// and store of p, m and v is coalesced.  The few per-Gaussian scalars
// (position, scale, rotation, opacity) are updated by their own thread.
template <int MAXC, int DEG, bool FUSED_ADAM>
__global__ __launch_bounds__(256, 2) void preprocess_bwd_multi_kernel(gs_gaussians G, const CamArgs* __restrict__ cams,
                                                                   CamArgs cam0, int V, const int* __restrict__ radii,
                                                                   gs_raster_grads rg, gs_param_grads out,
                                                                   AdamState ad) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    constexpr int E = MAXC * 3;  // SH floats per Gaussian
    constexpr int ROW = E + 1;   // padded LDS row
    // per wave: 64 parameter rows (the SH read once, by the owner thread),
    // then 32 gradient rows for the half of the wave's Gaussians being swept
    constexpr int WROWS = 96;
    constexpr bool KEEP_P = FUSED_ADAM && (NB == MAXC);
    __shared__ float s_sh[FUSED_ADAM ? 4 * WROWS * ROW : 1];
    if (FUSED_ADAM && ad.skip && *ad.skip) return;  // uniform over the grid
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = G.count;
    const bool valid = i < N;
    float tsh[E];
#pragma unroll
    for (int k = 0; k < E; k++) tsh[k] = 0.0f;
    if (valid) {
        Params P;
        load_params(G, i, true, P);
        float sh[NB * 3];
        load_sh<MAXC, NB>(G.sh_coeffs, i, sh);
        if constexpr (KEEP_P) {
            float* prow = s_sh + ((threadIdx.x >> 6) * WROWS + (threadIdx.x & 63)) * ROW;
#pragma unroll
            for (int k = 0; k < E; k++) prow[k] = sh[k];
        }
        float tpos[3] = {0.0f, 0.0f, 0.0f}, tsc[3] = {0.0f, 0.0f, 0.0f}, trot[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float top = 0.0f;
        // one view's gradients, summed into the totals
        auto view = [&](const CamArgs& c, size_t o) {
            Geom g;
            preprocess_geom(P.p, P.s, P.q, c, g);
            float gpos[3] = {0.0f, 0.0f, 0.0f};
            float gsh[E];
            // the screen-space sums are read once: non-temporal, so that they
            // do not push the parameters out of the caches
            const float drgb[3] = {__builtin_nontemporal_load(rg.d_rgb + o * 3),
                                   __builtin_nontemporal_load(rg.d_rgb + o * 3 + 1),
                                   __builtin_nontemporal_load(rg.d_rgb + o * 3 + 2)};
            sh_backward_from<MAXC, DEG>(sh, sh_dir(P.p, c.campos), drgb, gsh, gpos);
            RGrad r{__builtin_nontemporal_load(rg.d_means2D + o * 2), __builtin_nontemporal_load(rg.d_means2D + o * 2 + 1),
                    __builtin_nontemporal_load(rg.d_conics + o * 3), __builtin_nontemporal_load(rg.d_conics + o * 3 + 1),
                    __builtin_nontemporal_load(rg.d_conics + o * 3 + 2),
                    __builtin_nontemporal_load(rg.d_opacities_2d + o)};
            float gsc[3], grot[4], gop;
            geom_backward(P, g, c, r, gpos, gsc, grot, gop);
#pragma unroll
            for (int k = 0; k < 3; k++) { tpos[k] += gpos[k]; tsc[k] += gsc[k]; }
#pragma unroll
            for (int k = 0; k < 4; k++) trot[k] += grot[k];
            top += gop;
#pragma unroll
            for (int k = 0; k < E; k++) tsh[k] += gsh[k];
        };
        if (cams) {
            for (int v = 0; v < V; v++) {
                const size_t o = (size_t)v * N + i;
                if (radii[o] > 0) view(cams[v], o);
            }
        } else if (radii[i] > 0) {
            view(cam0, (size_t)i);  // one view, camera by value
        }
        if constexpr (FUSED_ADAM) {
            const gso::AdamDev h = ad.hp ? *ad.hp : ad.h0;
            adam_run<3>(ad.p[0], ad.m[0], ad.v[0], (size_t)i * 3, tpos, h.lr[0], h);
            adam_run<3>(ad.p[1], ad.m[1], ad.v[1], (size_t)i * 3, tsc, h.lr[1], h);
            adam_run<4>(ad.p[2], ad.m[2], ad.v[2], (size_t)i * 4, trot, h.lr[2], h);
            adam_run<1>(ad.p[4], ad.m[4], ad.v[4], (size_t)i, &top, h.lr[4], h);
        } else {
            out.d_positions[i * 3] = tpos[0]; out.d_positions[i * 3 + 1] = tpos[1];
            out.d_positions[i * 3 + 2] = tpos[2];
            out.d_scales[i * 3] = tsc[0]; out.d_scales[i * 3 + 1] = tsc[1]; out.d_scales[i * 3 + 2] = tsc[2];
            reinterpret_cast<float4*>(out.d_rotations)[i] = make_float4(trot[0], trot[1], trot[2], trot[3]);
            store_sh<MAXC>(out.d_sh_coeffs, i, tsh, false);
            out.d_opacities[i] = top;
        }
    }
    if constexpr (KEEP_P) {
        // two sweeps of 32 whole rows each (contiguous in memory): the
        // parameters come from the rows the owners wrote, the gradients from
        // the 32 gradient rows, which each half's owners fill in turn
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        float* wrows = s_sh + w * WROWS * ROW;
        const int g0 = blockIdx.x * blockDim.x + w * 64;  // first Gaussian of this wave
        const gso::AdamDev h = ad.hp ? *ad.hp : ad.h0;
#pragma unroll 1
        for (int half = 0; half < 2; half++) {
            if ((lane >> 5) == half) {
                float* grow = wrows + (64 + (lane & 31)) * ROW;
#pragma unroll
                for (int k = 0; k < E; k++) grow[k] = tsh[k];
            }
            __syncthreads();
            const int gh = g0 + 32 * half;
            const int cnt = N - gh < 32 ? N - gh : 32;
            if (cnt > 0) {
                float* __restrict__ p = ad.p[3] + (size_t)gh * E;
                float* __restrict__ m = ad.m[3] + (size_t)gh * E;
                float* __restrict__ v = ad.v[3] + (size_t)gh * E;
                const float* prows = wrows + 32 * half * ROW;
                const float* grows = wrows + 64 * ROW;
                const int total = cnt * E;
                constexpr int U = 8;  // elements per lane in flight
                for (int j0 = lane; j0 < total; j0 += 64 * U) {
                    float pp[U], mm[U], vv[U], gg[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int j = j0 + 64 * u;
                        if (j < total) {
                            const int gl = j / E, k = j - gl * E;
                            pp[u] = prows[gl * ROW + k];
                            gg[u] = grows[gl * ROW + k];
                            mm[u] = __builtin_nontemporal_load(m + j);
                            vv[u] = __builtin_nontemporal_load(v + j);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int j = j0 + 64 * u;
                        if (j < total) {
                            pp[u] = adam_math(pp[u], &mm[u], &vv[u], gg[u], h.lr[3], h);
                            p[j] = pp[u];
                            __builtin_nontemporal_store(mm[u], m + j);
                            __builtin_nontemporal_store(vv[u], v + j);
                        }
                    }
                }
            }
            __syncthreads();
        }
    } else if constexpr (FUSED_ADAM) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        float* row = s_sh + (w * WROWS + lane) * ROW;
#pragma unroll
        for (int k = 0; k < E; k++) row[k] = tsh[k];
        __syncthreads();
        const int g0 = blockIdx.x * blockDim.x + w * 64;  // first Gaussian of this wave
        const int cnt = N - g0 < 64 ? N - g0 : 64;
        if (cnt > 0) {
            const gso::AdamDev h = ad.hp ? *ad.hp : ad.h0;
            float* __restrict__ p = ad.p[3] + (size_t)g0 * E;
            float* __restrict__ m = ad.m[3] + (size_t)g0 * E;
            float* __restrict__ v = ad.v[3] + (size_t)g0 * E;
            const float* wsh = s_sh + w * WROWS * ROW;
            const int total = cnt * E;
            constexpr int U = 8;  // elements per lane in flight
            for (int j0 = lane; j0 < total; j0 += 64 * U) {
                float pp[U], mm[U], vv[U], gg[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int j = j0 + 64 * u;
                    if (j < total) {
                        const int gl = j / E;
                        gg[u] = wsh[gl * ROW + (j - gl * E)];
                        pp[u] = p[j];
                        mm[u] = m[j];
                        vv[u] = v[j];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int j = j0 + 64 * u;
                    if (j < total) {
                        pp[u] = adam_math(pp[u], &mm[u], &vv[u], gg[u], h.lr[3], h);
                        p[j] = pp[u];
                        m[j] = mm[u];
                        v[j] = vv[u];
                    }
                }
            }
        }
    }
}

}  // namespace gsd

using namespace gsd;

// Instantiates KERNEL<MAXC, DEG, ...> for the supported (coefficients, degree)
// pairs, DEG <= the degree MAXC holds.
#define GS_DISPATCH_SH(MAXC_, DEG_, CALL)                                        \
    switch (MAXC_) {                                                             \
        case 1: { constexpr int MC = 1; constexpr int DG = 0; CALL; } break;     \
        case 4:                                                                  \
            if (DEG_ <= 0) { constexpr int MC = 4; constexpr int DG = 0; CALL; } \
            else { constexpr int MC = 4; constexpr int DG = 1; CALL; }           \
            break;                                                               \
        case 9:                                                                  \
            if (DEG_ <= 0) { constexpr int MC = 9; constexpr int DG = 0; CALL; } \
            else if (DEG_ == 1) { constexpr int MC = 9; constexpr int DG = 1; CALL; } \
            else { constexpr int MC = 9; constexpr int DG = 2; CALL; }           \
            break;                                                               \
        case 16:                                                                 \
            if (DEG_ <= 0) { constexpr int MC = 16; constexpr int DG = 0; CALL; } \
            else if (DEG_ == 1) { constexpr int MC = 16; constexpr int DG = 1; CALL; } \
            else if (DEG_ == 2) { constexpr int MC = 16; constexpr int DG = 2; CALL; } \
            else { constexpr int MC = 16; constexpr int DG = 3; CALL; }          \
            break;                                                               \
        default: return hipErrorInvalidValue;                                    \
    }

// ---- launchers (called from capi.hip / the drivers) ----
hipError_t launch_preprocess_fwd(const gs_gaussians& G, const CamArgs& c, const gs_preprocessed& out, int deg,
                                 hipStream_t s) {
    if (G.count <= 0) return hipSuccess;  // no empty grids
    dim3 grid((G.count + 255) / 256), block(256);
    GS_DISPATCH_SH(G.max_coeffs, deg, (preprocess_fwd_kernel<MC, DG><<<grid, block, 0, s>>>(G, c, out)));
    return hipGetLastError();
}

hipError_t launch_activate(const gs_gaussians& G, float* pos, float* sc, float* rot, float* sh, float* op,
                           hipStream_t s) {
    if (G.count <= 0) return hipSuccess;
    activate_kernel<<<(G.count + 255) / 256, 256, 0, s>>>(G, pos, sc, rot, sh, op);
    return hipGetLastError();
}

hipError_t launch_preprocess_bwd(const gs_gaussians& G, const CamArgs& c, const int* radii, const gs_raster_grads& rg,
                                 const gs_param_grads& out, int accumulate, int sh_deg, hipStream_t s) {
    if (G.count <= 0) return hipSuccess;
    dim3 grid((G.count + 255) / 256), block(256);
    GS_DISPATCH_SH(G.max_coeffs, sh_deg,
                   (preprocess_bwd_kernel<MC, DG><<<grid, block, 0, s>>>(G, c, radii, rg, out, accumulate)));
    return hipGetLastError();
}

hipError_t launch_preprocess_fwd_multi(const gs_gaussians& G, const CamArgs* cams_dev, int V,
                                       const gs_preprocessed& out, uint64_t* dkeys, uint32_t* didx, int deg,
                                       hipStream_t s, uint32_t* rects) {
    if (G.count <= 0) return hipSuccess;
    dim3 grid((G.count + 255) / 256), block(256);
    GS_DISPATCH_SH(G.max_coeffs, deg,
                   (preprocess_fwd_multi_kernel<MC, DG>
                    <<<grid, block, 0, s>>>(G, cams_dev, V, out, dkeys, didx, rects)));
    return hipGetLastError();
}

// adam_p/m/v with hp (device) or hp_val (host, passed by value): the summed
// gradients update the parameters in place (G's arrays must be the same as
// adam_p) and `out` is unused.  cams_dev null: one view, cam_val by value.
hipError_t launch_preprocess_bwd_multi(const gs_gaussians& G, const CamArgs* cams_dev, const CamArgs* cam_val, int V,
                                       const int* radii, const gs_raster_grads& rg, const gs_param_grads& out,
                                       float* const adam_p[5], float* const adam_m[5], float* const adam_v[5],
                                       const gso::AdamDev* hp, const gso::AdamDev* hp_val, int sh_deg, hipStream_t s,
                                       const uint32_t* skip) {
    if (G.count <= 0) return hipSuccess;
    dim3 grid((G.count + 255) / 256), block(256);
    CamArgs cam0{};
    if (!cams_dev) {
        if (!cam_val || V != 1) return hipErrorInvalidValue;
        cam0 = *cam_val;
    }
    AdamState ad{};
    if (hp || hp_val) {
        for (int k = 0; k < 5; k++) {
            ad.p[k] = adam_p[k];
            ad.m[k] = adam_m[k];
            ad.v[k] = adam_v[k];
        }
        ad.hp = hp;
        ad.skip = skip;
        if (!hp) ad.h0 = *hp_val;
        GS_DISPATCH_SH(G.max_coeffs, sh_deg,
                       (preprocess_bwd_multi_kernel<MC, DG, true><<<grid, block, 0, s>>>(G, cams_dev, cam0, V, radii, rg,
                                                                                        out, ad)));
    } else {
        GS_DISPATCH_SH(G.max_coeffs, sh_deg,
                       (preprocess_bwd_multi_kernel<MC, DG, false><<<grid, block, 0, s>>>(G, cams_dev, cam0, V, radii,
                                                                                         rg, out, ad)));
    }
    return hipGetLastError();
}
