// k_raster.hip -- tile rasterization, forward and backward.
//
// One 256-thread workgroup (four wave64s) per 16x16 tile, one thread per
// pixel.  The tile's depth-sorted Gaussians are staged in LDS in batches of
// BATCH entries (position, conic+opacity, colour); every lane then reads the
// same LDS slot (a broadcast, no bank conflicts).  The forward stops a pixel
// when its transmittance would fall below 1e-4 and the workgroup when all its
// pixels stopped; it records T_final, the contributor count and the list index
// (+1) of the last contributor.
//
// The backward walks the same list back to front from the largest last
// contributor in the tile.  Each pixel recovers T_i = T/(1-alpha) exactly as
// the reference CPU backward (cpu/rasterizer_backward_cpu.cpp:69-134); the
// nine per-Gaussian gradient terms are summed across the wave with packed
// permlane-swap/DPP steps (no LDS traffic), across the four waves with LDS
// atomics, and land in HBM with one float atomic per (tile, Gaussian,
// component).
//
// Multi-view: the same tile bodies run over a global (view, tile) list, either
// one workgroup per global tile or a persistent grid that takes tiles from
// per-XCD atomic queues.  Workgroup-to-tile maps are XCD-aware throughout.
#include "gs_device.h"

namespace gsr {

using gsd::smin;

// Wave w of a tile (threads 64w .. 64w+63) covers one block of its pixels:
// in the forward the 8 x 8 quadrant (w & 1, w >> 1), in the backward the
// 16 x 4 strip of rows 4w .. 4w+3 (measured: each layout is the faster one
// for its kernel).  Per LDS batch, mask[w][c] bit b is set when entry 64c + b
// can pass the alpha cut somewhere in block w (see strip_hits); each wave
// visits only its set bits.
// One staged entry: 48 bytes, read with one base address and immediate
// offsets (mean, conic + opacity, colour).
struct Staged {
    float2 xy;
    float2 pad;
    float4 co;
    float4 rgb;
};

template <int BATCH>
struct FwdSmem {
    Staged e[BATCH];
    uint64_t mask[4][BATCH / 64];
};

// Block-hit bits of one staged entry: bit w set when the Gaussian can pass the
// 1/255 alpha cut at some pixel centre of block w of the tile whose top-left
// pixel is (x0, y0): QUAD, the 8 x 8 quadrant (w & 1, w >> 1); else the strip
// of rows y0 + 4w .. y0 + 4w + 3, columns x0 .. x0 + 15.
// alpha = o * exp(-Q/2) with Q = a u^2 + 2 b u v + c v^2 ((a, b, c) the conic,
// (u, v) the pixel minus the mean), so passing needs Q <= r2 = 2 ln(255 o).
// The least Q over a rectangle is 0 when the mean lies inside it, else the
// least of its minima along the four edges; along an edge Q is a quadratic
// in one variable whose minimum lies at its vertex clamped to the edge.
// Margins, so that a strip without the bit has no pixel that passes:
//  - the rasterizer's float Q has a relative rounding error below about
//    8 eps tr^2/det (its terms reach tr^2/det times Q), and so has the Q
//    computed here; the test is exact only where tr^2/det <= 1000 (together
//    at most 1e-3) and r2 is widened by 0.1 %, plus 1e-3 for the approximate
//    log used here;
//  - each rectangle is grown by half a pixel on every side.
// o < 1/255 never passes (no bits); a degenerate or too anisotropic conic
// gets every bit.  An entry with a non-finite mean or conic never passes in
// the rasterizer (its power is NaN) and gets no bits, so every visited entry
// is finite.
__device__ __forceinline__ float quad_form(float a, float b, float c, float u, float v) {
    return (a * u * u + c * v * v) + 2.0f * b * u * v;
}

template <bool QUAD>
__device__ __forceinline__ uint32_t strip_hits(float2 m, float4 co, float x0, float y0) {
    const float o = co.w;
    if (!(o >= 1.0f / 255.0f)) return 0u;
    const float a = co.x, b = co.y, c = co.z;
    if (!(__builtin_isfinite(m.x) && __builtin_isfinite(m.y) && __builtin_isfinite(a) && __builtin_isfinite(b) &&
          __builtin_isfinite(c) && __builtin_isfinite(o)))
        return 0u;
    const float det = a * c - b * b;
    const float tr = a + c;
    if (!(det > 0.0f) || !(tr * tr <= 1000.0f * det)) return 0xFu;
    // 2 ln(x) = 2 ln(2) log2(x)
    const float r2 = 1.38629436f * __builtin_amdgcn_logf(255.0f * o) * 1.001f + 1e-3f;
    constexpr float G = 0.5f;
    constexpr float BW = QUAD ? 7.0f : 15.0f, BH = QUAD ? 7.0f : 3.0f;  // block extent - 1
    // edge vertices: u* = -(b / a) v on the rows, v* = -(b / c) u on the columns
    const float ka = -b * __builtin_amdgcn_rcpf(a), kc = -b * __builtin_amdgcn_rcpf(c);
    uint32_t h = 0u;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const float bx = x0 + (QUAD ? (float)(8 * (w & 1)) : 0.0f);
        const float by = y0 + (QUAD ? (float)(8 * (w >> 1)) : (float)(4 * w));
        const float u0 = bx - G - m.x, u1 = bx + (BW + G) - m.x;
        const float vl = kc * u0, vr = kc * u1;
        const float v0 = by - G - m.y, v1 = by + (BH + G) - m.y;
        float q = (u0 <= 0.0f && u1 >= 0.0f && v0 <= 0.0f && v1 >= 0.0f) ? 0.0f : 3.0e38f;
        q = fminf(q, quad_form(a, b, c, fminf(fmaxf(ka * v0, u0), u1), v0));
        q = fminf(q, quad_form(a, b, c, fminf(fmaxf(ka * v1, u0), u1), v1));
        q = fminf(q, quad_form(a, b, c, u0, fminf(fmaxf(vl, v0), v1)));
        q = fminf(q, quad_form(a, b, c, u1, fminf(fmaxf(vr, v0), v1)));
        h |= (q <= r2) ? (1u << w) : 0u;
    }
    return h;
}

// Staged conics are kept as (-a/2, -b, -c/2, o), so that the power
// -0.5 (a dx^2 + c dy^2) - b dx dy of rasterizer_cpu.cpp:159-161 becomes
// (a' dx^2 + c' dy^2) + b' dx dy.  Scaling by -1/2 and -1 is exact, so this is
// the same float value (away from subnormals) in one multiply fewer.
__device__ __forceinline__ float4 prescale(float4 co) { return make_float4(-0.5f * co.x, -co.y, -0.5f * co.z, co.w); }

__device__ __forceinline__ float power_of(const float4& co, float dx, float dy) {
    return (co.x * dx * dx + co.z * dy * dy) + co.y * dx * dy;
}

// Ballots the hit bits of 64 staged entries (one per lane of the calling wave,
// entries 64c .. 64c+63) into the four strips' masks.
template <int BATCH>
__device__ __forceinline__ void store_masks(uint64_t (&mask)[4][BATCH / 64], int c, uint32_t h, int lane) {
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint64_t b = __ballot((h >> w) & 1u);
        if (lane == 0) mask[w][c] = b;
    }
}

__device__ __forceinline__ uint64_t wave_uniform(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// gradient sums per Gaussian: dop * (dx, dy, dx^2, dx dy, dy^2), rgb 3, opacity 1
constexpr int NG = 9;

template <int BATCH>
struct BwdSmem {
    Staged e[BATCH];
    uint64_t mask[4][BATCH / 64];
    uint32_t id[BATCH];
    float g[BATCH][NG];
    uint32_t maxlast;
};

struct PixOut {
    float* img;
    float* T_final;
    int* n_contrib;
    uint32_t* last_contrib;
};

template <int BATCH>
__device__ __forceinline__ void fwd_tile(const RasterArgs& a, int tile, FwdSmem<BATCH>& sm, const PixOut& o,
                                         const RasterSplit& sp) {
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int t = threadIdx.x;
    // 8 x 8 quadrant (w & 1, w >> 1) per wave
    const int px = tx * gsd::TILE + 8 * ((t >> 6) & 1) + (t & 7), py = ty * gsd::TILE + 8 * (t >> 7) + ((t >> 3) & 7);
    const bool inside = px < a.W && py < a.H;
    const int pix = py * a.W + px;
    const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
    // the part of the list this pass renders: [from, to)
    uint32_t from = start, to = end;
    if (sp.round != 0 && end > start) {
        const uint32_t fill = sp.scanned[(size_t)tile * sp.C + sp.g1];  // where group g1 begins
        if (sp.round == 1) {
            to = fill < end ? fill : end;
        } else {
            if (sp.tile_flag[tile] == 0u) return;  // finished in round 1 (uniform per workgroup)
            from = fill;
        }
    } else if (sp.round == 2) {
        return;
    }
    const float fpx = (float)px, fpy = (float)py;
    const float tx0 = (float)(tx * gsd::TILE), ty0 = (float)(ty * gsd::TILE);
    const int lane = t & 63, w = t >> 6;
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    int cnt = 0;
    uint32_t last = 0;
    bool done = !inside;
    if (sp.round == 2 && inside) {  // the state round 1 left
        T = o.T_final[pix];
        const uint32_t nc = (uint32_t)o.n_contrib[pix];
        cnt = (int)(nc & 0x7FFFFFFFu);
        done = (nc >> 31) != 0u;
        last = o.last_contrib[pix];
        C0 = sp.rc[pix * 3 + 0];
        C1 = sp.rc[pix * 3 + 1];
        C2 = sp.rc[pix * 3 + 2];
    }
    for (uint32_t base = from; base < to; base += BATCH) {
        if (__syncthreads_count(!done) == 0) break;
        for (int k = t; k < BATCH; k += 256) {
            uint32_t j = base + k;
            uint32_t h = 0u;
            if (j < to) {
                uint32_t id = a.values[(size_t)j * a.vstride];
                const float2 mm = a.means2D[id];
                const float4 cc = make_float4(a.conics[id * 3], a.conics[id * 3 + 1], a.conics[id * 3 + 2], a.opac[id]);
                sm.e[k].xy = mm;
                sm.e[k].co = prescale(cc);
                sm.e[k].rgb = make_float4(a.rgb[id * 3], a.rgb[id * 3 + 1], a.rgb[id * 3 + 2], 0.0f);
                h = strip_hits<true>(mm, cc, tx0, ty0);
            }
            store_masks<BATCH>(sm.mask, k >> 6, h, lane);
        }
        __syncthreads();
        const int nc = (int)(((to - base) < (uint32_t)BATCH ? (to - base) : (uint32_t)BATCH) + 63u) >> 6;
        // predicated: every lane runs every visited entry and selects, so the
        // only branches are uniform (the CU's scalar unit, shared by its
        // waves, is not spent on exec-mask bookkeeping of per-pixel branches)
        for (int c = 0; c < nc; c++) {
            if (__ballot(!done) == 0ull) break;
            uint64_t m = wave_uniform(sm.mask[w][c]);
            while (m != 0ull) {
                const int k = (c << 6) + __builtin_ctzll(m);
                m &= m - 1ull;
                const float2 mm = sm.e[k].xy;
                const float4 co = sm.e[k].co;
                const float4 cl = sm.e[k].rgb;
                const float dx = mm.x - fpx, dy = mm.y - fpy;
                const float power = power_of(co, dx, dy);
                float ep;
                const float alpha = gsd::raster_alpha_wave(co.w, power, &ep);
                // rasterizer_cpu.cpp: skip if power > 0 or alpha < 1/255; stop
                // (without this entry) once T would fall below 1e-4
                const bool use = !done && power <= 0.0f && alpha >= 1.0f / 255.0f;
                const float test_T = T * (1.0f - alpha);
                const bool take = use && test_T >= 0.0001f;  // T and alpha are finite here
                // used but no room left (one compare: the negated compare
                // would be a second, since the compiler keeps NaN semantics)
                done = done || (use != take);
                // colour += c alpha T, with the weight alpha T selected in or out;
                // the colour feeds no decision (T and alpha do), so its sums are
                // contracted to FMAs
                const float wt = take ? alpha * T : 0.0f;
                {
#pragma clang fp contract(fast)
                    C0 += cl.x * wt;
                    C1 += cl.y * wt;
                    C2 += cl.z * wt;
                }
                T = take ? test_T : T;
                cnt += take ? 1 : 0;
                last = take ? base + (uint32_t)k + 1u : last;
            }
        }
    }
    // round 1 stopped at the end of the placed part with pixels still open:
    // flag the tile (every tile's flag is written, so nothing needs clearing
    // beforehand) and keep the state (colour in rc, the done bit in
    // n_contrib's top bit) for round 2
    bool keep = false;
    if (sp.round == 1) {
        if (to < end) keep = __syncthreads_count(!done) != 0;
        if (t == 0) sp.tile_flag[tile] = keep ? 1u : 0u;
    }
    if (inside) {
        o.img[pix * 3 + 0] = C0 + T * a.bg0;
        o.img[pix * 3 + 1] = C1 + T * a.bg1;
        o.img[pix * 3 + 2] = C2 + T * a.bg2;
        if (o.T_final) o.T_final[pix] = T;
        if (o.n_contrib) o.n_contrib[pix] = keep && done ? (int)((uint32_t)cnt | 0x80000000u) : cnt;
        if (o.last_contrib) o.last_contrib[pix] = last;
        if (keep) {
            sp.rc[pix * 3 + 0] = C0;
            sp.rc[pix * 3 + 1] = C1;
            sp.rc[pix * 3 + 2] = C2;
        }
    }
}

// ---- packed wave reduction of the nine gradient terms ----------------------
// Summing nine values over 64 lanes one by one costs 9 x 6 cross-lane steps.
// Packed, two values share each exchange: v_permlane32_swap leaves value a
// summed over lane pairs {l, l+32} in lanes 0-31 and value b in lanes 32-63;
// v_permlane16_swap does the same between 16-lane rows; DPP row rotations
// finish each row.  Afterwards row r (lanes 16r..16r+15) holds, in every lane,
//   k0 = term A[r], k1 = term 4 + A[r], with A = {0, 2, 1, 3},
// and k2 = term 8 everywhere.
__device__ __forceinline__ float swap32_sum(float a, float b) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float swap16_sum(float a, float b) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the 16 lanes of each row (row_ror:8, 4, 2, 1).
__device__ __forceinline__ float row_sum(float v) {
    v += dpp_f<0x128>(v);
    v += dpp_f<0x124>(v);
    v += dpp_f<0x122>(v);
    v += dpp_f<0x121>(v);
    return v;
}

// Which reduced sum a lane delivers: in row r (lanes 16r .. 16r+15), lane
// 16r delivers k0 (term A[r]) and lane 16r+1 k1 (term 4 + A[r]); lane 2
// delivers k2 (term 8).  Fixed per lane, so it is set up once per thread and
// the delivery below is selects, not branches.
struct ReduceRole {
    bool is0, is1, any;
    int term;
};

__device__ __forceinline__ ReduceRole reduce_role(int lane) {
    const int r = lane >> 4, p = lane & 15;
    const int a = ((r & 1) << 1) | (r >> 1);
    ReduceRole q;
    q.is0 = p == 0;
    q.is1 = p == 1;
    q.any = q.is0 || q.is1 || lane == 2;
    q.term = q.is0 ? a : (q.is1 ? 4 + a : 8);
    return q;
}

// Reduces gv[0..8] over the wave; the nine lanes with q.any set return one
// distinct term's sum each (term q.term).
__device__ __forceinline__ float wave_reduce9(const float (&gv)[NG], const ReduceRole& q) {
    const float h0 = swap32_sum(gv[0], gv[1]);
    const float h1 = swap32_sum(gv[2], gv[3]);
    const float h2 = swap32_sum(gv[4], gv[5]);
    const float h3 = swap32_sum(gv[6], gv[7]);
    const float h4 = swap32_sum(gv[8], gv[8]);
    float k0 = swap16_sum(h0, h1);
    float k1 = swap16_sum(h2, h3);
    float k2 = swap16_sum(h4, h4);
    k0 = row_sum(k0);
    k1 = row_sum(k1);
    k2 = row_sum(k2);
    return q.is0 ? k0 : (q.is1 ? k1 : k2);
}

struct PixIn {
    const float* T_final;
    const uint32_t* last_contrib;
    const float* d_image;
};

template <int BATCH>
__device__ __forceinline__ void bwd_tile(const RasterArgs& a, int tile, BwdSmem<BATCH>& sm, const PixIn& in,
                                         const GradArgs& g) {
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int px = tx * gsd::TILE + (t & 15), py = ty * gsd::TILE + (t >> 4);
    const bool inside = px < a.W && py < a.H;
    const uint32_t start = a.ranges[2 * tile];
    const int pix = py * a.W + px;
    float T = 0.0f, dC0 = 0.0f, dC1 = 0.0f, dC2 = 0.0f;
    uint32_t last = 0;
    if (inside) {
        T = in.T_final[pix];
        last = in.last_contrib[pix];
        dC0 = in.d_image[pix * 3];
        dC1 = in.d_image[pix * 3 + 1];
        dC2 = in.d_image[pix * 3 + 2];
    }
    float acc0 = a.bg0, acc1 = a.bg1, acc2 = a.bg2;
    const float fpx = (float)px, fpy = (float)py;
    const float tx0 = (float)(tx * gsd::TILE), ty0 = (float)(ty * gsd::TILE);
    const int w = t >> 6;
    const ReduceRole role = reduce_role(lane);
    if (t == 0) sm.maxlast = 0;
    __syncthreads();
    if (last > 0) atomicMax(&sm.maxlast, last);
    // this wave's own largest last contributor: entries past it touch none of
    // its pixels, so the wave skips them
    uint32_t wlast = last;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t u = (uint32_t)__shfl_xor((int)wlast, o, 64);
        wlast = u > wlast ? u : wlast;
    }
    // every lane holds the maximum; readfirstlane lets the compiler keep it,
    // and the entry loop below, in scalar registers
    wlast = (uint32_t)__builtin_amdgcn_readfirstlane((int)wlast);
    __syncthreads();
    const uint32_t stop = sm.maxlast;

    for (uint32_t hi = stop; hi > start;) {
        const uint32_t lo = (hi - start > (uint32_t)BATCH) ? hi - BATCH : start;
        const int n = (int)(hi - lo);
        __syncthreads();
        for (int k = t; k < BATCH; k += 256) {
            uint32_t h = 0u;
            if (k < n) {
                uint32_t id = a.values[(size_t)(lo + k) * a.vstride];
                sm.id[k] = id;
                const float2 mm = a.means2D[id];
                const float4 cc = make_float4(a.conics[id * 3], a.conics[id * 3 + 1], a.conics[id * 3 + 2], a.opac[id]);
                sm.e[k].xy = mm;
                sm.e[k].co = prescale(cc);
                sm.e[k].rgb = make_float4(a.rgb[id * 3], a.rgb[id * 3 + 1], a.rgb[id * 3 + 2], 0.0f);
                h = strip_hits<false>(mm, cc, tx0, ty0);
            }
            store_masks<BATCH>(sm.mask, k >> 6, h, lane);
#pragma unroll
            for (int c = 0; c < NG; c++) sm.g[k][c] = 0.0f;
        }
        __syncthreads();
        // this wave's entries, back to front from its last contributor
        const int kmax = wlast <= lo ? -1 : ((int)(wlast - lo) < n ? (int)(wlast - lo) : n) - 1;
        for (int ch = kmax >> 6; ch >= 0; ch--) {
            uint64_t bits = wave_uniform(sm.mask[w][ch]);
            if (ch == (kmax >> 6)) bits &= ~0ull >> (63 - (kmax & 63));
            while (bits != 0ull) {
                const int k = (ch << 6) + (63 - __builtin_clzll(bits));
                bits &= ~(1ull << (k & 63));
                const uint32_t j = lo + (uint32_t)k;
                // predicated, as the forward: every lane evaluates the entry and
                // its contribution is selected in or out, through the factors
                // below (visited entries are finite, so a zero factor zeroes
                // every term)
                const float2 m = sm.e[k].xy;
                const float4 co = sm.e[k].co;
                const float4 cl = sm.e[k].rgb;
                const float dx = m.x - fpx, dy = m.y - fpy;
                const float power = power_of(co, dx, dy);
                float ep;
                const float alpha = gsd::raster_alpha_wave(co.w, power, &ep);
                const bool contrib = j < last && power <= 0.0f && alpha >= 1.0f / 255.0f;
                const float oma = 1.0f - alpha;
                // T_i = T / (1 - alpha) as the reference recovers it
                // (rasterizer_backward_cpu.cpp:85), with the hardware
                // reciprocal (1 ulp): gradients need 1e-3, not bits
                const float Ti = (oma > 1e-6f) ? T * __builtin_amdgcn_rcpf(oma) : 0.0f;
                float dal;
                {
                    // gradient-only arithmetic: contracted to FMAs (the
                    // contributor decisions above stay uncontracted)
#pragma clang fp contract(fast)
                    dal = Ti * ((cl.x - acc0) * dC0 + (cl.y - acc1) * dC1 + (cl.z - acc2) * dC2);
                }
                // dL/d(opacity * G) = dal * G; dL/dpower = opacity times that,
                // and the opacity is applied once per Gaussian when the sums
                // leave LDS (alpha >= 0.99 is the clamp: no gradient)
                const float dop = (contrib && alpha < 0.99f) ? dal * ep : 0.0f;
                const float aT = contrib ? alpha * Ti : 0.0f;
                const float pdx = dop * dx, pdy = dop * dy;
                float gv[NG];
                gv[0] = pdx;
                gv[1] = pdy;
                gv[2] = pdx * dx;
                gv[3] = pdx * dy;
                gv[4] = pdy * dy;
                gv[5] = aT * dC0;
                gv[6] = aT * dC1;
                gv[7] = aT * dC2;
                gv[8] = dop;
                // colour behind the next entry: alpha c + (1 - alpha) acc
                const float ac = contrib ? alpha : 0.0f;
                {
#pragma clang fp contract(fast)
                    acc0 += ac * (cl.x - acc0);
                    acc1 += ac * (cl.y - acc1);
                    acc2 += ac * (cl.z - acc2);
                }
                T = contrib ? Ti : T;
                if (__ballot(contrib) != 0ull) {
                    const float v = wave_reduce9(gv, role);
                    if (role.any && v != 0.0f) atomicAdd(&sm.g[k][role.term], v);
                }
            }
        }
        __syncthreads();
        for (int k = t; k < n; k += 256) {
            const uint32_t id = sm.id[k];
            // sums of dop * (dx, dy, dx^2, dx*dy, dy^2); times the opacity
            // they are those of dpow, and -> the reference's terms
            // (rasterizer_backward_cpu.cpp): d_mean = -(a Sx + b Sy, c Sy + b Sx),
            // d_conic = (-Sxx / 2, -Sxy, -Syy / 2); the staged conic is
            // (-a/2, -b, -c/2, o)
            const float4 cs = sm.e[k].co;
            const float ca = -2.0f * cs.x, cb = -cs.y, cc = -2.0f * cs.z, op = cs.w;
            const float sx = op * sm.g[k][0], sy = op * sm.g[k][1];
            float v;
            v = -(ca * sx + cb * sy); if (v != 0.0f) unsafeAtomicAdd(&g.d_means2D[id].x, v);
            v = -(cc * sy + cb * sx); if (v != 0.0f) unsafeAtomicAdd(&g.d_means2D[id].y, v);
            v = -0.5f * (op * sm.g[k][2]); if (v != 0.0f) unsafeAtomicAdd(&g.d_conics[id * 3 + 0], v);
            v = -(op * sm.g[k][3]); if (v != 0.0f) unsafeAtomicAdd(&g.d_conics[id * 3 + 1], v);
            v = -0.5f * (op * sm.g[k][4]); if (v != 0.0f) unsafeAtomicAdd(&g.d_conics[id * 3 + 2], v);
            v = sm.g[k][5]; if (v != 0.0f) unsafeAtomicAdd(&g.d_rgb[id * 3 + 0], v);
            v = sm.g[k][6]; if (v != 0.0f) unsafeAtomicAdd(&g.d_rgb[id * 3 + 1], v);
            v = sm.g[k][7]; if (v != 0.0f) unsafeAtomicAdd(&g.d_rgb[id * 3 + 2], v);
            v = sm.g[k][8]; if (v != 0.0f) unsafeAtomicAdd(&g.d_op[id], v);
        }
        hi = lo;
    }
}

template <int BATCH>
__global__ __launch_bounds__(256) void raster_fwd_kernel(RasterArgs a, PixOut o, RasterSplit sp) {
    __shared__ FwdSmem<BATCH> sm;
    fwd_tile<BATCH>(a, gsd::xcd_remap(blockIdx.x, gridDim.x), sm, o, sp);
}

template <int BATCH>
__global__ __launch_bounds__(256) void raster_bwd_kernel(RasterArgs a, PixIn in, GradArgs g) {
    __shared__ BwdSmem<BATCH> sm;
    bwd_tile<BATCH>(a, gsd::xcd_remap(blockIdx.x, gridDim.x), sm, in, g);
}

// ------------------------------------------------------------ multi-view ----
// View of a global tile (views are few; every thread of the group scans the
// same table, so the loop is uniform).
__device__ __forceinline__ int view_of(const MultiRaster& m, int gtile) {
    int v = 0;
    while (v + 1 < m.V && m.views[v + 1].w <= gtile) v++;
    return v;
}

__device__ __forceinline__ RasterArgs view_args(const MultiRaster& m, int v) {
    const int4 vi = m.views[v];  // (W, H, gx, first global tile)
    RasterArgs a;
    a.W = vi.x;
    a.H = vi.y;
    a.gx = vi.z;
    a.means2D = m.means2D + (size_t)v * m.N;
    a.conics = m.conics + (size_t)v * m.N * 3;
    a.opac = m.opac + (size_t)v * m.N;
    a.rgb = m.rgb + (size_t)v * m.N * 3;
    a.values = m.values;
    a.vstride = m.vstride;
    a.ranges = m.ranges + 2 * (size_t)vi.w;
    a.bg0 = m.bg0;
    a.bg1 = m.bg1;
    a.bg2 = m.bg2;
    return a;
}

// Next global tile: blockIdx.x (one group per tile, XCD-remapped), or from the
// persistent work queues.  The queues are 8 counters over 8 contiguous ranges of
// global tiles; a group starts on the queue of its own XCD (so neighbouring
// tiles share an L2) and moves to the next queue when one runs dry.  Every
// pass takes a fresh counter value, so each group leaves after at most
// total_tiles + 8 passes.
constexpr int NQ = 8;

struct QueueState {
    int q;     // queues tried so far
    int tile;  // broadcast slot
};

__device__ __forceinline__ bool next_tile(const MultiRaster& m, int& g, int& iter, QueueState* s) {
    if (m.counter == nullptr) {
        g = gsd::xcd_remap(blockIdx.x, gridDim.x);
        return iter++ == 0 && blockIdx.x < (unsigned)m.total_tiles;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (iter++ == 0) s->q = 0;
        const int k = m.total_tiles / NQ, r = m.total_tiles % NQ;
        int tile = -1;
        while (s->q < NQ) {
            const int qi = ((int)(blockIdx.x % NQ) + s->q) % NQ;
            const int len = k + (qi < r ? 1 : 0);
            const int i = (int)atomicAdd(&m.counter[qi], 1u);
            if (i < len) {
                tile = qi * k + (qi < r ? qi : r) + i;
                break;
            }
            s->q++;
        }
        s->tile = tile;
    }
    __syncthreads();
    g = s->tile;
    return g >= 0;
}

template <int BATCH>
__global__ __launch_bounds__(256) void raster_fwd_multi_kernel(MultiRaster m, MultiPix o) {
    __shared__ FwdSmem<BATCH> sm;
    __shared__ QueueState qs;
    int g, iter = 0;
    while (next_tile(m, g, iter, &qs)) {
        const int v = view_of(m, g);
        RasterArgs a = view_args(m, v);
        const size_t p0 = (size_t)m.pix_off[v];
        PixOut po{o.img + p0 * 3, o.T_final + p0, o.n_contrib + p0, o.last_contrib + p0};
        fwd_tile<BATCH>(a, g - m.views[v].w, sm, po, RasterSplit{});
    }
}

template <int BATCH>
__global__ __launch_bounds__(256) void raster_bwd_multi_kernel(MultiRaster m, MultiPix o, MultiGrads gr) {
    __shared__ BwdSmem<BATCH> sm;
    __shared__ QueueState qs;
    int g, iter = 0;
    while (next_tile(m, g, iter, &qs)) {
        const int v = view_of(m, g);
        RasterArgs a = view_args(m, v);
        const size_t p0 = (size_t)m.pix_off[v];
        PixIn pi{o.T_final + p0, o.last_contrib + p0, o.d_image + p0 * 3};
        const size_t n0 = (size_t)v * m.N;
        GradArgs ga{gr.d_means2D + n0, gr.d_conics + n0 * 3, gr.d_rgb + n0 * 3, gr.d_op + n0};
        bwd_tile<BATCH>(a, g - m.views[v].w, sm, pi, ga);
    }
}

}  // namespace gsr

using namespace gsr;

hipError_t launch_raster_fwd(const RasterArgs& a, int num_tiles, int batch, float* img, float* T_final, int* n_contrib,
                             uint32_t* last_contrib, hipStream_t s, const RasterSplit* split) {
    if (num_tiles <= 0) return hipSuccess;
    PixOut o{img, T_final, n_contrib, last_contrib};
    const RasterSplit sp = split ? *split : RasterSplit{};
    if (sp.round != 0 && (!T_final || !n_contrib || !last_contrib)) return hipErrorInvalidValue;  // the round state
    switch (batch) {
        case 64: raster_fwd_kernel<64><<<num_tiles, 256, 0, s>>>(a, o, sp); break;
        case 128: raster_fwd_kernel<128><<<num_tiles, 256, 0, s>>>(a, o, sp); break;
        case 256: raster_fwd_kernel<256><<<num_tiles, 256, 0, s>>>(a, o, sp); break;
        case 512: raster_fwd_kernel<512><<<num_tiles, 256, 0, s>>>(a, o, sp); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_raster_bwd(const RasterArgs& a, int num_tiles, int batch, const float* T_final,
                             const uint32_t* last_contrib, const float* d_image, const GradArgs& g, hipStream_t s) {
    if (num_tiles <= 0) return hipSuccess;
    PixIn in{T_final, last_contrib, d_image};
    switch (batch) {
        case 64: raster_bwd_kernel<64><<<num_tiles, 256, 0, s>>>(a, in, g); break;
        case 128: raster_bwd_kernel<128><<<num_tiles, 256, 0, s>>>(a, in, g); break;
        case 256: raster_bwd_kernel<256><<<num_tiles, 256, 0, s>>>(a, in, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// grid: one group per global tile, or `persistent_groups` groups pulling
// tiles from the 8 queue counters at m.counter (zero at launch).
hipError_t launch_raster_fwd_multi(const MultiRaster& m, const MultiPix& o, int batch, int persistent_groups,
                                   hipStream_t s) {
    if (m.total_tiles <= 0) return hipSuccess;
    int grid = m.counter ? persistent_groups : m.total_tiles;
    switch (batch) {
        case 64: raster_fwd_multi_kernel<64><<<grid, 256, 0, s>>>(m, o); break;
        case 128: raster_fwd_multi_kernel<128><<<grid, 256, 0, s>>>(m, o); break;
        case 256: raster_fwd_multi_kernel<256><<<grid, 256, 0, s>>>(m, o); break;
        case 512: raster_fwd_multi_kernel<512><<<grid, 256, 0, s>>>(m, o); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_raster_bwd_multi(const MultiRaster& m, const MultiPix& o, const MultiGrads& g, int batch,
                                   int persistent_groups, hipStream_t s) {
    if (m.total_tiles <= 0) return hipSuccess;
    int grid = m.counter ? persistent_groups : m.total_tiles;
    switch (batch) {
        case 64: raster_bwd_multi_kernel<64><<<grid, 256, 0, s>>>(m, o, g); break;
        case 128: raster_bwd_multi_kernel<128><<<grid, 256, 0, s>>>(m, o, g); break;
        case 256: raster_bwd_multi_kernel<256><<<grid, 256, 0, s>>>(m, o, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
