// k_optim.hip -- streaming kernels: L1 loss + image gradient, fused Adam over
// the five parameter tensors, and a zero-fill usable inside graph capture.
// All three are HBM-bound; they use 16-byte accesses where alignment allows.
#include "gs_device.h"

namespace gso {

constexpr int L1_THREADS = 256;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float l1_grad(float d, float inv_n) {
    return (d > 0.0f) ? inv_n : ((d < 0.0f) ? -inv_n : 0.0f);
}

// loss.cpp:4-30: d_image = sign(r - g) / n (0 on ties); |r - g| summed in FP64.
__global__ __launch_bounds__(L1_THREADS) void l1_kernel(const float* __restrict__ r, const float* __restrict__ g,
                                                        float* __restrict__ d, uint32_t n, float inv_n,
                                                        double* __restrict__ partial) {
    __shared__ double s_w[L1_THREADS / 64];
    double s = 0.0;
    const uint32_t n4 = n / 4;
    const float4* r4 = reinterpret_cast<const float4*>(r);
    const float4* g4 = reinterpret_cast<const float4*>(g);
    float4* d4 = reinterpret_cast<float4*>(d);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x) {
        float4 a = r4[i], b = g4[i];
        float4 df = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
        s += (double)fabsf(df.x) + (double)fabsf(df.y) + (double)fabsf(df.z) + (double)fabsf(df.w);
        d4[i] = make_float4(l1_grad(df.x, inv_n), l1_grad(df.y, inv_n), l1_grad(df.z, inv_n), l1_grad(df.w, inv_n));
    }
    if (blockIdx.x == 0) {
        for (uint32_t i = n4 * 4 + threadIdx.x; i < n; i += blockDim.x) {
            float df = r[i] - g[i];
            s += (double)fabsf(df);
            d[i] = l1_grad(df, inv_n);
        }
    }
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < L1_THREADS / 64; k++) t += s_w[k];
        partial[blockIdx.x] = t;
    }
}

// One block row per view (blockIdx.y); each view's loss is its own mean.
__global__ __launch_bounds__(L1_THREADS) void l1_multi_kernel(const float* __restrict__ r, const float* __restrict__ g,
                                                              float* __restrict__ d, const int* __restrict__ pix_off,
                                                              double* __restrict__ partial, int nb) {
    __shared__ double s_w[L1_THREADS / 64];
    const int v = blockIdx.y;
    const size_t e0 = (size_t)pix_off[v] * 3, e1 = (size_t)pix_off[v + 1] * 3;
    const uint32_t n = (uint32_t)(e1 - e0);
    const float inv_n = 1.0f / (float)n;
    double s = 0.0;
    for (size_t e = e0 + blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < e1; e += (size_t)gridDim.x * blockDim.x) {
        float df = r[e] - g[e];
        s += (double)fabsf(df);
        d[e] = l1_grad(df, inv_n);
    }
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < L1_THREADS / 64; k++) t += s_w[k];
        partial[(size_t)v * nb + blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(256) void l1_multi_finish_kernel(const double* __restrict__ partial, int nb,
                                                              const int* __restrict__ pix_off,
                                                              double* __restrict__ loss) {
    __shared__ double s_w[4];
    const int v = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += partial[(size_t)v * nb + i];
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        loss[v] = (s_w[0] + s_w[1] + s_w[2] + s_w[3]) / (double)((size_t)(pix_off[v + 1] - pix_off[v]) * 3);
}

__global__ __launch_bounds__(256) void l1_finish_kernel(const double* __restrict__ partial, int nb, uint32_t n,
                                                        double* __restrict__ loss) {
    __shared__ double s_w[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += partial[i];
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *loss = (s_w[0] + s_w[1] + s_w[2] + s_w[3]) / (double)n;
}

// optimizer.cpp:32-43, same operation order (no fused multiply-add).
struct AdamH {
    float b1, b2, eps, bc1, bc2;
};

__device__ __forceinline__ void adam1(float& p, float g, float& m, float& v, float lr, const AdamH& a) {
    m = a.b1 * m + (1.0f - a.b1) * g;
    v = a.b2 * v + (1.0f - a.b2) * g * g;
    float mh = m * a.bc1;
    float vh = v * a.bc2;
    p -= lr * mh / (sqrtf(vh) + a.eps);
}

// Non-temporal 16-byte load / store (the builtins take native vector types).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 nt_load4(const float4* p) {
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void nt_store4(float4 x, float4* p) {
    const nt_f4 v = {x.x, x.y, x.z, x.w};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_f4*>(p));
}

__global__ __launch_bounds__(256) void adam_kernel(AdamArgs args) {
    if (args.skip && *args.skip) return;
    const AdamSeg& s = args.seg[blockIdx.y];
    AdamH a{args.b1, args.b2, args.eps, args.bc1, args.bc2};
    float lr = s.lr;
    if (args.hp) {
        const AdamDev h = *args.hp;
        a = AdamH{h.b1, h.b2, h.eps, h.bc1, h.bc2};
        lr = h.lr[blockIdx.y];
    }
    const uint32_t n = s.n;
    const uint32_t n4 = n / 4;
    float4* p4 = reinterpret_cast<float4*>(s.p);
    const float4* g4 = reinterpret_cast<const float4*>(s.g);
    float4* m4 = reinterpret_cast<float4*>(s.m);
    float4* v4 = reinterpret_cast<float4*>(s.v);
    // the moments and gradients are touched once per step: non-temporal, so
    // that the parameters the next forward reads stay in the caches
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x) {
        float4 p = p4[i], g = nt_load4(g4 + i), m = nt_load4(m4 + i), v = nt_load4(v4 + i);
        adam1(p.x, g.x, m.x, v.x, lr, a);
        adam1(p.y, g.y, m.y, v.y, lr, a);
        adam1(p.z, g.z, m.z, v.z, lr, a);
        adam1(p.w, g.w, m.w, v.w, lr, a);
        p4[i] = p;
        nt_store4(m, m4 + i);
        nt_store4(v, v4 + i);
    }
    if (blockIdx.x == 0) {
        for (uint32_t i = n4 * 4 + threadIdx.x; i < n; i += blockDim.x) {
            float p = s.p[i], m = s.m[i], v = s.v[i];
            adam1(p, s.g[i], m, v, lr, a);
            s.p[i] = p;
            s.m[i] = m;
            s.v[i] = v;
        }
    }
}

// optimizer.cpp:9-26: p -= lr * g, per tensor.
__global__ __launch_bounds__(256) void sgd_kernel(AdamArgs a) {
    const AdamSeg& s = a.seg[blockIdx.y];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x)
        s.p[i] -= s.lr * s.g[i];
}

__global__ __launch_bounds__(256) void zero_kernel(uint4* __restrict__ p16, size_t n16, unsigned char* __restrict__ tail,
                                                   size_t ntail) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p16[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0)
        for (size_t i = threadIdx.x; i < ntail; i += blockDim.x) tail[i] = 0;
}

}  // namespace gso

using namespace gso;

int l1_partials_needed(uint32_t n) {
    uint32_t n4 = n / 4;
    int nb = (int)((n4 + L1_THREADS - 1) / L1_THREADS);
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    return nb;
}

hipError_t launch_l1(const float* r, const float* g, float* d, uint32_t n, double* partial, double* loss,
                     hipStream_t s) {
    int nb = l1_partials_needed(n);
    float inv_n = 1.0f / (float)n;
    l1_kernel<<<nb, L1_THREADS, 0, s>>>(r, g, d, n, inv_n, partial);
    l1_finish_kernel<<<1, 256, 0, s>>>(partial, nb, n, loss);
    return hipGetLastError();
}

hipError_t launch_l1_multi(const float* r, const float* g, float* d, const int* pix_off_dev, int V, int nb,
                           double* partial, double* loss, hipStream_t s) {
    if (V <= 0) return hipSuccess;
    l1_multi_kernel<<<dim3(nb, V), L1_THREADS, 0, s>>>(r, g, d, pix_off_dev, partial, nb);
    l1_multi_finish_kernel<<<V, 256, 0, s>>>(partial, nb, pix_off_dev, loss);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t s) {
    uint32_t mx = 0;
    for (int k = 0; k < 5; k++) mx = a.seg[k].n > mx ? a.seg[k].n : mx;
    uint32_t blocks = (mx / 4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    adam_kernel<<<dim3(blocks, 5), 256, 0, s>>>(a);
    return hipGetLastError();
}

__global__ void count_if_kernel(const uint32_t* flag, uint32_t* counter) {
    if (*flag) *counter += 1u;
}

hipError_t launch_count_if(const uint32_t* flag, uint32_t* counter, hipStream_t s) {
    count_if_kernel<<<1, 1, 0, s>>>(flag, counter);
    return hipGetLastError();
}

hipError_t launch_sgd(const AdamArgs& a, hipStream_t s) {
    uint32_t mx = 0;
    for (int k = 0; k < 5; k++) mx = a.seg[k].n > mx ? a.seg[k].n : mx;
    uint32_t blocks = (mx + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    sgd_kernel<<<dim3(blocks, 5), 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_zero(void* ptr, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    uintptr_t p = reinterpret_cast<uintptr_t>(ptr);
    size_t head = (16 - (p & 15)) & 15;
    if (head > bytes) head = bytes;
    // unaligned head bytes are handled by a second tiny launch over the tail path
    if (head) {
        zero_kernel<<<1, 64, 0, s>>>(nullptr, 0, reinterpret_cast<unsigned char*>(ptr), head);
    }
    size_t rest = bytes - head;
    size_t n16 = rest / 16;
    size_t ntail = rest - n16 * 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    zero_kernel<<<(unsigned)blocks, 256, 0, s>>>(reinterpret_cast<uint4*>(p + head), n16,
                                                 reinterpret_cast<unsigned char*>(p + head + n16 * 16), ntail);
    return hipGetLastError();
}
