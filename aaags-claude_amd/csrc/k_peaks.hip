// k_peaks.hip -- measured ceilings of this device, the denominators the
// rasterizers' and the HBM-bound stages' rooflines are quoted against next to
// the published figures:
//   - HBM: a 16-byte-per-lane grid-stride copy of 1 GiB (read + write bytes);
//   - LDS: conflict-free 16-byte reads (ds_read_b128), each lane a different
//     address, eight waves per CU;
//   - VALU: eight independent chains of one wave64 FP32 instruction kind per
//     lane (multiply; exp2 for the transcendental rate), eight waves per CU.
// Each figure is the best of five timed launches (HIP events on the context's
// stream) after one untimed launch.  The buffers are allocated and freed by
// the call; it is a measurement, run outside any timed region.
#include "gs_device.h"
#include "gs_internal.h"

namespace gsp {

// four independent 16-byte loads in flight per lane before their stores
__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ in, float4* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const float4 a = in[i], b = in[i + stride], c = in[i + 2 * stride], d = in[i + 3 * stride];
        out[i] = a;
        out[i + stride] = b;
        out[i + 2 * stride] = c;
        out[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) out[i] = in[i];
}

constexpr int kLdsF4 = 2048;  // 32 KiB of float4 (five workgroups fit in a CU's LDS)

__global__ __launch_bounds__(256) void lds_read_kernel(int iters, float* __restrict__ out) {
    __shared__ float4 buf[kLdsF4];
    const int t = threadIdx.x;
    for (int i = t; i < kLdsF4; i += 256) buf[i] = make_float4((float)i, (float)(i + 1), (float)(i + 2), (float)(i + 3));
    __syncthreads();
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0, a3 = a0;
    int idx = t;
    for (int it = 0; it < iters; it++) {
        // four reads at constant offsets from one address: consecutive lanes
        // read consecutive 16-byte words (no bank conflicts)
        const float4 v0 = buf[idx & (kLdsF4 - 1)];
        const float4 v1 = buf[(idx + 256) & (kLdsF4 - 1)];
        const float4 v2 = buf[(idx + 512) & (kLdsF4 - 1)];
        const float4 v3 = buf[(idx + 768) & (kLdsF4 - 1)];
        a0.x += v0.x; a0.y += v0.y; a0.z += v0.z; a0.w += v0.w;
        a1.x += v1.x; a1.y += v1.y; a1.z += v1.z; a1.w += v1.w;
        a2.x += v2.x; a2.y += v2.y; a2.z += v2.z; a2.w += v2.w;
        a3.x += v3.x; a3.y += v3.y; a3.z += v3.z; a3.w += v3.w;
        idx += 1024;
    }
    const float s = a0.x + a0.y + a0.z + a0.w + a1.x + a1.y + a1.z + a1.w + a2.x + a2.y + a2.z + a2.w + a3.x + a3.y +
                    a3.z + a3.w;
    out[blockIdx.x * 256 + t] = s;
}

template <int KIND>
__global__ __launch_bounds__(256) void valu_kernel(float* __restrict__ out, float a, int iters) {
    float x[8];
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = threadIdx.x * 0.001f + j;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = KIND == 0 ? x[j] * a : __builtin_amdgcn_exp2f(x[j]);
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; j++) s += x[j];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace gsp

namespace {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

template <typename F>
int best_ms(hipStream_t s, Timer& tm, F launch, float* out) {
    hipError_t e = launch();
    if (e != hipSuccess) return gs_hip_status(e, "peak probe launch");
    float best = 1e30f;
    for (int rep = 0; rep < 5; rep++) {
        if ((e = hipEventRecord(tm.a, s)) != hipSuccess) return gs_hip_status(e, "hipEventRecord");
        if ((e = launch()) != hipSuccess) return gs_hip_status(e, "peak probe launch");
        if ((e = hipEventRecord(tm.b, s)) != hipSuccess) return gs_hip_status(e, "hipEventRecord");
        if ((e = hipEventSynchronize(tm.b)) != hipSuccess) return gs_hip_status(e, "hipEventSynchronize");
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, tm.a, tm.b)) != hipSuccess) return gs_hip_status(e, "hipEventElapsedTime");
        if (ms < best) best = ms;
    }
    *out = best;
    return GS_OK;
}

}  // namespace

extern "C" int gs_device_peaks(gs_ctx* c, gs_peaks* out) {
    if (!c || !out) return gs_set_error(GS_ERR_INVALID, "null argument");
    hipStream_t s = c->stream;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, c->device);
    if (e != hipSuccess) return gs_hip_status(e, "hipGetDeviceProperties");
    const int cus = prop.multiProcessorCount;
    Timer tm;
    if ((e = hipEventCreate(&tm.a)) != hipSuccess || (e = hipEventCreate(&tm.b)) != hipSuccess)
        return gs_hip_status(e, "hipEventCreate");
    *out = gs_peaks{};
    out->compute_units = cus;
    out->clock_mhz = prop.clockRate / 1000.0;

    // HBM copy: 1 GiB in, 1 GiB out
    const size_t bytes = size_t(1) << 30, n4 = bytes / 16;
    DevBuf src, dst;
    if ((e = hipMalloc(&src.p, bytes)) != hipSuccess) return gs_hip_status(e, "hipMalloc");
    if ((e = hipMalloc(&dst.p, bytes)) != hipSuccess) return gs_hip_status(e, "hipMalloc");
    if ((e = launch_zero(src.p, bytes, s)) != hipSuccess) return gs_hip_status(e, "zero");
    float ms = 0.f;
    int st = GS_OK;
    for (int per_cu : {4, 8, 16, 32}) {  // the best of a few grid sizes
        const int copy_blocks = cus * per_cu;
        st = best_ms(s, tm, [&] {
            gsp::copy_kernel<<<copy_blocks, 256, 0, s>>>(static_cast<const float4*>(src.p), static_cast<float4*>(dst.p),
                                                         n4);
            return hipGetLastError();
        }, &ms);
        if (st != GS_OK) return st;
        const double gbs = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
        if (gbs > out->hbm_copy_gbs) out->hbm_copy_gbs = gbs;
    }

    // LDS reads: eight 256-thread workgroups per CU (five resident), 32 KiB of LDS each
    const int blocks = cus * 8, lds_iters = 4096;
    DevBuf sink;
    if ((e = hipMalloc(&sink.p, (size_t)blocks * 256 * sizeof(float))) != hipSuccess) return gs_hip_status(e, "hipMalloc");
    float* so = static_cast<float*>(sink.p);
    st = best_ms(s, tm, [&] {
        gsp::lds_read_kernel<<<blocks, 256, 0, s>>>(lds_iters, so);
        return hipGetLastError();
    }, &ms);
    if (st != GS_OK) return st;
    out->lds_read_gbs = (double)blocks * 256 * 16.0 * 4 * lds_iters / (ms * 1e-3) / 1e9;

    // VALU issue: wave64 instructions per second
    const int valu_iters = 2048;
    const double waves = (double)blocks * 4;
    st = best_ms(s, tm, [&] {
        gsp::valu_kernel<0><<<blocks, 256, 0, s>>>(so, 0.999f, valu_iters);
        return hipGetLastError();
    }, &ms);
    if (st != GS_OK) return st;
    out->valu_wave_insts_per_s = waves * valu_iters * 8 / (ms * 1e-3);
    st = best_ms(s, tm, [&] {
        gsp::valu_kernel<1><<<blocks, 256, 0, s>>>(so, 0.999f, valu_iters);
        return hipGetLastError();
    }, &ms);
    if (st != GS_OK) return st;
    out->exp_wave_insts_per_s = waves * valu_iters * 8 / (ms * 1e-3);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return gs_hip_status(e, "hipStreamSynchronize");
    return GS_OK;
}
