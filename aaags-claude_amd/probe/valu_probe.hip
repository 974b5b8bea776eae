// valu_probe.hip -- issue rate of wave64 VALU instructions on this GPU, the
// ceiling of the VALU-bound rasterizers.  Each kernel runs 8 independent
// chains of one instruction kind per lane over four full grids of
// workgroups (8 of 256 threads per CU each); the rate is wave-instructions per second, and per
// SIMD-cycle at the nominal 2.4 GHz.  Built with the library's flags (no
// packed FP32, no contraction).
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHAINS(OP)                                                        \
    float x[8];                                                           \
    _Pragma("unroll") for (int j = 0; j < 8; j++) x[j] = threadIdx.x * 0.001f + j; \
    for (int i = 0; i < iters; i++) {                                     \
        _Pragma("unroll") for (int j = 0; j < 8; j++) { OP; }             \
    }                                                                     \
    float s = 0.f;                                                        \
    _Pragma("unroll") for (int j = 0; j < 8; j++) s += x[j];              \
    out[blockIdx.x * 256 + threadIdx.x] = s;

__global__ __launch_bounds__(256) void p_mul(float* out, float a, float b, int iters) { CHAINS(x[j] = x[j] * a) }
__global__ __launch_bounds__(256) void p_fma(float* out, float a, float b, int iters) {
    CHAINS(x[j] = __builtin_fmaf(x[j], a, b))
}
__global__ __launch_bounds__(256) void p_cnd(float* out, float a, float b, int iters) { CHAINS(x[j] = (x[j] < a) ? x[j] : b) }
__global__ __launch_bounds__(256) void p_xor(float* out, float a, float b, int iters) {
    CHAINS(x[j] = __int_as_float(__float_as_int(x[j]) ^ i))
}
__global__ __launch_bounds__(256) void p_exp(float* out, float a, float b, int iters) { CHAINS(x[j] = __builtin_amdgcn_exp2f(x[j])) }
__global__ __launch_bounds__(256) void p_rcp(float* out, float a, float b, int iters) { CHAINS(x[j] = __builtin_amdgcn_rcpf(x[j])) }
__global__ __launch_bounds__(256) void p_dpp(float* out, float a, float b, int iters) {
    CHAINS(x[j] = x[j] + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x[j]), 0x124, 0xF, 0xF, false)))
}

typedef void (*Probe)(float*, float, float, int);

int main() {
    const int blocks = 256 * 8 * 4, iters = 2048;
    float* out = nullptr;
    if (hipMalloc(&out, sizeof(float) * blocks * 256) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // VALU instructions per loop iteration of each kernel (one per chain;
    // the compare-and-select kernel issues a compare and a select)
    const Probe ks[] = {p_mul, p_fma, p_cnd, p_xor, p_exp, p_rcp, p_dpp};
    const char* nm[] = {"v_mul_f32", "v_fma_f32", "v_cmp+v_cndmask", "v_xor_b32", "v_exp_f32", "v_rcp_f32",
                        "v_add_f32 (DPP)"};
    const int per[] = {8, 8, 16, 8, 8, 8, 8};
    for (int q = 0; q < 7; q++) {
        ks[q]<<<blocks, 256>>>(out, 0.999f, 0.001f, 16);
        if (hipDeviceSynchronize() != hipSuccess) return 1;
        float best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            (void)hipEventRecord(e0);
            ks[q]<<<blocks, 256>>>(out, 0.999f, 0.001f, iters);
            (void)hipEventRecord(e1);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) best = ms;
        }
        const double insts = (double)blocks * 4 * iters * per[q];
        printf("%-18s %8.3f ms  %7.1f G wave-instructions/s  %.3f per SIMD-cycle at 2.4 GHz\n", nm[q], best,
               insts / best / 1e6, insts / (best * 1e-3) / (256.0 * 4 * 2.4e9));
    }
    (void)hipFree(out);
    return 0;
}
