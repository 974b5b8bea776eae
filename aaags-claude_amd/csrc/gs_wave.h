// gs_wave.h -- wave64 building blocks shared by the sort and binning kernels:
// lane masks, shuffle scans, the block-wide exclusive scan, and the stable
// digit match that ranks radix digits with ballots.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gsw {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(v, o, 64);
        if (l >= o) v += u;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Exclusive block scan of one value per thread (s_wave: one entry per wave of
// the block, at most 16); returns the
// exclusive prefix and writes the block total to *total.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* s_wave, T* total) {
    int w = threadIdx.x >> 6, l = lane_id();
    T inc = wave_inclusive_scan(v);
    if (l == 63) s_wave[w] = inc;
    __syncthreads();
    T wbase = 0, tot = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
        T s = s_wave[k];
        if (k < w) wbase += s;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return wbase + inc - v;
}

// ---------------------------------------------------------- radix digits ----
template <int BITS>
__device__ __forceinline__ uint32_t digit_of(uint64_t k, int shift) {
    return (uint32_t)(k >> shift) & ((1u << BITS) - 1u);
}

// Lanes (among `active`) whose digit equals this lane's: one ballot per digit
// bit; a lane differs from this one where its bit differs, i.e. where the
// ballot XOR this lane's bit (spread to all ones) is set.  Two bits per step,
// so the two XORs fold into one three-input OR.
template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
    uint32_t dlo = 0u, dhi = 0u;
#pragma unroll
    for (int b = 0; b < BITS; b += 2) {
        const uint32_t m0 = (uint32_t)((int32_t)(d << (31 - b)) >> 31);  // all ones where bit b is set
        const uint64_t bal0 = __builtin_amdgcn_ballot_w64(m0 != 0u);
        if (b + 1 < BITS) {
            const uint32_t m1 = (uint32_t)((int32_t)(d << (30 - b)) >> 31);
            const uint64_t bal1 = __builtin_amdgcn_ballot_w64(m1 != 0u);
            dlo |= ((uint32_t)bal0 ^ m0) | ((uint32_t)bal1 ^ m1);
            dhi |= ((uint32_t)(bal0 >> 32) ^ m0) | ((uint32_t)(bal1 >> 32) ^ m1);
        } else {
            dlo |= (uint32_t)bal0 ^ m0;
            dhi |= (uint32_t)(bal0 >> 32) ^ m0;
        }
    }
    return active & ~(((uint64_t)dhi << 32) | dlo);
}

}  // namespace gsw
