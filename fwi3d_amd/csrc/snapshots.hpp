// Snapshot-store kernels: the forward wavefield every `sub` steps, kept in HBM for the
// imaging condition of the adjoint run (the role of zfpWrapper::set/get in the
// reference, src/include/zfpWrapper.hpp:78-116). FP32 slots are plain copies; 16-bit
// slots hold BF16 / FP16 codes of u / s with the per-snapshot scale of SnapView.
#pragma once
#include "common.hpp"

namespace fwi3d {

// max|u| over n values into *maxbits (float bits; the word starts at zero).
__global__ __launch_bounds__(256) void snap_maxabs(const float* __restrict__ u, unsigned* __restrict__ maxbits,
                                                   long long n) {
    float m = 0.f;
    const long long n4 = n >> 2;
    const float4* __restrict__ u4 = (const float4*)u;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
        const float4 v = u4[i];
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
    for (long long i = 4 * n4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(u[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, __float_as_uint(m));
}

// 16-bit codes of u / s, s = max|u| / 2^15 (round to nearest even).
template <int FMT>
__global__ __launch_bounds__(256) void snap_store(const float* __restrict__ u, void* __restrict__ dst,
                                                  const unsigned* __restrict__ maxbits, long long n) {
    const float m = __uint_as_float(*maxbits);
    const float inv = m > 0.f ? kSnapTop / m : 0.f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float v = u[i] * inv;
        if (FMT == SNAP_BF16) ((unsigned short*)dst)[i] = bf16_rne(v);
        else ((__half*)dst)[i] = __float2half_rn(v);
    }
}

}  // namespace fwi3d
