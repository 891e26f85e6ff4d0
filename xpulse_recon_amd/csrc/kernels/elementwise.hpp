// Elementwise kernels: Error (KernelError.cl:8-24), preprocessing, fills,
// conversions.
// (library-internal; every header here is included once, by xpr_kernels.hip,
// into one translation unit)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/xpr_hip.h"
#include "common.hpp"

namespace {

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
template <typename IS, typename AccT, typename ES>
__global__ __launch_bounds__(256) void k_error(const typename IS::T *__restrict__ img,
                                               const AccT *sums, typename ES::T *err,
                                               uint64_t n, float weight)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float ref = IS::load(img, i);
        const float got = (float)sums[i] / kQ22One;                 // FIXED_TO_FLOAT
        const float e = logf(fmaxf(ref / fmaxf(got, kQ22Eps), kQ22Eps)) * weight;
        ES::store(err, i, e);
    }
}

template <typename AccT>
__global__ __launch_bounds__(256) void k_q22_to_float(const AccT *__restrict__ sums,
                                                      float *__restrict__ out, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (float)sums[i] / kQ22One;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_raw(T *__restrict__ dst, T value, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = value;
}

template <typename DS, typename SS>
__global__ __launch_bounds__(256) void k_convert(typename DS::T *__restrict__ dst,
                                                 const typename SS::T *__restrict__ src,
                                                 uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        DS::store(dst, i, SS::load(src, i));
}

// Order key of a float: signed-int comparison of keys matches float order.
__device__ __forceinline__ int32_t fkey(float f)
{
    int32_t b = __builtin_bit_cast(int32_t, f);
    return b < 0 ? (b ^ 0x7fffffff) : b;
}

__global__ __launch_bounds__(256) void k_minmax(const float *__restrict__ data, uint64_t n,
                                                int32_t *minmax)
{
    int32_t lo = 0x7fffffff, hi = (int32_t)0x80000000;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = data[i];
        if (__builtin_isnormal(v)) {
            const int32_t kk = fkey(v);
            lo = min(lo, kk);
            hi = max(hi, kk);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(minmax + 0, lo);
        atomicMax(minmax + 1, hi);
    }
}

template <typename OS>
__global__ __launch_bounds__(256) void k_preprocess(const float *__restrict__ in,
                                                    typename OS::T *__restrict__ out,
                                                    uint64_t n, int normalize, int mlog,
                                                    float mn, float mx)
{
    const float hi_clamp = 1.0f - kQ22Eps;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float v = in[i];
        if (normalize)
            v = __builtin_isnormal(v) ? (v - mn) / (mx - mn) : (mx + mn) / 2.0f;
        if (mlog) {
            float t = (hi_clamp < v) ? hi_clamp : v;   // std::min
            t = (t < kQ22Eps) ? kQ22Eps : t;           // std::max
            v = -logf(t);
        }
        OS::store(out, i, v);
    }
}

} // namespace
