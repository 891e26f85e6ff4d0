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
        const float got = sum_value(sums[i]);                       // FIXED_TO_FLOAT
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
        out[i] = sum_value(sums[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_raw(T *__restrict__ dst, T value, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = value;
}

__global__ __launch_bounds__(256) void k_scale(float *__restrict__ x, uint64_t n, float factor)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        x[i] = x[i] * factor;
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

// one pixel: normalisation (non-normal values to the midpoint) and -log with
// its clamp (data_loader.hpp:55-86)
__device__ __forceinline__ float preprocess_value(float v, int normalize, int mlog, float mn, float mx)
{
    const float hi_clamp = 1.0f - kQ22Eps;
    if (normalize)
        v = __builtin_isnormal(v) ? (v - mn) / (mx - mn) : (mx + mn) / 2.0f;
    if (mlog) {
        float t = (hi_clamp < v) ? hi_clamp : v;   // std::min
        t = (t < kQ22Eps) ? kQ22Eps : t;           // std::max
        v = -logf(t);
    }
    return v;
}

template <typename OS>
__global__ __launch_bounds__(256) void k_preprocess(const float *__restrict__ in,
                                                    typename OS::T *__restrict__ out,
                                                    uint64_t n, int normalize, int mlog,
                                                    float mn, float mx)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        OS::store(out, i, preprocess_value(in[i], normalize, mlog, mn, mx));
}

// k_preprocess over n_views views of 4 n4 pixels, view j read from
// in + j * in_stride4 (16-byte aligned rows) and written to out + 4 j n4:
// four consecutive pixels per access and two accesses in flight per thread
// (one pixel per thread and access leaves the launch waiting on memory
// latency: 36 views of 768^2 took 140 us).  The same arithmetic per pixel.
template <typename OS>
__device__ __forceinline__ void store_preprocessed4(typename OS::T *__restrict__ out, uint64_t o,
                                                    float4 v, int normalize, int mlog, float mn,
                                                    float mx)
{
    OS::store(out, o + 0, preprocess_value(v.x, normalize, mlog, mn, mx));
    OS::store(out, o + 1, preprocess_value(v.y, normalize, mlog, mn, mx));
    OS::store(out, o + 2, preprocess_value(v.z, normalize, mlog, mn, mx));
    OS::store(out, o + 3, preprocess_value(v.w, normalize, mlog, mn, mx));
}

template <typename OS>
__global__ __launch_bounds__(256) void k_preprocess_views4(const float4 *__restrict__ in,
                                                           uint64_t in_stride4,
                                                           typename OS::T *__restrict__ out,
                                                           uint64_t n4, uint32_t n_views,
                                                           int normalize, int mlog, float mn,
                                                           float mx)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t j = blockIdx.y; j < n_views; j += gridDim.y) {
        const float4 *__restrict__ src = in + (uint64_t)j * in_stride4;
        const uint64_t o = (uint64_t)j * n4 * 4u;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += 2 * stride) {
            const uint64_t i2 = i + stride;
            const float4 a = src[i];
            const float4 b = i2 < n4 ? src[i2] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            store_preprocessed4<OS>(out, o + 4u * i, a, normalize, mlog, mn, mx);
            if (i2 < n4) store_preprocessed4<OS>(out, o + 4u * i2, b, normalize, mlog, mn, mx);
        }
    }
}

// k_preprocess over n_views views of n elements each, view j read from
// in + j * in_stride and written to out + j * n (one launch for the views of
// a sub-iteration, which lie at a fixed stride in the projection stack)
template <typename OS>
__global__ __launch_bounds__(256) void k_preprocess_views(const float *__restrict__ in,
                                                          uint64_t in_stride,
                                                          typename OS::T *__restrict__ out,
                                                          uint64_t n, uint32_t n_views,
                                                          int normalize, int mlog, float mn,
                                                          float mx)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t j = blockIdx.y; j < n_views; j += gridDim.y) {
        const float *__restrict__ src = in + (uint64_t)j * in_stride;
        const uint64_t o = (uint64_t)j * n;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            OS::store(out, o + i, preprocess_value(src[i], normalize, mlog, mn, mx));
    }
}

// Self-check of the exact fast forms (recip's fast branch and pixel_of)
// over every float32 bit pattern; bad[0..2] as xpr_check_fast_paths counts.
__global__ __launch_bounds__(256) void k_check_fast_paths(unsigned long long *__restrict__ bad)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t n0 = 0, n1 = 0, n2 = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
        const float w = __builtin_bit_cast(float, (uint32_t)i);
        const float aw = fabsf(w);
        if (aw > 0.0f && aw <= 3.4e38f) {                           // finite, non-zero
            const float r0 = __builtin_amdgcn_rcpf(w);
            const float r = __builtin_fmaf(__builtin_fmaf(-w, r0, 1.0f), r0, r0);
            const float ref = 1.0f / w;                               // IEEE quotient
            if (__builtin_bit_cast(uint32_t, r) != __builtin_bit_cast(uint32_t, ref)) {
                if (aw >= 1e-6f && aw <= 1e6f) ++n0;
                else ++n2;
            }
        }
        if (w > -0.5f && w < 8388608.0f && pixel_of(w) != (int32_t)roundf(w)) ++n1;
    }
    // one atomic per wave and counter
    for (int off = 32; off > 0; off >>= 1) {
        n0 += __shfl_xor(n0, off);
        n1 += __shfl_xor(n1, off);
        n2 += __shfl_xor(n2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (n0) atomicAdd(&bad[0], (unsigned long long)n0);
        if (n1) atomicAdd(&bad[1], (unsigned long long)n1);
        if (n2) atomicAdd(&bad[2], (unsigned long long)n2);
    }
}

} // namespace
