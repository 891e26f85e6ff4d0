// VALU issue ceiling of the device (no counterpart in the reference): the
// denominator of the projectors' VALU roofline.  Every lane runs 16
// independent chains of one instruction kind; each instruction reads one
// VGPR (the other operands are a scalar register or an inline constant, so
// no VGPR bank conflicts), the trip count is uniform (loop control on the
// scalar unit) and 8 waves per SIMD keep the chains' latency hidden.  Kinds:
// 0 v_fma_f32 (VOP3), 1 v_mul_f32 and 2 v_add_f32 (VOP2); exactly 64 of them
// per trip of the 4x unrolled loop (integer chains are not used: the
// compiler folds unrolled integer adds).
// (library-internal; included once, by xpr_kernels.hip)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kValuChains = 16;

template <int Kind>
__global__ __launch_bounds__(256) void k_valu_ceiling(float *__restrict__ out, uint32_t iters,
                                                      float m)
{
    float a[kValuChains];
#pragma unroll
    for (int j = 0; j < kValuChains; ++j) a[j] = float(threadIdx.x + j) * 1e-3f;
    // iters is a multiple of 4
#pragma unroll 4
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kValuChains; ++j) {
            if constexpr (Kind == 0) a[j] = fmaf(a[j], m, 0.5f);
            else if constexpr (Kind == 1) a[j] = a[j] * m;
            else a[j] = a[j] + m;
        }
    }
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < kValuChains; ++j) sum += a[j];
    out[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

} // namespace
