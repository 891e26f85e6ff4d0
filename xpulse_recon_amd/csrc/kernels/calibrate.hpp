// VALU issue ceiling of the device (no counterpart in the reference): the
// denominator of the projectors' VALU roofline.  Every lane runs 16
// independent instruction chains; the trip count is uniform (loop control on
// the scalar unit) and 8 waves per SIMD keep the chains' latency hidden.
// gfx950 issues simple integer instructions faster than FMAs (measured: two-
// VGPR v_xor_b32 streams reach about 1.4 wave-instructions per CU-cycle,
// v_fma_f32 streams about 0.95), so the ceiling is the best of several kinds:
//   0 v_fma_f32 (VOP3), 1 v_mul_f32, 2 v_add_f32 (VOP2): one VGPR read each,
//     the other operands a scalar register or an inline constant;
//   3, 4, 5: 8, 12 and 16 of the 16 chains v_xor_b32 of two VGPRs, the rest
//     v_fma_f32.  The xors are written as single-instruction inline asm: the
//     compiler would fold unrolled integer chains.
// Exactly 16 vector instructions per chain step (per loop trip, or 64 per
// trip where the loop is unrolled 4x).
// (library-internal; included once, by xpr_kernels.hip)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kValuChains = 16;
constexpr int kValuKinds = 6;

template <int Kind>
__global__ __launch_bounds__(256) void k_valu_ceiling(float *__restrict__ out, uint32_t iters,
                                                      float m)
{
    constexpr int kXor = Kind == 3 ? 8 : Kind == 4 ? 12 : Kind == 5 ? 16 : 0;
    constexpr int kFp = kValuChains - kXor;
    float a[kFp > 0 ? kFp : 1];
    uint32_t b[kXor > 0 ? kXor : 1], c[kXor > 0 ? kXor : 1];
#pragma unroll
    for (int j = 0; j < kFp; ++j) a[j] = float(threadIdx.x + j) * 1e-3f;
#pragma unroll
    for (int j = 0; j < kXor; ++j) {
        b[j] = threadIdx.x * 7u + (uint32_t)j;
        c[j] = threadIdx.x * 13u + 5u * (uint32_t)j + 1u;
    }
    // iters is a multiple of 4
#pragma unroll 4
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kFp; ++j) {
            if constexpr (Kind == 1) a[j] = a[j] * m;
            else if constexpr (Kind == 2) a[j] = a[j] + m;
            else a[j] = fmaf(a[j], m, 0.5f);
        }
#pragma unroll
        for (int j = 0; j < kXor; ++j) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(b[j]) : "v"(c[j]));
    }
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < kFp; ++j) sum += a[j];
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < kXor; ++j) x += b[j];
    out[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = sum + (float)x;
}

} // namespace
