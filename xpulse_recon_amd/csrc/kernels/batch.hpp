// Persistent kernels of the batched step engine.
// (library-internal; every header here is included once, by xpr_kernels.hip,
// into one translation unit)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/xpr_hip.h"
#include "../batch.hpp"
#include "backward.hpp"
#include "elementwise.hpp"
#include "forward.hpp"

namespace {

// ---------------------------------------------------------------------------
// Batched steps: persistent kernels over a table of work items of several
// reconstructions (one kernel per phase, units taken from an atomic counter).
// Every unit is the same arithmetic as the corresponding single-job launch, so
// a job's result does not depend on the other items of its steps.
// ---------------------------------------------------------------------------
using xpr_batch::Item;
using xpr_batch::Table;

// Work units are dealt per XCD: workgroups are dispatched to the 8 XCDs
// round-robin by blockIdx, so the workgroups of XCD x = blockIdx % 8 take the
// units u = x + 8k, k from the phase's counter for x.  Each XCD's L2 then
// serves the same slice of the volume and detector as in a single-job launch
// (whose block b runs on XCD b % 8); within an XCD the units balance
// dynamically.  Every unit is taken exactly once whatever the real dispatch.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcd_of_block() { return blockIdx.x % kXcds; }

// next work unit of the workgroup (uniform); the barrier before the fetch
// also ends the previous unit's use of shared memory
__device__ __forceinline__ uint32_t next_unit(uint32_t *counters, uint32_t ph, uint32_t *slot)
{
    const uint32_t x = xcd_of_block();
    __syncthreads();
    if (threadIdx.x == 0) *slot = x + kXcds * atomicAdd(&counters[ph * kXcds + x], 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);
}

// The table is read-only for the kernels: reading it through the constant
// address space gives scalar loads into registers shared by the wave (the
// compiler cannot otherwise prove it unchanged by the kernels' stores and
// would read it with per-lane vector loads).
using ConstU32 = const __attribute__((address_space(4))) uint32_t;

__device__ __forceinline__ uint32_t uniform_u32(const uint32_t *p) { return *(ConstU32 *)p; }

// item of phase ph holding unit u (items are in prefix order)
__device__ __forceinline__ uint32_t item_of(const Table *__restrict__ t, uint32_t ph, uint32_t u)
{
    uint32_t i = 0;
    const uint32_t n = uniform_u32(&t->count[ph]);
    while (i + 1 < n && uniform_u32(&t->items[ph][i + 1].first) <= u) ++i;
    return i;
}

__device__ __forceinline__ Item uniform_item(const Item *__restrict__ p)
{
    constexpr int kWords = (int)(sizeof(Item) / 4);
    static_assert(sizeof(Item) % 4 == 0, "item size");
    uint32_t w[kWords];
    ConstU32 *src = (ConstU32 *)p;
#pragma unroll
    for (int i = 0; i < kWords; ++i) w[i] = src[i];
    Item it;
    __builtin_memcpy(&it, w, sizeof(Item));
    return it;
}

__device__ __forceinline__ BrickShape batch_brick_shape(const xpr_proj_args &a)
{
    BrickShape s;
    s.nx = a.vmax[0] - a.vmin[0];
    s.nl = a.vmax[1] - a.vmin[1];
    s.nk = a.vmax[2] - a.vmin[2];
    s.nbx = (s.nx + 63) / 64;
    s.nbk = (s.nk + kBK - 1) / kBK;
    s.cap = kTileCap;
    s.exact = 1;
    s.brick_base = 0;
    s.nbricks = (int64_t)s.nbx * s.nbk * ((s.nl + kWaves - 1) / kWaves);
    return s;
}

__device__ __forceinline__ SlabShape batch_slab_shape(const xpr_proj_args &a, int J)
{
    SlabShape s;
    s.nx = a.vmax[0] - a.vmin[0];
    s.nl = a.vmax[1] - a.vmin[1];
    s.nk = a.vmax[2] - a.vmin[2];
    s.nseg = (s.nx + 64 * J - 1) / (64 * J);
    s.row_base = 0;
    s.nrows = backward_rows(s.nl, s.nk, kBwdL);
    return s;
}

__global__ __launch_bounds__(kThreads) void k_batch_reset(uint32_t *counters)
{
    if (threadIdx.x < (unsigned)xpr_batch::kPhases * kXcds) counters[threadIdx.x] = 0u;
}

// One kernel per volume storage type VS, walking the forward phase of that type.
template <typename VS>
__global__ __launch_bounds__(kThreads, 4) void k_batch_forward(const Table *__restrict__ t,
                                                               uint32_t *counters, uint32_t ph)
{
    __shared__ unsigned long long lds[2 * kTileCap + kThreads];
    __shared__ ViewBin bins[kViewChunk];
    __shared__ uint32_t slot;
    const uint32_t total = uniform_u32(&t->total[ph]);
    if (total == 0) return;
    for (int i = threadIdx.x; i < 2 * kTileCap; i += kThreads) lds[i] = 0ull;
    for (;;) {
        const uint32_t u = next_unit(counters, ph, &slot);
        if (u >= total) break;                              // uniform: every workgroup ends here
        const Item it = uniform_item(&t->items[ph][item_of(t, ph, u)]);
        const BrickShape s = batch_brick_shape(it.a);
        const int64_t b = (int64_t)(u - it.first);
        forward_brick<VS, uint64_t, false>((const typename VS::T *)it.dst, (uint64_t *)it.sums, it.pd, it.a, s,
                                           b, lds, bins);
    }
}

// Backward units are the single-job launch's workgroups: segment unit % nseg
// of the kWaves adjacent wave rows (layer group, depth) of row block
// unit / nseg, one row per wave, run together so the four rows (which gather
// nearly the same detector pixels) share the CU's L1.  One kernel per volume
// storage type VS (the error samples share it).
template <typename VS>
__global__ __launch_bounds__(kThreads, 7) void k_batch_backward(const Table *__restrict__ t,
                                                                uint32_t *counters, uint32_t ph)
{
    __shared__ uint32_t slot;
    constexpr int kRec = 64 * BwdRecord<kBwdL>::kQuads;
    __shared__ float4 rec[kWaves * kRec];             // per-wave view-record tables
    float4 *wrec = rec + (threadIdx.x >> 6) * kRec;
    const uint32_t total = uniform_u32(&t->total[ph]);
    if (total == 0) return;
    for (;;) {
        const uint32_t u = next_unit(counters, ph, &slot);
        if (u >= total) break;
        const Item it = uniform_item(&t->items[ph][item_of(t, ph, u)]);
        const SlabShape s = batch_slab_shape(it.a, kBwdJ);
        const int64_t unit = (int64_t)(u - it.first);
        const int64_t row = (unit / s.nseg) * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int seg = (int)(unit % s.nseg);
        backward_row<VS, VS, kBwdJ, kBwdL, false>((typename VS::T *)it.dst, (const typename VS::T *)it.err, it.pd,
                                                  it.a, s, row, seg, nullptr, wrec);
    }
}

template <typename IS, typename ES>
__device__ __forceinline__ void error_range(const Item &it, uint64_t i0, uint64_t i1)
{
    const typename IS::T *img = (const typename IS::T *)it.src;
    const Glob<uint64_t> *sums = (const Glob<uint64_t> *)it.sums;
    typename ES::T *err = (typename ES::T *)it.dst;
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        const float ref = IS::load(img, i);
        const float got = (float)sums[i] / kQ22One;                 // FIXED_TO_FLOAT
        ES::store(err, i, logf(fmaxf(ref / fmaxf(got, kQ22Eps), kQ22Eps)) * it.weight);
    }
}

template <typename SS>
__device__ __forceinline__ void convert_range(const Item &it, uint64_t i0, uint64_t i1)
{
    const typename SS::T *src = (const typename SS::T *)it.src;
    float *dst = (float *)it.dst;
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kThreads) F32::store(dst, i, SS::load(src, i));
}

// convert (to float32), error or fill items, one chunk of kElemUnit elements per unit
__global__ __launch_bounds__(kThreads) void k_batch_elementwise(const Table *__restrict__ t,
                                                                uint32_t *counters, uint32_t ph)
{
    __shared__ uint32_t slot;
    const uint32_t total = uniform_u32(&t->total[ph]);
    if (total == 0) return;
    for (;;) {
        const uint32_t u = next_unit(counters, ph, &slot);
        if (u >= total) break;
        const Item it = uniform_item(&t->items[ph][item_of(t, ph, u)]);
        const uint64_t i0 = (uint64_t)(u - it.first) * xpr_batch::kElemUnit;
        const uint64_t i1 = min(it.n, i0 + xpr_batch::kElemUnit);
        if (ph == xpr_batch::kFill) {
            Glob<uint64_t> *dst = (Glob<uint64_t> *)it.dst;
            for (uint64_t i = i0 + threadIdx.x; i < i1; i += kThreads) dst[i] = it.value;
        } else if (ph == xpr_batch::kConvert) {
            if (it.dt_src == XPR_F16) convert_range<F16>(it, i0, i1);
            else if (it.dt_src == XPR_BF16) convert_range<BF16>(it, i0, i1);
            else convert_range<F32>(it, i0, i1);
        } else {                                            // error: image and error dtypes
            if (it.dt_src == XPR_F32 && it.dt_dst == XPR_F32) error_range<F32, F32>(it, i0, i1);
            else if (it.dt_src == XPR_F16 && it.dt_dst == XPR_F16) error_range<F16, F16>(it, i0, i1);
            else if (it.dt_src == XPR_BF16 && it.dt_dst == XPR_BF16) error_range<BF16, BF16>(it, i0, i1);
            else error_range<F32, F32>(it, i0, i1);     // other pairs are rejected by the host
        }
    }
}

} // namespace
