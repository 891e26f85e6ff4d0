// Backward projector with the multiplicative update (KernelBackward.cl:8-41).
// (library-internal; every header here is included once, by xpr_kernels.hip,
// into one translation unit)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/xpr_hip.h"
#include "common.hpp"

namespace {

// ---------------------------------------------------------------------------
// Backward: one wave per (layer, depth) row segment, lane owns J voxels at
// x = seg*64*J + lane + 64*j.  Per view the (y, z, w) part of the three dot
// products is computed once per row; a voxel-view then costs 3 FMA, the
// reciprocal (2 more FMA for its Newton step), 2 MUL, 2 conversions, the
// index, one gather and the add, plus 4 compares and the hit count in views
// that do not hold the whole segment inside.  The gather reads through an
// indexed buffer over the view, so out-of-bounds voxels read 0 without a
// branch and all J loads stay in flight.
// ---------------------------------------------------------------------------
// Projection of a voxel known to be in bounds (fast reciprocal range checked).
template <bool CROP>
__device__ __forceinline__ void project_inside(const ViewConsts &c, float px, float line_off,
                                               uint32_t &iu, uint32_t &iv)
{
    const float X = __builtin_fmaf(c.c0x, px, c.k0);
    const float Y = __builtin_fmaf(c.c1x, px, c.k1);
    const float W = __builtin_fmaf(c.c3x, px, c.k3);
    const float r0 = __builtin_amdgcn_rcpf(W);
    const float r = __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
    const float u = X * r;
    const float v = CROP ? Y * r - line_off : Y * r;
    iu = (uint32_t)pixel_of(u);
    iv = (uint32_t)pixel_of(v);
}

__device__ __forceinline__ float lane_float(float x, uint32_t lane)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
}

// Set-up of one (row segment, view record) pair, done by one lane per record:
// the row terms, the thresholds, whether the fast reciprocal is exact over the
// segment, and whether every voxel of the segment is in bounds.
//
// "Inside" rests on a rounding bound.  At fixed (py, pz) the exact X, Y, W of
// the evaluated FMAs are affine in px, so with W > 0 the exact u = X/W and
// v = Y/W are monotone along the segment and lie between their values at the
// segment's ends.  Each computed u carries the relative error of one FMA (X),
// one FMA (W), the correctly rounded reciprocal and one product: at most
// 4 * 2^-24 (1 + 2^-22), and v one subtraction of the line offset more.  The
// segment is inside when its computed end values, widened by 2^-20 (U + 1)
// (U the largest end |u|, about four times the two ends' errors together),
// pass the canonical half-open test.  Every voxel of the segment then passes
// it too, so its test can be skipped without changing any result.
struct RowSetup {
    float k0, k1, k3, ulo, uhi, vlo, vhi;
    bool fast, inside;
};

__device__ __forceinline__ RowSetup row_setup(const float *__restrict__ p, const xpr_proj_args &a,
                                              float py, float pz, float px_first, float px_last,
                                              float line_off, bool inside_ok)
{
    RowSetup r;
    const float c0x = p[0], c1x = p[4], c3x = p[12];
    r.k0 = __builtin_fmaf(p[1], py, __builtin_fmaf(p[2], pz, p[3]));
    r.k1 = __builtin_fmaf(p[5], py, __builtin_fmaf(p[6], pz, p[7]));
    r.k3 = __builtin_fmaf(p[13], py, __builtin_fmaf(p[14], pz, p[15]));
    r.ulo = lo_threshold(fmaxf(p[16], 0.0f));
    r.uhi = hi_threshold(fminf(p[17], (float)a.det_width));
    r.vlo = lo_threshold(fmaxf(p[18], 0.0f));
    r.vhi = hi_threshold(fminf(p[19], (float)a.line_rows));
    const float w0 = __builtin_fmaf(c3x, px_first, r.k3);
    const float w1 = __builtin_fmaf(c3x, px_last, r.k3);
    const float t = fabsf(c3x) * fmaxf(fabsf(px_first), fabsf(px_last)) + fabsf(p[13] * py) +
                    fabsf(p[14] * pz) + fabsf(p[15]);
    r.fast = fast_rcp_box(w0, t) & fast_rcp_box(w1, t);
    // end values, computed exactly like a voxel's (fast reciprocal)
    const float q0 = __builtin_fmaf(__builtin_fmaf(-w0, __builtin_amdgcn_rcpf(w0), 1.0f),
                                    __builtin_amdgcn_rcpf(w0), __builtin_amdgcn_rcpf(w0));
    const float q1 = __builtin_fmaf(__builtin_fmaf(-w1, __builtin_amdgcn_rcpf(w1), 1.0f),
                                    __builtin_amdgcn_rcpf(w1), __builtin_amdgcn_rcpf(w1));
    const float u0 = __builtin_fmaf(c0x, px_first, r.k0) * q0;
    const float u1 = __builtin_fmaf(c0x, px_last, r.k0) * q1;
    const float y0 = __builtin_fmaf(c1x, px_first, r.k1) * q0;
    const float y1 = __builtin_fmaf(c1x, px_last, r.k1) * q1;
    const float v0 = y0 - line_off, v1 = y1 - line_off;
    constexpr float kMargin = 1.0f / 1048576.0f;
    const float mu = kMargin * (fmaxf(fabsf(u0), fabsf(u1)) + 1.0f);
    const float mv = kMargin * (fmaxf(fabsf(y0), fabsf(y1)) + fabsf(line_off) + 1.0f);
    r.inside = inside_ok & r.fast & (fminf(u0, u1) - mu >= r.ulo) & (fmaxf(u0, u1) + mu < r.uhi) &
               (fminf(v0, v1) - mv >= r.vlo) & (fmaxf(v0, v1) + mv < r.vhi);
    return r;
}

// Segment seg of one (layer, depth) row of the backward, run by one wave.
// The per-view set-up of the row is spread over the lanes (one view record
// each, 64 at a time).  The view loop reads each record's row terms and
// matrix coefficients back from the wave's LDS table rec[128] as broadcast
// loads straight into VGPRs (v_readlane would need three reads and three
// copies, since an FMA takes one scalar operand); LDS operations of a wave
// run in order, so the table needs no barrier.  Both branches of a view leave
// their J samples in smp[] and one accumulation follows, in view order, so
// the sums equal the canonical ones and no register copies join the branches.
template <typename VS, typename ES, int J, bool CROP>
__device__ __forceinline__ void backward_row(typename VS::T *__restrict__ vol,
                                             const typename ES::T *__restrict__ err,
                                             const float *__restrict__ pd,
                                             const xpr_proj_args &a, const SlabShape &s,
                                             int64_t row, int seg,
                                             unsigned long long *__restrict__ empty_count,
                                             float4 *__restrict__ rec)
{
    const int lane = threadIdx.x & 63;
    if (row >= s.nrows) return;                       // whole wave leaves together
    const int32_t l = (int32_t)(row / s.nk), k = (int32_t)(row % s.nk);
    const float py = __builtin_fmaf((float)(l + a.vmin[1]), a.voxel[1], a.origin[1]);
    const float pz = __builtin_fmaf((float)(k + a.vmin[2]), a.voxel[2], a.origin[2]);
    const int32_t xb = seg * 64 * J + lane;
    float px[J], sum[J];
    uint32_t cnt[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        px[j] = __builtin_fmaf((float)(xb + 64 * j + a.vmin[0]), a.voxel[0], a.origin[0]);
        sum[j] = 0.0f;
        cnt[j] = 0;
    }
    // first and last voxel positions of the row segment (they bound u, v, W over it)
    const float px_first = __builtin_fmaf((float)(seg * 64 * J + a.vmin[0]), a.voxel[0], a.origin[0]);
    const float px_last = __builtin_fmaf((float)(min(seg * 64 * J + 64 * J, s.nx) - 1 + a.vmin[0]),
                                         a.voxel[0], a.origin[0]);
    const float line_off = (float)a.line_offset;
    const uint32_t view_elems = a.det_width * (uint32_t)a.line_rows;
    const uint32_t esz = (uint32_t)sizeof(typename ES::T);
    const uint32_t nrec = a.n_views * a.modules;
    uint32_t n_inside = 0;                  // records in which every voxel is in bounds
    uint32_t view = 0, mm = 0;
    for (uint32_t m0 = 0; m0 < nrec; m0 += 64) {
        const RowSetup su = row_setup(pd + 20u * min(m0 + (uint32_t)lane, nrec - 1), a, py, pz,
                                      px_first, px_last, line_off, s.inside != 0);
        const uint64_t fast_mask = __builtin_amdgcn_ballot_w64(su.fast);
        const uint64_t inside_mask = __builtin_amdgcn_ballot_w64(su.inside);
        const uint32_t n_in = min(64u, nrec - m0);
        {
            const float *__restrict__ pl = pd + 20u * min(m0 + (uint32_t)lane, nrec - 1);
            rec[2 * lane] = make_float4(su.k0, su.k1, su.k3, pl[0]);
            rec[2 * lane + 1] = make_float4(pl[4], pl[12], 0.0f, 0.0f);
        }
        for (uint32_t i = 0; i < n_in; ++i) {
            const float4 r = rec[2 * i], r2 = rec[2 * i + 1];
            const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(err + (uint64_t)view_of(a, view) * view_elems), (short)esz,
                (int)view_elems, kBufferWord3);
            ViewConsts c;
            c.c0x = r.w;
            c.c1x = r2.x;
            c.c3x = r2.y;
            c.k0 = r.x;
            c.k1 = r.y;
            c.k3 = r.z;
            float smp[J];
            if ((inside_mask >> i) & 1u) {
                ++n_inside;
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    uint32_t iu, iv;
                    project_inside<CROP>(c, px[j], line_off, iu, iv);
                    smp[j] = sample_or_zero<ES>(rsrc, __umul24(iv, a.det_width) + iu);
                }
            } else {
                c.ulo = lane_float(su.ulo, i);
                c.uhi = lane_float(su.uhi, i);
                c.vlo = lane_float(su.vlo, i);
                c.vhi = lane_float(su.vhi, i);
                if ((fast_mask >> i) & 1u) {
#pragma unroll
                    for (int j = 0; j < J; ++j) {
                        uint32_t iu, iv;
                        const bool inb = project<CROP, true>(c, px[j], line_off, iu, iv);
                        const uint32_t idx = __umul24(iv, a.det_width) + iu;
                        smp[j] = sample_or_zero<ES>(rsrc, inb ? idx : kNoSample);
                        cnt[j] += inb ? 1u : 0u;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < J; ++j) {
                        uint32_t iu, iv;
                        const bool inb = project<CROP, false>(c, px[j], line_off, iu, iv);
                        const uint32_t idx = __umul24(iv, a.det_width) + iu;
                        smp[j] = sample_or_zero<ES>(rsrc, inb ? idx : kNoSample);
                        cnt[j] += inb ? 1u : 0u;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < J; ++j) sum[j] += smp[j];
            if (++mm == a.modules) {
                mm = 0;
                ++view;
            }
        }
    }
    uint32_t empty = 0;
    const uint64_t rowoff = (uint64_t)row * (uint64_t)s.nx;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int32_t x = xb + 64 * j;
        const uint32_t n = cnt[j] + n_inside;
        if (x < s.nx) {
            if (n == 0) {
                ++empty;              // the reference computes exp(0/0) here
            } else {
                const float v = VS::load(vol, rowoff + x);
                VS::store(vol, rowoff + x, v * expf(sum[j] / (float)n));
            }
        }
    }
    if (empty_count != nullptr) {
        uint32_t e = empty;
        for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
        if (lane == 0 && e) atomicAdd(empty_count, (unsigned long long)e);
    }
}

template <typename VS, typename ES, int J, bool CROP>
__global__ __launch_bounds__(kThreads) void k_backward(
    typename VS::T *__restrict__ vol, const typename ES::T *__restrict__ err,
    const float *__restrict__ pd, xpr_proj_args a, SlabShape s,
    unsigned long long *__restrict__ empty_count)
{
    __shared__ float4 rec[kWaves * 128];              // per-wave view-record tables
    // segment blockIdx % nseg of the kWaves rows of row block blockIdx / nseg
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    const int64_t row = s.row_base + (int64_t)(blockIdx.x / s.nseg) * kWaves + wave;
    backward_row<VS, ES, J, CROP>(vol, err, pd, a, s, row, (int)(blockIdx.x % s.nseg), empty_count,
                                  rec + (threadIdx.x >> 6) * 128);
}

} // namespace
