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
// Backward: one wave per (layer group, depth) row segment.  A lane owns J
// voxels along x, x = seg*64*J + lane + 64*j, in each of the L layers of its
// group.  The gather reads through an indexed buffer over the view, so
// out-of-bounds voxels read 0 without a branch and all J*L loads stay in
// flight.
//
// Layer-coherent views.  The canonical W of a voxel is fmaf(c3x, px, k3)
// with the row term k3 = fmaf(p13, py, t3) of its layer.  When k3 is the same
// float in every layer of a wave's group, so is W: W and its reciprocal are
// then computed once per lane and x position and serve the L layers.  That
// holds whenever the layer coefficient p13 (ProjData float 13) is zero (every
// detector the reference builds) and also when it is the rounding residue of
// a roll about the beam axis (|p13 py| below half an ulp of t3).  (Equal
// floats may differ in the sign of a zero, and a W of +0 or -0 leaves the
// voxel out of bounds either way.)  When the u row terms k0 agree across the
// group as well (an upright detector), so does the column u (a zero X gives
// the same pixel and bounds result with either sign).  Other groups take the
// canonical projection per voxel.
//
// Bilinear sampling (opt-in, BILIN): the same projection and the same
// voxel-views as the nearest-pixel test accepts, each sampling the error
// image at the continuous (u, v): taps columns floor u and floor u + 1 and
// rows floor v and floor v + 1, each clamped to the accepted pixels
// [ceil(vpl), ceil(vpr) - 1] (rows likewise; the bounds clamped to the
// detector and the rows held), weights fx = u - floor u, fy = v - floor v,
// three lerps:
//   top = fmaf(fx, E01 - E00, E00), bot = fmaf(fx, E11 - E10, E10),
//   sample = fmaf(fy, bot - top, top)     (E_rc: row +r, column +c)
// so equal taps give that value exactly, and an edge pixel's value holds over
// the half pixel beyond its centre.  The accepted pixels come from the
// thresholds: ceil(ulo) and floor(uhi).
// ---------------------------------------------------------------------------
// LDS record of one view for one wave: k3 (the w row term at the group's first
// layer), the x coefficients of rows 0, 1, 3, and the L row terms of u and of
// v, padded to float4s
template <int L>
struct BwdRecord {
    static constexpr int kFloats = 4 + 2 * L;
    static constexpr int kQuads = (kFloats + 3) / 4;
};

// Set-up of one (row segment, view record) pair, done by one lane per record.
//
// "Inside" rests on a rounding bound.  The exact X, Y, W of a voxel are affine
// in its position, so with W > 0 over the segment's box (x ends by the
// group's first and last layer at one depth) the exact u = X/W and v = Y/W lie
// between their corner values.  With
//   du = 8 eps ((Tx + U Tw) / Wmin + U),  eps = 2^-24,
// bounding the rounding of the three FMA dot products, the reciprocal and the
// product for any voxel or corner (Tx, Tw: sums of |coefficient| x the largest
// |coordinate| of the box; U = max |u| + 1; Wmin = min W), every computed
// u lies within 2 du of the computed corners' range (likewise v with dv).  A
// segment is inside when that range, widened by 2 du plus a term that covers
// the float evaluation of the bound, passes the canonical half-open test:
// every voxel-view of the segment then passes it too, so its test is skipped
// without changing any result.
template <int L>
struct RowSetup {
    float k3, t3, p13;                   // w row term at the first layer; pz part; coefficient
    float k0[L], k1[L];
    float ulo, uhi, vlo, vhi;
    bool fast, inside, wcol, ucol;
};

template <int L>
__device__ __forceinline__ RowSetup<L> row_setup(const float *__restrict__ p,
                                                 const xpr_proj_args &a, const float (&py)[L],
                                                 float pz, float px_first, float px_last,
                                                 float line_off, bool inside_ok)
{
    RowSetup<L> r;
    const float c0x = p[0], c1x = p[4], c3x = p[12];
    r.p13 = p[13];
    const float t0 = __builtin_fmaf(p[2], pz, p[3]);
    const float t1 = __builtin_fmaf(p[6], pz, p[7]);
    r.t3 = __builtin_fmaf(p[14], pz, p[15]);
    r.k3 = __builtin_fmaf(p[13], py[0], r.t3);
    r.wcol = true;
    r.ucol = true;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        r.k0[i] = __builtin_fmaf(p[1], py[i], t0);
        r.k1[i] = __builtin_fmaf(p[5], py[i], t1);
        r.wcol &= __builtin_fmaf(p[13], py[i], r.t3) == r.k3;
        r.ucol &= r.k0[i] == r.k0[0];
    }
    r.ucol &= r.wcol;
    r.ulo = lo_threshold(fmaxf(p[16], 0.0f));
    r.uhi = hi_threshold(fminf(p[17], (float)a.det_width));
    r.vlo = lo_threshold(fmaxf(p[18], 0.0f));
    r.vhi = hi_threshold(fminf(p[19], (float)a.line_rows));
    // magnitudes for the rounding bound: sums of |coefficient x coordinate|,
    // bounded over the box by the largest |coordinate| of each axis
    const float mx = fmaxf(fabsf(px_first), fabsf(px_last));
    const float my = fmaxf(fabsf(py[0]), fabsf(py[L - 1]));
    const float tx = fabsf(c0x) * mx + fabsf(p[1]) * my + fabsf(p[2] * pz) + fabsf(p[3]);
    const float ty = fabsf(c1x) * mx + fabsf(p[5]) * my + fabsf(p[6] * pz) + fabsf(p[7]);
    const float tw = fabsf(c3x) * mx + fabsf(p[13]) * my + fabsf(p[14] * pz) + fabsf(p[15]);
    // the four corners of the segment's box, computed like a voxel's
    float umin = 3.4e38f, umax = -3.4e38f, vmin = 3.4e38f, vmax = -3.4e38f;
    float wmin = 3.4e38f, ymax = 0.0f;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float cx = (c & 1) ? px_last : px_first;
        const float cy = (c & 2) ? py[L - 1] : py[0];
        const float k0 = (c & 2) ? r.k0[L - 1] : r.k0[0];
        const float k1 = (c & 2) ? r.k1[L - 1] : r.k1[0];
        const float k3 = __builtin_fmaf(p[13], cy, r.t3);
        const float X = __builtin_fmaf(c0x, cx, k0);
        const float Y = __builtin_fmaf(c1x, cx, k1);
        const float W = __builtin_fmaf(c3x, cx, k3);
        const float q0 = __builtin_amdgcn_rcpf(W);
        const float q = __builtin_fmaf(__builtin_fmaf(-W, q0, 1.0f), q0, q0);
        const float u = X * q, y = Y * q, v = y - line_off;
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        vmin = fminf(vmin, v);
        vmax = fmaxf(vmax, v);
        ymax = fmaxf(ymax, fabsf(y));
        wmin = fminf(wmin, W);
        ok &= fast_rcp_box(W, tw);
    }
    // every voxel W of the box lies between the corner values (fast_rcp_box)
    r.fast = ok;
    ok &= (umin == umin) & (vmin == vmin) & (umax == umax) & (vmax == vmax);
    constexpr float eps8 = 8.0f / 16777216.0f;
    const float U = fmaxf(fabsf(umin), fabsf(umax)) + 1.0f;
    const float V = ymax + fabsf(line_off) + 1.0f;
    const float iw = 1.0f / wmin;
    const float du = eps8 * ((tx + U * tw) * iw + U);
    const float dv = eps8 * ((ty + V * tw) * iw + V);
    const float mu = 2.0f * du + 1e-6f * U + (1.0f / 1024.0f);
    const float mv = 2.0f * dv + 1e-6f * V + (1.0f / 1024.0f);
    r.inside = inside_ok & ok & r.wcol & (umin - mu >= r.ulo) & (umax + mu < r.uhi) &
               (vmin - mv >= r.vlo) & (vmax + mv < r.vhi);
    return r;
}

// One bilinear sample of the view's error image at (u, v), taps clamped to
// the accepted columns [c0, c1] and rows [r0, r1].  The taps are read through
// the indexed buffer, whose bound makes any index safe; the caller discards
// the sample of a voxel-view that is out of bounds.
template <typename ES>
__device__ __forceinline__ float bilinear_sample(__amdgpu_buffer_rsrc_t rsrc, float u, float v,
                                                 uint32_t dw, float c0, float c1, float r0,
                                                 float r1)
{
    const float fu = floorf(u), fv = floorf(v);
    const float fx = u - fu, fy = v - fv;       // exact: floor(u) is within a factor 2 of u
    const uint32_t i0 = (uint32_t)(int32_t)fmaxf(fu, c0), i1 = (uint32_t)(int32_t)fminf(fu + 1.0f, c1);
    const uint32_t j0 = __umul24((uint32_t)(int32_t)fmaxf(fv, r0), dw);
    const uint32_t j1 = __umul24((uint32_t)(int32_t)fminf(fv + 1.0f, r1), dw);
    const float e00 = sample_f32<ES>(rsrc, j0 + i0), e01 = sample_f32<ES>(rsrc, j0 + i1);
    const float e10 = sample_f32<ES>(rsrc, j1 + i0), e11 = sample_f32<ES>(rsrc, j1 + i1);
    const float top = __builtin_fmaf(fx, e01 - e00, e00);
    const float bot = __builtin_fmaf(fx, e11 - e10, e10);
    return __builtin_fmaf(fy, bot - top, top);
}

// Segment seg of one (layer group, depth) row of the backward, run by one
// wave.  The per-view set-up is spread over the lanes (one view record each,
// 64 at a time).  The view loop reads each record back from the wave's LDS
// table as broadcast loads straight into VGPRs; LDS operations of a wave run
// in order, so the table needs no barrier.  Every branch of a view leaves its
// samples in smp[] and one accumulation follows, in view order, so the sums
// equal the canonical ones.  max_bits (optional, XPR_MAX_SLOTS entries): the
// slot of this wave is raised to the bits of the row's largest voxel value
// after the update (updated or kept; before a 16-bit store rounds it).
template <typename VS, typename ES, int J, int L, bool CROP, bool BILIN = false>
__device__ __forceinline__ void backward_row(typename VS::T *__restrict__ vol,
                                             const typename ES::T *__restrict__ err,
                                             const float *__restrict__ pd,
                                             const xpr_proj_args &a, const SlabShape &s,
                                             int64_t row, int seg,
                                             unsigned long long *__restrict__ empty_count,
                                             float4 *__restrict__ rec,
                                             uint32_t *__restrict__ max_bits = nullptr)
{
    using R = BwdRecord<L>;
    const int lane = threadIdx.x & 63;
    if (row >= s.nrows) return;                       // whole wave leaves together
    const int32_t g = (int32_t)(row / s.nk), k = (int32_t)(row % s.nk);
    const int32_t l0 = g * L;
    const int32_t nlv = min(L, s.nl - l0);           // layers of the group in the slab
    float py[L];
#pragma unroll
    for (int i = 0; i < L; ++i)                       // a short last group repeats its last layer
        py[i] = __builtin_fmaf((float)(l0 + min(i, nlv - 1) + a.vmin[1]), a.voxel[1], a.origin[1]);
    const float pz = __builtin_fmaf((float)(k + a.vmin[2]), a.voxel[2], a.origin[2]);
    const int32_t xb = seg * 64 * J + lane;
    float px[J], sum[J][L];
    uint32_t cnt[J][L];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        px[j] = __builtin_fmaf((float)(xb + 64 * j + a.vmin[0]), a.voxel[0], a.origin[0]);
#pragma unroll
        for (int i = 0; i < L; ++i) {
            sum[j][i] = 0.0f;
            cnt[j][i] = 0;
        }
    }
    // first and last voxel positions of the row segment
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
        const float *__restrict__ pl = pd + 20u * min(m0 + (uint32_t)lane, nrec - 1);
        const RowSetup<L> su = row_setup<L>(pl, a, py, pz, px_first, px_last, line_off,
                                            s.inside != 0);
        const uint64_t fast_mask = __builtin_amdgcn_ballot_w64(su.fast);
        const uint64_t inside_mask = __builtin_amdgcn_ballot_w64(su.inside);
        const uint64_t wcol_mask = __builtin_amdgcn_ballot_w64(su.wcol);
        const uint64_t ucol_mask = __builtin_amdgcn_ballot_w64(su.ucol);
        const uint32_t n_in = min(64u, nrec - m0);
        {
            float f[4 * R::kQuads];
            f[0] = su.k3;
            f[1] = pl[0];
            f[2] = pl[4];
            f[3] = pl[12];
#pragma unroll
            for (int i = 0; i < L; ++i) {
                f[4 + i] = su.k0[i];
                f[4 + L + i] = su.k1[i];
            }
#pragma unroll
            for (int i = R::kFloats; i < 4 * R::kQuads; ++i) f[i] = 0.0f;
#pragma unroll
            for (int q = 0; q < R::kQuads; ++q)
                rec[R::kQuads * lane + q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
        }
        for (uint32_t i = 0; i < n_in; ++i) {
            float f[4 * R::kQuads];
#pragma unroll
            for (int q = 0; q < R::kQuads; ++q) {
                const float4 t = rec[R::kQuads * i + q];
                f[4 * q] = t.x;
                f[4 * q + 1] = t.y;
                f[4 * q + 2] = t.z;
                f[4 * q + 3] = t.w;
            }
            const float k3 = f[0], c0x = f[1], c1x = f[2], c3x = f[3];
            const float *k0 = f + 4, *k1 = f + 4 + L;
            const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(err + (uint64_t)view_of(a, view) * view_elems), (short)esz,
                (int)view_elems, kBufferWord3);
            const uint32_t dw = a.det_width;
            if constexpr (BILIN) {
                const bool inside = (inside_mask >> i) & 1u;
                n_inside += inside ? 1u : 0u;
                const float ulo = lane_float(su.ulo, i), uhi = lane_float(su.uhi, i);
                const float vlo = lane_float(su.vlo, i), vhi = lane_float(su.vhi, i);
                // the accepted pixels: columns ceil(ulo) .. floor(uhi), rows likewise
                const float c0 = ceilf(ulo), c1 = floorf(uhi), r0 = ceilf(vlo), r1 = floorf(vhi);
                const bool fast = (fast_mask >> i) & 1u;
                const bool wcol = (wcol_mask >> i) & 1u;
                const float p13 = lane_float(su.p13, i), t3 = lane_float(su.t3, i);
                auto rcp = [&](float W) {
                    if (fast) {
                        const float r0 = __builtin_amdgcn_rcpf(W);
                        return __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
                    }
                    return recip(W);
                };
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    const float rw = wcol ? rcp(__builtin_fmaf(c3x, px[j], k3)) : 0.0f;
#pragma unroll
                    for (int l = 0; l < L; ++l) {
                        const float r = wcol ? rw
                                             : rcp(__builtin_fmaf(c3x, px[j],
                                                                  __builtin_fmaf(p13, py[l], t3)));
                        const float u = __builtin_fmaf(c0x, px[j], k0[l]) * r;
                        const float Y = __builtin_fmaf(c1x, px[j], k1[l]);
                        const float v = CROP ? Y * r - line_off : Y * r;
                        const float smp = bilinear_sample<ES>(rsrc, u, v, dw, c0, c1, r0, r1);
                        if (inside) {
                            sum[j][l] = sum[j][l] + smp;
                        } else {
                            const bool inb = (u >= ulo) & (u < uhi) & (v >= vlo) & (v < vhi);
                            sum[j][l] = sum[j][l] + (inb ? smp : 0.0f);
                            cnt[j][l] += inb ? 1u : 0u;
                        }
                    }
                }
                if (++mm == a.modules) {
                    mm = 0;
                    ++view;
                }
                continue;
            }
            typename SampleOf<ES>::type smp[J][L];
            if ((inside_mask >> i) & 1u) {
                ++n_inside;
                if ((ucol_mask >> i) & 1u) {
                    // upright: one column per x position
#pragma unroll
                    for (int j = 0; j < J; ++j) {
                        const float X = __builtin_fmaf(c0x, px[j], k0[0]);
                        const float W = __builtin_fmaf(c3x, px[j], k3);
                        const float r0 = __builtin_amdgcn_rcpf(W);
                        const float r = __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
                        const uint32_t iu = (uint32_t)pixel_of(X * r);
#pragma unroll
                        for (int l = 0; l < L; ++l) {
                            const float Y = __builtin_fmaf(c1x, px[j], k1[l]);
                            const float v = CROP ? Y * r - line_off : Y * r;
                            smp[j][l] = sample_raw<ES>(rsrc, __umul24((uint32_t)pixel_of(v), dw) + iu);
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < J; ++j) {
                        const float W = __builtin_fmaf(c3x, px[j], k3);
                        const float r0 = __builtin_amdgcn_rcpf(W);
                        const float r = __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
#pragma unroll
                        for (int l = 0; l < L; ++l) {
                            const float X = __builtin_fmaf(c0x, px[j], k0[l]);
                            const float Y = __builtin_fmaf(c1x, px[j], k1[l]);
                            const float v = CROP ? Y * r - line_off : Y * r;
                            smp[j][l] = sample_raw<ES>(
                                rsrc, __umul24((uint32_t)pixel_of(v), dw) + (uint32_t)pixel_of(X * r));
                        }
                    }
                }
            } else {
                ViewConsts c;
                c.c0x = c0x;
                c.c1x = c1x;
                c.c3x = c3x;
                c.ulo = lane_float(su.ulo, i);
                c.uhi = lane_float(su.uhi, i);
                c.vlo = lane_float(su.vlo, i);
                c.vhi = lane_float(su.vhi, i);
                const bool fast = (fast_mask >> i) & 1u;
                if ((wcol_mask >> i) & 1u) {
#pragma unroll
                    for (int j = 0; j < J; ++j) {
                        const float W = __builtin_fmaf(c3x, px[j], k3);
                        float r;
                        if (fast) {
                            const float r0 = __builtin_amdgcn_rcpf(W);
                            r = __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
                        } else {
                            r = recip(W);
                        }
#pragma unroll
                        for (int l = 0; l < L; ++l) {
                            const float u = __builtin_fmaf(c0x, px[j], k0[l]) * r;
                            const float Y = __builtin_fmaf(c1x, px[j], k1[l]);
                            const float v = CROP ? Y * r - line_off : Y * r;
                            const bool inb = (u >= c.ulo) & (u < c.uhi) & (v >= c.vlo) & (v < c.vhi);
                            const uint32_t idx = __umul24((uint32_t)pixel_of(v), dw) + (uint32_t)pixel_of(u);
                            smp[j][l] = sample_raw<ES>(rsrc, inb ? idx : kNoSample);
                            cnt[j][l] += inb ? 1u : 0u;
                        }
                    }
                } else {
                    // canonical projection per voxel: the w row term of every layer
                    const float p13 = lane_float(su.p13, i), t3 = lane_float(su.t3, i);
#pragma unroll
                    for (int l = 0; l < L; ++l) {
                        c.k0 = k0[l];
                        c.k1 = k1[l];
                        c.k3 = __builtin_fmaf(p13, py[l], t3);
#pragma unroll
                        for (int j = 0; j < J; ++j) {
                            uint32_t iu, iv;
                            const bool inb = fast ? project<CROP, true>(c, px[j], line_off, iu, iv)
                                                  : project<CROP, false>(c, px[j], line_off, iu, iv);
                            const uint32_t idx = __umul24(iv, dw) + iu;
                            smp[j][l] = sample_raw<ES>(rsrc, inb ? idx : kNoSample);
                            cnt[j][l] += inb ? 1u : 0u;
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < J; ++j)
#pragma unroll
                for (int l = 0; l < L; ++l) sum[j][l] = add_sample<ES>(sum[j][l], smp[j][l]);
            if (++mm == a.modules) {
                mm = 0;
                ++view;
            }
        }
    }
    uint32_t empty = 0, top = 0;           // top: the largest value's bits (max_bits)
    const bool track = max_bits != nullptr;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        if (l >= nlv) break;                          // uniform
        const uint64_t rowoff = ((uint64_t)(l0 + l) * (uint64_t)s.nk + (uint64_t)k) * (uint64_t)s.nx;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int32_t x = xb + 64 * j;
            const uint32_t n = cnt[j][l] + n_inside;
            if (x < s.nx) {
                if (n == 0) {
                    ++empty;              // the reference computes exp(0/0) here
                    if (track) top = max(top, __float_as_uint(VS::load(vol, rowoff + x)));
                } else {
                    const float v = VS::load(vol, rowoff + x);
                    const float nv = v * expf(sum[j][l] / (float)n);
                    VS::store(vol, rowoff + x, nv);
                    top = max(top, __float_as_uint(nv));
                }
            }
        }
    }
    if (empty_count != nullptr) {
        uint32_t e = empty;
        for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
        if (lane == 0 && e) atomicAdd(empty_count, (unsigned long long)e);
    }
    if (track) {
        for (int off = 32; off > 0; off >>= 1) top = max(top, (uint32_t)__shfl_xor((int)top, off));
        static_assert((XPR_MAX_SLOTS & (XPR_MAX_SLOTS - 1)) == 0, "slots: a power of two");
        const uint32_t slot = (uint32_t)row & (XPR_MAX_SLOTS - 1u);
        if (lane == 0) atomicMax(max_bits + slot, top);
    }
}

// wave rows of a slab: (layer group, depth) pairs
__host__ __device__ __forceinline__ int64_t backward_rows(int32_t nl, int32_t nk, int L)
{
    return (int64_t)((nl + L - 1) / L) * nk;
}

template <typename VS, typename ES, int J, int L, bool CROP, bool BILIN = false>
__global__ __launch_bounds__(kThreads) void k_backward(
    typename VS::T *__restrict__ vol, const typename ES::T *__restrict__ err,
    const float *__restrict__ pd, xpr_proj_args a, SlabShape s,
    unsigned long long *__restrict__ empty_count, uint32_t *__restrict__ max_bits)
{
    __shared__ float4 rec[kWaves * 64 * BwdRecord<L>::kQuads];   // per-wave view-record tables
    // segment blockIdx % nseg of the kWaves rows of row block blockIdx / nseg
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    const int64_t row = s.row_base + (int64_t)(blockIdx.x / s.nseg) * kWaves + wave;
    backward_row<VS, ES, J, L, CROP, BILIN>(vol, err, pd, a, s, row, (int)(blockIdx.x % s.nseg),
                                            empty_count,
                                            rec + (threadIdx.x >> 6) * 64 * BwdRecord<L>::kQuads,
                                            max_bits);
}

} // namespace
