// Forward projector with LDS binning (KernelForward.cl:10-55).
// (library-internal; every header here is included once, by xpr_kernels.hip,
// into one translation unit)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/xpr_hip.h"
#include "common.hpp"

namespace {

// ---------------------------------------------------------------------------
// Forward with LDS binning.  A workgroup owns a brick of 64 (x) x 4 (layers,
// one per wave) x 16 (depths).  Per view the brick's 8 corners bound its
// detector footprint (a rounding bound, or one pixel of margin, covers float
// rounding, see bin_view); voxel-views accumulate there with LDS u64 atomics and the non-zero footprint pixels go
// to the detector sums with one global atomic each.  Two tiles alternate: the
// flush of view v also clears its tile, so one barrier per view separates
// splat and flush.  Footprints above the tile capacity and corners behind the
// source take the direct path (one global atomic per voxel-view); bricks
// holding Q.22 values above 2^32 take it with 64-bit values.  Every path adds
// the same integers: the sums are bit-identical whatever the path, the lane
// order or the launch order.
//
// Lane order per view.  Voxels along x project to neighbouring detector
// columns when the ray is across x, but onto one or two columns when the ray
// runs along x (views near 90 and 270 degrees): 64 lanes of one x-row would
// then hit the same LDS entries in one atomic.  Each wave therefore holds its
// 1024 voxels twice, as 64 x-lanes by 16 depths and as 16 depth-lanes x 4
// x-groups by 16 x-steps, and a view uses the order whose lanes spread more
// across columns.  The per-(layer, depth) part of the dot products is
// computed once per view by the depth-lanes; x-order reads it with readlane.
//
// The footprint, bounds thresholds, matrix rows, order and fast-path flags
// of a chunk of 32 views are computed once per workgroup (8 corner lanes per
// view) into LDS.
// ---------------------------------------------------------------------------
constexpr int kTileCap = 1536;
constexpr int kBK = 16;          // depths per brick
constexpr int kViewChunk = 32;   // views per footprint pass (256 threads / 8 corners)

struct BrickShape {
    int32_t nx, nl, nk;      // slab extent (x, layers, depths)
    int32_t nbx, nbk;        // bricks along x and depth
    int32_t cap;             // tile entries in use (<= kTileCap); 0 forces the direct path
    int32_t exact;           // 1: views may take the exact-tile splat (kBinExact)
    int64_t brick_base;      // first brick of this launch
    int64_t nbricks;         // bricks in the slab
};

template <typename AccT>
__device__ __forceinline__ void global_add(AccT *p, AccT v)
{
    __hip_atomic_fetch_add((Glob<AccT> *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

enum : uint32_t {
    kBinSkip = 1u,     // no voxel of the brick lands inside the view's bounds
    kBinTile = 2u,     // splat into an LDS tile (else direct global atomics)
    kBinFast = 4u,     // every voxel's W is inside the fast reciprocal's range
    kBinDepth = 8u,    // depth-lane order
    kBinExact = 16u,   // the tile test alone decides the canonical bounds (see bin_view)
    kBinInside = 32u,  // exact, and every voxel-view of the brick lands in the tile
};

// A view takes the depth-lane order when the brick's columns spread along
// depth at least this fraction of their spread along x.  The depth order needs
// no v_readlane per depth; below about half the spread its 16 depth lanes
// crowd too few tile entries and the LDS atomics conflict (0.35 / 0.5 / 0.7 /
// 1.0 measured 8.84 / 8.73 / 8.75 / 8.85 ms per 768^3 forward).
constexpr float kDepthOrderRatio = 0.5f;

struct ViewBin {                   // one (view, module) record of the brick, in LDS
    int32_t u0, v0, tw, th;        // tile origin and extent on the detector
    float ulo, uhi, vlo, vhi;      // in-bounds thresholds (view_thresholds)
    uint32_t flags;
    uint32_t pad[3];
    // ProjData rows 0, 1 and 3 of the matrix (record floats 0-7, 12-15): the
    // splat reads them from here with the rest of the record, not from memory
    float m[12];
};

// Footprint of the brick for view record m; run by 8 consecutive lanes (one
// per corner: bit 0 x, bit 1 layer, bit 2 depth), written by the first.
//
// The tile is the footprint (with one pixel of margin unless the rounding
// bound below holds), clipped to the columns and rows the canonical bounds
// accept (integers n with ceil(lo) <= n < ceil(hi)), so every tile entry is in
// bounds.  Exact views: the voxel
// positions are monotone in their indices, so the exact projection of every
// voxel of the brick lies between the corners' (W > 0 over the box); with
//   du = 8 eps ((Tx + U Tw) / Wmin + U),  eps = 2^-24,
// bounding the rounding of the three FMA dot products, the reciprocal and the
// product for any voxel or corner (Tx, Tw: sums of |coefficient x coordinate|
// over the box; U = max |u|; Wmin = min W), every computed u lies within 2 du
// of [umin, umax].  With 2 du <= 0.5 every in-bounds voxel-view rounds into
// the tile, and a voxel-view in the tile is in bounds: v_cvt_rpi equals roundf
// for u > -0.5, which every tile column >= 1 implies.  So when du, dv <= 0.25
// and the tile starts at column and row >= 1, the tile test alone is the
// canonical bounds test.  When in addition the viewport clips no side of the
// tile and the brick lies wholly in the slab (so the corners bound every
// voxel of it), every voxel-view rounds into the tile and is in bounds: the
// splat needs no test at all (kBinInside).
template <bool CROP>
__device__ __forceinline__ void bin_view(const float *__restrict__ pd, uint32_t m, bool live,
                                         int corner, const int32_t (&lo)[3],
                                         const int32_t (&hi)[3], int cap, bool exact_ok,
                                         bool full, const xpr_proj_args &a,
                                         ViewBin *__restrict__ out)
{
    const float *__restrict__ p = pd + 20u * m;
    const float cpx = __builtin_fmaf((float)(((corner & 1) ? hi[0] : lo[0]) + a.vmin[0]), a.voxel[0], a.origin[0]);
    const float cpy = __builtin_fmaf((float)(((corner & 2) ? hi[1] : lo[1]) + a.vmin[1]), a.voxel[1], a.origin[1]);
    const float cpz = __builtin_fmaf((float)(((corner & 4) ? hi[2] : lo[2]) + a.vmin[2]), a.voxel[2], a.origin[2]);
    const float line_off = (float)a.line_offset;
    const float X = __builtin_fmaf(p[0], cpx, __builtin_fmaf(p[1], cpy, __builtin_fmaf(p[2], cpz, p[3])));
    const float Y = __builtin_fmaf(p[4], cpx, __builtin_fmaf(p[5], cpy, __builtin_fmaf(p[6], cpz, p[7])));
    const float Wc = __builtin_fmaf(p[12], cpx, __builtin_fmaf(p[13], cpy, __builtin_fmaf(p[14], cpz, p[15])));
    const float r = 1.0f / Wc;
    const float u = X * r, v = CROP ? Y * r - line_off : Y * r;
    float umin = u, vmin = v, umax = u, vmax = v;
    int ok = (Wc > 0.0f) & (u == u) & (v == v) & (fabsf(u) < 1e7f) & (fabsf(v) < 1e7f);
    // every voxel W of the brick lies between the corner values (fast_rcp_box)
    const float tmag = fabsf(p[12] * cpx) + fabsf(p[13] * cpy) + fabsf(p[14] * cpz) + fabsf(p[15]);
    int fast = fast_rcp_box(Wc, tmag);
    // magnitudes for the rounding bound (maxima over the corners)
    float tx = fabsf(p[0] * cpx) + fabsf(p[1] * cpy) + fabsf(p[2] * cpz) + fabsf(p[3]);
    float ty = fabsf(p[4] * cpx) + fabsf(p[5] * cpy) + fabsf(p[6] * cpz) + fabsf(p[7]);
    float tw = tmag, wmin = Wc;
    // column spread per voxel step along x (corner 1) and along depth (corner 4)
    const int base = (int)(threadIdx.x & 63) & ~7;     // lane of corner 0
    const float du_x = fabsf(__shfl(u, base + 1) - __shfl(u, base)) / (float)max(hi[0] - lo[0], 1);
    const float du_k = fabsf(__shfl(u, base + 4) - __shfl(u, base)) / (float)max(hi[2] - lo[2], 1);
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        umin = fminf(umin, __shfl_xor(umin, off));
        umax = fmaxf(umax, __shfl_xor(umax, off));
        vmin = fminf(vmin, __shfl_xor(vmin, off));
        vmax = fmaxf(vmax, __shfl_xor(vmax, off));
        tx = fmaxf(tx, __shfl_xor(tx, off));
        ty = fmaxf(ty, __shfl_xor(ty, off));
        tw = fmaxf(tw, __shfl_xor(tw, off));
        wmin = fminf(wmin, __shfl_xor(wmin, off));
        ok &= __shfl_xor(ok, off);
        fast &= __shfl_xor(fast, off);
    }
    if (!live || corner != 0) return;
    ViewBin b{};
    uint32_t flags = fast ? kBinFast : 0u;
    if (ok && du_k > kDepthOrderRatio * du_x) flags |= kBinDepth;
    const float vpl = fmaxf(p[16], 0.0f), vpr = fminf(p[17], (float)a.det_width);
    const float vlo = fmaxf(p[18], 0.0f), vhi = fminf(p[19], (float)a.line_rows);
    b.ulo = lo_threshold(vpl);
    b.uhi = hi_threshold(vpr);
    b.vlo = lo_threshold(vlo);
    b.vhi = hi_threshold(vhi);
    if (ok) {
        constexpr float eps8 = 8.0f / 16777216.0f;
        const float U = fmaxf(fabsf(umin), fabsf(umax)) + 1.0f;
        const float V = fmaxf(fabsf(vmin), fabsf(vmax)) + fabsf(line_off) + 1.0f;
        const float du = eps8 * ((tx + U * tw) / wmin + U);
        const float dv = eps8 * ((ty + V * tw) / wmin + V);
        // With du, dv <= 0.25 every computed u lies in [umin - 2du, umax + 2du],
        // so its column floor(u + 0.5) lies in [floor(umin - mu + 0.5),
        // floor(umax + mu + 0.5)] with mu = 2du plus a term that covers the
        // float evaluation of these bounds (likewise rows): the tile is the
        // rounded footprint itself.  Otherwise one pixel of margin each side.
        const bool tight = exact_ok & (fast != 0) & (du <= 0.25f) & (dv <= 0.25f);
        const float mu = 2.0f * du + 1e-6f * U + (1.0f / 1024.0f);
        const float mv = 2.0f * dv + 1e-6f * V + (1.0f / 1024.0f);
        const float fu0 = tight ? floorf(umin - mu + 0.5f) : floorf(umin) - 1.0f;
        const float fv0 = tight ? floorf(vmin - mv + 0.5f) : floorf(vmin) - 1.0f;
        const float fu1 = tight ? floorf(umax + mu + 0.5f) : ceilf(umax) + 1.0f;
        const float fv1 = tight ? floorf(vmax + mv + 0.5f) : ceilf(vmax) + 1.0f;
        const bool unclipped = (fu0 >= ceilf(vpl)) & (fv0 >= ceilf(vlo)) &
                               (fu1 <= ceilf(vpr) - 1.0f) & (fv1 <= ceilf(vhi) - 1.0f);
        b.u0 = (int32_t)fmaxf(fu0, ceilf(vpl));
        b.v0 = (int32_t)fmaxf(fv0, ceilf(vlo));
        const int32_t u1 = (int32_t)fminf(fu1, ceilf(vpr) - 1.0f);
        const int32_t v1 = (int32_t)fminf(fv1, ceilf(vhi) - 1.0f);
        b.tw = u1 - b.u0 + 1;
        b.th = v1 - b.v0 + 1;
        if (b.tw <= 0 || b.th <= 0) {
            flags |= kBinSkip;
        } else if (b.tw * b.th <= cap) {
            flags |= kBinTile;
            if (tight && b.u0 >= 1 && b.v0 >= 1) {
                flags |= kBinExact;
                if (full && unclipped) flags |= kBinInside;
            }
        }
    }
    b.flags = flags;
#pragma unroll
    for (int i = 0; i < 8; ++i) b.m[i] = p[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) b.m[8 + i] = p[12 + i];
    *out = b;
}

// Workgroup barrier that orders LDS only.  The tile protocol needs no
// ordering of the global atomics of a flush, so the wave does not wait for
// them here (__syncthreads would also drain every outstanding global access).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// one brick with some Q.22 value of 2^32 or more: one global atomic per
// voxel-view with the exact 64-bit value (x-lane order)
template <typename VS, typename AccT, bool CROP>
__device__ __noinline__ void forward_direct_u64(const typename VS::T *__restrict__ vol,
                                                AccT *__restrict__ sums,
                                                const float *__restrict__ pd,
                                                const xpr_proj_args a, const BrickShape s,
                                                int32_t x0, int32_t l0, int32_t k0)
{
    const int lane = threadIdx.x & 63;
    const int32_t l = l0 + (threadIdx.x >> 6);
    if (l >= s.nl) return;
    const float py = __builtin_fmaf((float)(l + a.vmin[1]), a.voxel[1], a.origin[1]);
    const float line_off = (float)a.line_offset;
    const uint32_t view_elems = a.det_width * (uint32_t)a.line_rows;
    const int32_t x = x0 + lane;
    const float px = __builtin_fmaf((float)(x + a.vmin[0]), a.voxel[0], a.origin[0]);
    for (int i = 0; i < kBK; ++i) {
        const int32_t k = k0 + i;
        if (x >= s.nx || k >= s.nk) continue;
        const float v = VS::load(vol, ((uint64_t)l * s.nk + k) * s.nx + x);
        const AccT q = sizeof(AccT) == 8 ? (AccT)q22_u64(v) : (AccT)q22_u32(v);
        if (q == 0) continue;
        const float pz = __builtin_fmaf((float)(k + a.vmin[2]), a.voxel[2], a.origin[2]);
        for (uint32_t view = 0, m = 0; view < a.n_views; ++view) {
            AccT *vsum = sums + (uint64_t)view_of(a, view) * view_elems;
            for (uint32_t mm = 0; mm < a.modules; ++mm, ++m) {
                const float *__restrict__ p = pd + 20u * m;
                ViewConsts c;
                view_thresholds(p, a, c);
                row_terms(p, py, pz, c);
                uint32_t iu, iv;
                if (project<CROP, false>(c, px, line_off, iu, iv))
                    global_add(vsum + (__umul24(iv, a.det_width) + iu), q);
            }
        }
    }
}

// one voxel-view of a splat: its Q.22 value goes to its footprint tile entry
// (voxel-views out of bounds add 0 to the lane's sink slot); one off the tile
// goes straight to the detector sums.  Entries are counted from the tile
// base; sink_e is the sink slot's entry relative to it.
template <bool CROP, bool FAST, typename AccT>
__device__ __forceinline__ void splat_one(const ViewConsts &c, float px, uint32_t qq,
                                          float line_off, int32_t u0, int32_t v0, int32_t tw,
                                          int32_t th, bool use_tile,
                                          unsigned long long *__restrict__ tl, uint32_t sink_e,
                                          AccT *__restrict__ vsum, uint32_t det_width)
{
    uint32_t iu, iv;
    const bool inb = project<CROP, FAST>(c, px, line_off, iu, iv);
    const uint32_t tu = iu - (uint32_t)u0;
    const uint32_t tv = iv - (uint32_t)v0;
    const bool hit = inb & use_tile & (tu < (uint32_t)tw) & (tv < (uint32_t)th);
    const uint32_t e = __umul24(tv, (uint32_t)tw) + tu;
    atomicAdd(&tl[hit ? e : sink_e], (unsigned long long)(hit ? qq : 0u));
    if (inb & !hit & (qq != 0u))      // off the tile: direct global add
        global_add(vsum + (__umul24(iv, det_width) + iu), (AccT)qq);
}

// one voxel-view of an exact view (kBinExact): the tile test is the bounds test
template <bool CROP>
__device__ __forceinline__ void splat_exact(const ViewConsts &c, float px, uint32_t qq,
                                            float line_off, int32_t u0, int32_t v0, int32_t tw,
                                            int32_t th, unsigned long long *__restrict__ tl,
                                            uint32_t sink_e)
{
    const float X = __builtin_fmaf(c.c0x, px, c.k0);
    const float Y = __builtin_fmaf(c.c1x, px, c.k1);
    const float W = __builtin_fmaf(c.c3x, px, c.k3);
    const float r0 = __builtin_amdgcn_rcpf(W);
    const float r = __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
    const float u = X * r;
    const float v = CROP ? Y * r - line_off : Y * r;
    const uint32_t tu = (uint32_t)pixel_of(u) - (uint32_t)u0;
    const uint32_t tv = (uint32_t)pixel_of(v) - (uint32_t)v0;
    const bool hit = (tu < (uint32_t)tw) & (tv < (uint32_t)th);
    const uint32_t e = __umul24(tv, (uint32_t)tw) + tu;
    atomicAdd(&tl[hit ? e : sink_e], (unsigned long long)(hit ? qq : 0u));
}

using LdsU64 = __attribute__((address_space(3))) unsigned long long;

// LDS byte address of a pointer into shared memory
__device__ __forceinline__ uint32_t lds_address(const unsigned long long *p)
{
    return (uint32_t)(uintptr_t)(const LdsU64 *)p;
}

// one voxel-view of an inside view (kBinInside): it lands in the tile.  tb is
// the LDS address of the tile base minus the entry of pixel (0, 0), so the
// entry needs no offset; tw8 is the tile row in bytes.  The address is one
// shift-add and one 24-bit multiply-add (left to itself the compiler spends
// a multiply, two shifts and a three-way add).
template <bool CROP>
__device__ __forceinline__ void splat_inside(const ViewConsts &c, float px, uint32_t qq,
                                             float line_off, uint32_t tw8, uint32_t tb)
{
    const float X = __builtin_fmaf(c.c0x, px, c.k0);
    const float Y = __builtin_fmaf(c.c1x, px, c.k1);
    const float W = __builtin_fmaf(c.c3x, px, c.k3);
    const float r0 = __builtin_amdgcn_rcpf(W);
    const float r = __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
    const float u = X * r;
    const float v = CROP ? Y * r - line_off : Y * r;
    const uint32_t col = ((uint32_t)pixel_of(u) << 3) + tb;
    uint32_t addr;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(addr) : "v"((uint32_t)pixel_of(v)), "s"(tw8), "v"(col));
    __hip_atomic_fetch_add((LdsU64 *)(uintptr_t)addr, (unsigned long long)qq, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One brick of the forward (b < s.nbricks, uniform over the workgroup).  lds
// holds two tiles and one sink slot per lane ([2*kTileCap + kThreads]); the
// tiles are zero on entry and left zero.
template <typename VS, typename AccT, bool CROP>
__device__ __forceinline__ void forward_brick(const typename VS::T *__restrict__ vol,
                                              AccT *__restrict__ sums,
                                              const float *__restrict__ pd,
                                              const xpr_proj_args &a, const BrickShape &s,
                                              int64_t b, unsigned long long *__restrict__ lds,
                                              ViewBin *__restrict__ bins)
{
    const int lane = threadIdx.x & 63;
    // wave-uniform: lets the compiler keep row and flush addressing scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int32_t bx = (int32_t)(b % s.nbx);
    b /= s.nbx;
    const int32_t bk = (int32_t)(b % s.nbk);
    const int32_t bl = (int32_t)(b / s.nbk);
    const int32_t x0 = bx * 64, l0 = bl * kWaves, k0 = bk * kBK;
    const int32_t l = l0 + wave;
    const int32_t dk = lane & 15, xg = lane >> 4;     // depth-lane order

    // Q.22 voxel values in both lane orders (32 bits suffice unless some
    // value is >= 1024).  Raw buffer loads from this wave's brick corner
    // (x0, l, k0): out-of-slab voxels read 0 without a branch, so all 32
    // loads are in flight together.
    uint32_t qx[kBK], qd[kBK];
    bool big = false;
    {
        constexpr uint32_t esz = (uint32_t)sizeof(typename VS::T);
        const int64_t base = ((int64_t)l * s.nk + k0) * s.nx + x0;
        const int64_t left = (int64_t)s.nl * s.nk * s.nx - base;      // voxels from the corner on
        const uint32_t bytes = l < s.nl ? (uint32_t)min<int64_t>(left * esz, 0x7fffffff) : 0u;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(vol + (l < s.nl ? base : 0)), 0, (int)bytes, kBufferWord3);
        auto q22 = [&](float v, bool track) {
            const float sc = v * kQ22One;               // FLOAT_TO_FIXED (constants.cl:22)
            const bool pos = sc > 0.0f, over = sc >= 4294967296.0f;
            if (sizeof(AccT) == 8 && track) big |= pos & over;
            return pos ? (over ? 0xffffffffu : (uint32_t)sc) : 0u;
        };
        float vx_[kBK], vd_[kBK];
#pragma unroll
        for (int i = 0; i < kBK; ++i) {
            const bool in = (x0 + lane < s.nx) & (k0 + i < s.nk);
            const uint32_t off = (uint32_t)(i * s.nx + lane) * esz;
            vx_[i] = load_or_zero<VS>(rsrc, in ? off : kOutOfRange);
        }
#pragma unroll
        for (int j = 0; j < kBK; ++j) {
            const bool in = (x0 + xg * 16 + j < s.nx) & (k0 + dk < s.nk);
            const uint32_t off = (uint32_t)(dk * s.nx + xg * 16 + j) * esz;
            vd_[j] = load_or_zero<VS>(rsrc, in ? off : kOutOfRange);
        }
#pragma unroll
        for (int i = 0; i < kBK; ++i) qx[i] = q22(vx_[i], true);
#pragma unroll
        for (int j = 0; j < kBK; ++j) qd[j] = q22(vd_[j], false);
    }
    if (__syncthreads_or(big)) {
        forward_direct_u64<VS, AccT, CROP>(vol, sums, pd, a, s, x0, l0, k0);
        return;
    }
    const float px_x = __builtin_fmaf((float)(x0 + lane + a.vmin[0]), a.voxel[0], a.origin[0]);
    float px_d[kBK];
#pragma unroll
    for (int j = 0; j < kBK; ++j)
        px_d[j] = __builtin_fmaf((float)(x0 + xg * 16 + j + a.vmin[0]), a.voxel[0], a.origin[0]);
    const float py = __builtin_fmaf((float)(l + a.vmin[1]), a.voxel[1], a.origin[1]);
    const float pz_d = __builtin_fmaf((float)(k0 + dk + a.vmin[2]), a.voxel[2], a.origin[2]);

    // brick corners (voxel indices inside the slab)
    const int32_t lo[3] = {x0, l0, k0};
    const int32_t hi[3] = {min(x0 + 63, s.nx - 1), min(l0 + kWaves - 1, s.nl - 1),
                           min(k0 + kBK - 1, s.nk - 1)};
    const bool full = (x0 + 64 <= s.nx) & (l0 + kWaves <= s.nl) & (k0 + kBK <= s.nk);
    const float line_off = (float)a.line_offset;
    const uint32_t view_elems = a.det_width * (uint32_t)a.line_rows;
    const uint32_t nm = a.n_views * a.modules;

    int buf = 0;
    for (uint32_t c0 = 0; c0 < nm; c0 += kViewChunk) {
        const uint32_t cnt = min((uint32_t)kViewChunk, nm - c0);
        __syncthreads();                              // previous chunk's records read
        if ((uint32_t)wave * 8u < cnt) {              // a wave with no live record skips
            const uint32_t vi = threadIdx.x >> 3;
            const bool live = vi < cnt;
            bin_view<CROP>(pd, live ? c0 + vi : c0, live, threadIdx.x & 7, lo, hi, s.cap,
                           s.exact != 0, full, a, &bins[vi]);
        }
        __syncthreads();
        for (uint32_t vi = 0; vi < cnt; ++vi) {
            const uint32_t m = c0 + vi;
            const uint32_t flags = __builtin_amdgcn_readfirstlane(bins[vi].flags);
            if (flags & kBinSkip) continue;
            const int32_t u0 = __builtin_amdgcn_readfirstlane(bins[vi].u0);
            const int32_t v0 = __builtin_amdgcn_readfirstlane(bins[vi].v0);
            const int32_t tw = __builtin_amdgcn_readfirstlane(bins[vi].tw);
            const int32_t th = __builtin_amdgcn_readfirstlane(bins[vi].th);
            const bool use_tile = (flags & kBinTile) != 0;
            // (a uniform branch spares the single-module case the integer division)
            const uint32_t view =
                __builtin_amdgcn_readfirstlane(a.modules == 1 ? m : m / a.modules);
            AccT *__restrict__ vsum = sums + (uint64_t)view_of(a, view) * view_elems;
            unsigned long long *__restrict__ tl = lds + buf * kTileCap;
            const uint32_t sk = (uint32_t)((2 - buf) * kTileCap) + threadIdx.x;
            // this lane's (layer, depth) row terms: depth k0 + dk (row_terms
            // on the record's copy of the matrix)
            const ViewBin &bv = bins[vi];
            ViewConsts c;
            c.ulo = bv.ulo;
            c.uhi = bv.uhi;
            c.vlo = bv.vlo;
            c.vhi = bv.vhi;
            c.c0x = bv.m[0];
            c.c1x = bv.m[4];
            c.c3x = bv.m[8];
            c.k0 = __builtin_fmaf(bv.m[1], py, __builtin_fmaf(bv.m[2], pz_d, bv.m[3]));
            c.k1 = __builtin_fmaf(bv.m[5], py, __builtin_fmaf(bv.m[6], pz_d, bv.m[7]));
            c.k3 = __builtin_fmaf(bv.m[9], py, __builtin_fmaf(bv.m[10], pz_d, bv.m[11]));
            const uint32_t dw = a.det_width;
            if (flags & kBinInside) {
                const uint32_t tb = lds_address(tl) -
                                    8u * (__umul24((uint32_t)v0, (uint32_t)tw) + (uint32_t)u0);
                if (flags & kBinDepth) {
#pragma unroll
                    for (int j = 0; j < kBK; ++j)
                        splat_inside<CROP>(c, px_d[j], qd[j], line_off, 8u * (uint32_t)tw, tb);
                } else {
                    ViewConsts ci = c;
#pragma unroll
                    for (int i = 0; i < kBK; ++i) {
                        ci.k0 = lane_float(c.k0, i);
                        ci.k1 = lane_float(c.k1, i);
                        ci.k3 = lane_float(c.k3, i);
                        splat_inside<CROP>(ci, px_x, qx[i], line_off, 8u * (uint32_t)tw, tb);
                    }
                }
            } else if (flags & kBinExact) {
                if (flags & kBinDepth) {
#pragma unroll
                    for (int j = 0; j < kBK; ++j)
                        splat_exact<CROP>(c, px_d[j], qd[j], line_off, u0, v0, tw, th, tl, sk);
                } else {
                    ViewConsts ci = c;
#pragma unroll
                    for (int i = 0; i < kBK; ++i) {
                        ci.k0 = lane_float(c.k0, i);
                        ci.k1 = lane_float(c.k1, i);
                        ci.k3 = lane_float(c.k3, i);
                        splat_exact<CROP>(ci, px_x, qx[i], line_off, u0, v0, tw, th, tl, sk);
                    }
                }
            } else if (flags & kBinDepth) {
                if (flags & kBinFast) {
#pragma unroll
                    for (int j = 0; j < kBK; ++j)
                        splat_one<CROP, true, AccT>(c, px_d[j], qd[j], line_off, u0, v0, tw, th,
                                                    use_tile, tl, sk, vsum, dw);
                } else {
#pragma unroll
                    for (int j = 0; j < kBK; ++j)
                        splat_one<CROP, false, AccT>(c, px_d[j], qd[j], line_off, u0, v0, tw, th,
                                                     use_tile, tl, sk, vsum, dw);
                }
            } else {
                // x-lane order: the row terms of depth k0 + i sit in lane i
                ViewConsts ci = c;
                if (flags & kBinFast) {
#pragma unroll
                    for (int i = 0; i < kBK; ++i) {
                        ci.k0 = lane_float(c.k0, i);
                        ci.k1 = lane_float(c.k1, i);
                        ci.k3 = lane_float(c.k3, i);
                        splat_one<CROP, true, AccT>(ci, px_x, qx[i], line_off, u0, v0, tw, th,
                                                    use_tile, tl, sk, vsum, dw);
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < kBK; ++i) {
                        ci.k0 = lane_float(c.k0, i);
                        ci.k1 = lane_float(c.k1, i);
                        ci.k3 = lane_float(c.k3, i);
                        splat_one<CROP, false, AccT>(ci, px_x, qx[i], line_off, u0, v0, tw, th,
                                                     use_tile, tl, sk, vsum, dw);
                    }
                }
            }
            if (use_tile) {
                lds_barrier();                // splat of this view complete
                // flush and clear; the next view splats into the other tile
                // (row pointers are wave-uniform; a lane adds a 32-bit column offset)
                for (int rr = wave; rr < th; rr += kWaves) {
                    AccT *dst = vsum + (uint32_t)(v0 + rr) * dw + (uint32_t)u0;
                    unsigned long long *src = tl + rr * tw;
                    for (uint32_t cc = (uint32_t)lane; cc < (uint32_t)tw; cc += 64u) {
                        const unsigned long long t = src[cc];
                        if (t != 0ull) {
                            global_add(dst + cc, (AccT)t);
                            src[cc] = 0ull;
                        }
                    }
                }
                buf ^= 1;
            }
        }
    }
}

template <typename VS, typename AccT, bool CROP>
__global__ __launch_bounds__(kThreads, 4) void k_forward_bin(
    const typename VS::T *__restrict__ vol, AccT *__restrict__ sums,
    const float *__restrict__ pd, xpr_proj_args a, BrickShape s)
{
    // two tiles, one private slot per lane (voxel-views outside the footprint
    // add 0 there, so the splat needs no branch) and the view records
    __shared__ unsigned long long lds[2 * kTileCap + kThreads];
    __shared__ ViewBin bins[kViewChunk];
    const int64_t b = s.brick_base + blockIdx.x;
    if (b >= s.nbricks) return;                       // uniform over the block
    for (int t = threadIdx.x; t < 2 * kTileCap; t += kThreads) lds[t] = 0ull;
    forward_brick<VS, AccT, CROP>(vol, sums, pd, a, s, b, lds, bins);
}

} // namespace
