// Forward projector with LDS binning (KernelForward.cl:10-55).
// (library-internal; every header here is included once, by xpr_kernels.hip,
// into one translation unit)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../../include/xpr_hip.h"
#include "common.hpp"

namespace {

// ---------------------------------------------------------------------------
// Forward with LDS binning.  A workgroup owns a brick of 64 (x) x 4 (layers)
// x 16 (depths).  Per view the brick's 8 corners bound its detector footprint
// (a rounding bound, or one pixel of margin, covers float rounding, see
// bin_view); voxel-views accumulate there with LDS integer atomics and the
// non-zero footprint pixels go to the detector sums with one global atomic
// each.  Two tiles alternate: the flush of view v also clears its tile, so
// one barrier per view separates splat and flush.  Footprints above the tile
// capacity and corners behind the source take the direct path (one global
// atomic per voxel-view); bricks holding Q.22 values above 2^32 take it with
// 64-bit values.  Every path adds the same integers: the sums are
// bit-identical whatever the path, the lane order or the launch order.
//
// Each wave splats columns of the brick's 4 layers at one (x, depth), so a
// view whose W does not change across the layers computes W and its
// reciprocal once per column (see the coherence modes below).
//
// Float sums (AccT = float, the opt-in XPR_ACC_F32): the same bricks, Q.22
// values, integer tiles and paths; only the detector sums are float.  A
// flushed tile entry, or a direct voxel-view, adds its real value (the Q.22
// integer times 2^-22, rounded once to float) with a global float atomic.
// The additions of a pixel happen in whatever order the atomics land, so the
// sums are within float rounding of the exact Q.22 sum, not bit-reproducible.
// (The tiles stay integer: ds_add_f32 issues about 40 times slower than
// ds_add_u32 on gfx950.)
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

// (float: the hardware global_atomic_add_f32; the sums are coarse-grained
// device memory, where it is supported)
template <typename AccT>
__device__ __forceinline__ void global_add(AccT *p, AccT v)
{
    __hip_atomic_fetch_add((Glob<AccT> *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A Q.22 integer (a tile entry or a voxel-view's value) in the sums' type:
// the integer itself, or for float sums its real value q x 2^-22 rounded once
// to float (a 64-bit q through double, exact below 2^53: a tile entry sums at
// most 4096 values below 2^32).
template <typename AccT, typename Q>
__device__ __forceinline__ AccT sum_term(Q q)
{
    if constexpr (std::is_same<AccT, float>::value) {
        if constexpr (sizeof(Q) == 8) return (float)((double)q * (double)kQ22Eps);
        else return (float)q * kQ22Eps;
    } else {
        return (AccT)q;
    }
}

enum : uint32_t {
    kBinSkip = 1u,     // no voxel of the brick lands inside the view's bounds
    kBinTile = 2u,     // splat into an LDS tile (else direct global atomics)
    kBinFast = 4u,     // every voxel's W is inside the fast reciprocal's range
    kBinDepth = 8u,    // depth-lane order
    kBinExact = 16u,   // the tile test alone decides the canonical bounds (see bin_view)
    kBinInside = 32u,  // exact, and every voxel-view of the brick lands in the tile
    kBinColW = 64u,    // the w row terms of every (depth, layer) agree across the 4 layers
    kBinColU = 128u,   // and so do the u row terms
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
    uint32_t magic;                // ceil(2^32 / tw) (0 when tw is 1): tile entry -> row
    uint32_t pad[2];
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
    const float r = recip(Wc);                          // 1.0f / Wc, exactly
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
    // Layer coherence of the brick's rows (see the coherence modes below),
    // with the positions and FMA chains the splat evaluates.  A row term
    // fmaf(c, py, t) is the correctly rounded value of a function affine in
    // py, and py = fmaf(layer, vx, origin) rounds monotonically in the layer:
    // the 4 layers' terms are monotone in the layer, so they all agree when
    // the first and last layer's do.  The 8 lanes of the view check 2 of the
    // 16 depths each.
    int wsame = 1, usame = 1;
    const float py0 = __builtin_fmaf((float)(lo[1] + a.vmin[1]), a.voxel[1], a.origin[1]);
    const float py3 = __builtin_fmaf((float)(lo[1] + 3 + a.vmin[1]), a.voxel[1], a.origin[1]);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int d = corner * 2 + i;
        const float pz = __builtin_fmaf((float)(lo[2] + d + a.vmin[2]), a.voxel[2], a.origin[2]);
        const float t0 = __builtin_fmaf(p[2], pz, p[3]), t3 = __builtin_fmaf(p[14], pz, p[15]);
        wsame &= __builtin_fmaf(p[13], py3, t3) == __builtin_fmaf(p[13], py0, t3);
        usame &= __builtin_fmaf(p[1], py3, t0) == __builtin_fmaf(p[1], py0, t0);
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        wsame &= __shfl_xor(wsame, off);
        usame &= __shfl_xor(usame, off);
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
    if (wsame) flags |= kBinColW;
    if (wsame & usame) flags |= kBinColU;
    b.flags = flags;
    b.magic = b.tw > 1 ? 0xffffffffu / (uint32_t)b.tw + 1u : 0u;
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

// one brick with some Q.22 value of 2^32 or more (64-bit or float sums): one
// global atomic per voxel-view with the exact 64-bit value, or for float sums
// its real value (x-lane order)
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
        AccT q;
        if constexpr (sizeof(AccT) == 8 || std::is_same<AccT, float>::value) q = sum_term<AccT>(q22_u64(v));
        else q = (AccT)q22_u32(v);
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

// LDS byte address of a pointer into shared memory
template <typename T>
__device__ __forceinline__ uint32_t lds_address(const T *p)
{
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) T *)p;
}

// How a view's per-voxel projection is evaluated for one brick (uniform over
// the workgroup; bin_view decides it from the brick's row terms).
// The canonical X, Y, W of a voxel are fmaf(c, px, k) with the row terms k of
// its (layer, depth).  When the w row terms of a depth are the same float in
// the brick's 4 layers (a zero layer coefficient in ProjData float 13, as in
// every detector the reference builds, or the rounding residue a roll about
// the beam axis leaves there), W and its reciprocal are computed once per
// (x, depth) column and serve the 4 layers (kColW); when the u row terms agree
// too (an upright detector), so does u (kColU).  Equal floats may differ in
// the sign of a zero: a zero X gives the same pixel and bounds result with
// either sign, and the views that share W are fast views, whose W is not 0.
enum : int { kColU = 0, kColW = 1, kColNone = 2 };
// what a voxel-view does with its pixel
enum : int {
    kSplatInside = 0,    // kBinInside: lands in the tile, no test
    kSplatExact = 1,     // kBinExact: the tile test is the bounds test
    kSplatChecked = 2,   // bounds test; off the tile (or no tile): global atomic
};

template <typename AccT, typename TileT>
struct SplatTarget {
    int32_t u0, v0, tw, th;      // tile origin and extent
    uint32_t twb, tb;            // inside: tile row in bytes, LDS address of pixel (0, 0)
    TileT *tl;                   // tile
    uint32_t sk;                 // this lane's sink entry, relative to tl
    bool use_tile;
    AccT *vsum;                  // the view's detector sums
    uint32_t det_width;
};

template <typename TileT>
__device__ __forceinline__ void lds_add(uint32_t addr, uint32_t q)
{
    using L = __attribute__((address_space(3))) TileT;
    __hip_atomic_fetch_add((L *)(uintptr_t)addr, (TileT)q, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// one voxel-view: its Q.22 value goes to its pixel's tile entry (inside: the
// address is one shift-add and one 24-bit multiply-add); voxel-views out of
// bounds add 0 to the lane's sink slot, so the splat needs no branch; a
// checked voxel-view in bounds but off the tile goes straight to the sums
template <int SPLAT, typename AccT, typename TileT>
__device__ __forceinline__ void deposit(float u, float v, uint32_t q, const ViewConsts &c,
                                        const SplatTarget<AccT, TileT> &t)
{
    constexpr uint32_t kShift = sizeof(TileT) == 8 ? 3u : 2u;
    if (SPLAT == kSplatInside) {
        const uint32_t col = ((uint32_t)pixel_of(u) << kShift) + t.tb;
        uint32_t addr;
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(addr) : "v"((uint32_t)pixel_of(v)), "s"(t.twb), "v"(col));
        lds_add<TileT>(addr, q);
    } else if (SPLAT == kSplatExact) {
        const uint32_t tu = (uint32_t)pixel_of(u) - (uint32_t)t.u0;
        const uint32_t tv = (uint32_t)pixel_of(v) - (uint32_t)t.v0;
        const bool hit = (tu < (uint32_t)t.tw) & (tv < (uint32_t)t.th);
        const uint32_t e = __umul24(tv, (uint32_t)t.tw) + tu;
        atomicAdd(&t.tl[hit ? e : t.sk], (TileT)(hit ? q : 0u));
    } else {
        const bool inb = (u >= c.ulo) & (u < c.uhi) & (v >= c.vlo) & (v < c.vhi);
        const uint32_t iu = (uint32_t)pixel_of(u), iv = (uint32_t)pixel_of(v);
        const uint32_t tu = iu - (uint32_t)t.u0;
        const uint32_t tv = iv - (uint32_t)t.v0;
        const bool hit = inb & t.use_tile & (tu < (uint32_t)t.tw) & (tv < (uint32_t)t.th);
        const uint32_t e = __umul24(tv, (uint32_t)t.tw) + tu;
        atomicAdd(&t.tl[hit ? e : t.sk], (TileT)(hit ? q : 0u));
        if (inb & !hit & (q != 0u))      // off the tile: direct global add
            global_add(t.vsum + (__umul24(iv, t.det_width) + iu), sum_term<AccT>(q));
    }
}

// The 4 layers of one (x, depth) column of the brick for one view: px, the
// Q.22 values q[l] and the row terms k0/k1/k3[l] of each layer.  FAST: every
// W of the view is in the fast reciprocal's range.
template <int MODE, int SPLAT, bool FAST, bool CROP, typename AccT, typename TileT>
__device__ __forceinline__ void splat_column(float px, const uint32_t (&q)[kWaves],
                                             const float (&k0)[kWaves], const float (&k1)[kWaves],
                                             const float (&k3)[kWaves], const ViewConsts &c,
                                             float line_off, const SplatTarget<AccT, TileT> &t)
{
    auto rcp = [](float w) {
        if (FAST) {
            const float r0 = __builtin_amdgcn_rcpf(w);
            return __builtin_fmaf(__builtin_fmaf(-w, r0, 1.0f), r0, r0);
        }
        return recip(w);
    };
    float r = 0.0f, u = 0.0f;
    if (MODE != kColNone) r = rcp(__builtin_fmaf(c.c3x, px, k3[0]));
    if (MODE == kColU) u = __builtin_fmaf(c.c0x, px, k0[0]) * r;
#pragma unroll
    for (int l = 0; l < kWaves; ++l) {
        if (MODE == kColNone) r = rcp(__builtin_fmaf(c.c3x, px, k3[l]));
        if (MODE != kColU) u = __builtin_fmaf(c.c0x, px, k0[l]) * r;
        const float y = __builtin_fmaf(c.c1x, px, k1[l]) * r;
        deposit<SPLAT, AccT, TileT>(u, CROP ? y - line_off : y, q[l], c, t);
    }
}

template <int V>
using IntTag = std::integral_constant<int, V>;

// Calls f(mode, splat, fast) as compile-time tags for the compiled
// combinations: inside and exact views (fast) in every mode; checked views
// with the per-layer projection, which is correct for every view.
template <typename F>
__device__ __forceinline__ void splat_dispatch(int splat, bool fast, int mode, F &&f)
{
    if (splat == kSplatChecked) {
        if (fast) f(IntTag<kColNone>{}, IntTag<kSplatChecked>{}, std::true_type{});
        else f(IntTag<kColNone>{}, IntTag<kSplatChecked>{}, std::false_type{});
    } else if (splat == kSplatInside) {
        if (mode == kColU) f(IntTag<kColU>{}, IntTag<kSplatInside>{}, std::true_type{});
        else if (mode == kColW) f(IntTag<kColW>{}, IntTag<kSplatInside>{}, std::true_type{});
        else f(IntTag<kColNone>{}, IntTag<kSplatInside>{}, std::true_type{});
    } else {
        if (mode == kColU) f(IntTag<kColU>{}, IntTag<kSplatExact>{}, std::true_type{});
        else if (mode == kColW) f(IntTag<kColW>{}, IntTag<kSplatExact>{}, std::true_type{});
        else f(IntTag<kColNone>{}, IntTag<kSplatExact>{}, std::true_type{});
    }
}

constexpr int kD = kBK / kWaves;     // depths per wave in x-lane order, x steps in depth-lane order
static_assert(kD == 4 && kWaves == 4, "a wave holds 4 x 4 columns of the brick's 4 layers");

// x-lane order: depth 4 wave + d of the brick; lane 4 d + l holds the row
// terms of (depth d, layer l)
template <int MODE, int SPLAT, bool FAST, bool CROP, typename AccT, typename TileT>
__device__ __forceinline__ void splat_x_order(float px, const uint32_t (&q)[kD][kWaves], float k0t,
                                              float k1t, float k3t, const ViewConsts &c,
                                              float line_off, const SplatTarget<AccT, TileT> &t)
{
#pragma unroll
    for (int d = 0; d < kD; ++d) {
        float K0[kWaves], K1[kWaves], K3[kWaves];
#pragma unroll
        for (int l = 0; l < kWaves; ++l) {
            K0[l] = lane_float(k0t, 4 * d + l);
            K1[l] = lane_float(k1t, 4 * d + l);
            K3[l] = lane_float(k3t, 4 * d + l);
        }
        splat_column<MODE, SPLAT, FAST, CROP, AccT, TileT>(px, q[d], K0, K1, K3, c, line_off, t);
    }
}

// depth-lane order: the lane's depth, x step xs
template <int MODE, int SPLAT, bool FAST, bool CROP, typename AccT, typename TileT>
__device__ __forceinline__ void splat_depth_order(const float (&px)[kD],
                                                  const uint32_t (&q)[kD][kWaves],
                                                  const float (&K0)[kWaves], const float (&K1)[kWaves],
                                                  const float (&K3)[kWaves], const ViewConsts &c,
                                                  float line_off, const SplatTarget<AccT, TileT> &t)
{
#pragma unroll
    for (int xs = 0; xs < kD; ++xs)
        splat_column<MODE, SPLAT, FAST, CROP, AccT, TileT>(px[xs], q[xs], K0, K1, K3, c, line_off, t);
}

// The views of one brick (see forward_brick), with TileT tile entries.
template <typename VS, typename AccT, bool CROP, typename TileT>
__device__ __forceinline__ void forward_views(AccT *__restrict__ sums, const float *__restrict__ pd,
                                           const xpr_proj_args a, const BrickShape s, int32_t x0,
                                           int32_t l0, int32_t k0, const uint32_t (&qx)[kD][kWaves],
                                           const uint32_t (&qd)[kD][kWaves],
                                           unsigned long long *__restrict__ lds,
                                           ViewBin *__restrict__ bins)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int32_t dk = lane & 15, xg = lane >> 4;     // depth-lane order
    TileT *__restrict__ tiles = (TileT *)lds;
    const float px_x = __builtin_fmaf((float)(x0 + lane + a.vmin[0]), a.voxel[0], a.origin[0]);
    float px_d[kD];
#pragma unroll
    for (int xs = 0; xs < kD; ++xs)
        px_d[xs] = __builtin_fmaf((float)(x0 + wave + 16 * xg + 4 * xs + a.vmin[0]), a.voxel[0],
                                  a.origin[0]);
    float py[kWaves];
#pragma unroll
    for (int l = 0; l < kWaves; ++l)
        py[l] = __builtin_fmaf((float)(l0 + l + a.vmin[1]), a.voxel[1], a.origin[1]);
    const float pz_d = __builtin_fmaf((float)(k0 + dk + a.vmin[2]), a.voxel[2], a.origin[2]);
    // x-lane order row terms: lane i computes (depth 4 wave + i % 16 / 4, layer i % 4)
    const int32_t ti = lane & 15;
    const float pz_t = __builtin_fmaf((float)(k0 + wave * kD + (ti >> 2) + a.vmin[2]), a.voxel[2],
                                      a.origin[2]);
    const float py_ti = __builtin_fmaf((float)(l0 + (ti & 3) + a.vmin[1]), a.voxel[1], a.origin[1]);

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
            SplatTarget<AccT, TileT> t;
            t.u0 = __builtin_amdgcn_readfirstlane(bins[vi].u0);
            t.v0 = __builtin_amdgcn_readfirstlane(bins[vi].v0);
            t.tw = __builtin_amdgcn_readfirstlane(bins[vi].tw);
            t.th = __builtin_amdgcn_readfirstlane(bins[vi].th);
            t.use_tile = (flags & kBinTile) != 0;
            // (a uniform branch spares the single-module case the integer division)
            const uint32_t view =
                __builtin_amdgcn_readfirstlane(a.modules == 1 ? m : m / a.modules);
            t.vsum = sums + (uint64_t)view_of(a, view) * view_elems;
            t.det_width = a.det_width;
            t.tl = tiles + buf * kTileCap;
            t.sk = (uint32_t)((2 - buf) * kTileCap) + threadIdx.x;
            t.twb = (uint32_t)sizeof(TileT) * (uint32_t)t.tw;
            t.tb = lds_address(t.tl) - (uint32_t)sizeof(TileT) *
                                           (__umul24((uint32_t)t.v0, (uint32_t)t.tw) + (uint32_t)t.u0);
            const ViewBin &bv = bins[vi];
            ViewConsts c;
            c.ulo = bv.ulo;
            c.uhi = bv.uhi;
            c.vlo = bv.vlo;
            c.vhi = bv.vhi;
            c.c0x = bv.m[0];
            c.c1x = bv.m[4];
            c.c3x = bv.m[8];
            c.k0 = c.k1 = c.k3 = 0.0f;
            const int splat = (flags & kBinInside) ? kSplatInside
                              : (flags & kBinExact) ? kSplatExact : kSplatChecked;
            const bool fast = (flags & kBinFast) != 0;
            const int mode = (flags & kBinColU) ? kColU : (flags & kBinColW) ? kColW : kColNone;
            if (flags & kBinDepth) {
                // depth-lane order: this lane's depth, the 4 layers' row terms
                const float t0 = __builtin_fmaf(bv.m[2], pz_d, bv.m[3]);
                const float t1 = __builtin_fmaf(bv.m[6], pz_d, bv.m[7]);
                const float t3 = __builtin_fmaf(bv.m[10], pz_d, bv.m[11]);
                float K0[kWaves], K1[kWaves], K3[kWaves];
#pragma unroll
                for (int l = 0; l < kWaves; ++l) {
                    K0[l] = __builtin_fmaf(bv.m[1], py[l], t0);
                    K1[l] = __builtin_fmaf(bv.m[5], py[l], t1);
                    K3[l] = __builtin_fmaf(bv.m[9], py[l], t3);
                }
                splat_dispatch(splat, fast, mode, [&](auto M, auto S, auto F) __attribute__((always_inline)) {
                    splat_depth_order<decltype(M)::value, decltype(S)::value, decltype(F)::value,
                                      CROP, AccT, TileT>(px_d, qd, K0, K1, K3, c, line_off, t);
                });
            } else {
                // x-lane order: lane 4 d + l computes the row terms of (depth
                // 4 wave + d, layer l); the splat reads them with v_readlane
                const float k0t = __builtin_fmaf(bv.m[1], py_ti, __builtin_fmaf(bv.m[2], pz_t, bv.m[3]));
                const float k1t = __builtin_fmaf(bv.m[5], py_ti, __builtin_fmaf(bv.m[6], pz_t, bv.m[7]));
                const float k3t = __builtin_fmaf(bv.m[9], py_ti, __builtin_fmaf(bv.m[10], pz_t, bv.m[11]));
                splat_dispatch(splat, fast, mode, [&](auto M, auto S, auto F) __attribute__((always_inline)) {
                    splat_x_order<decltype(M)::value, decltype(S)::value, decltype(F)::value, CROP,
                                  AccT, TileT>(px_x, qx, k0t, k1t, k3t, c, line_off, t);
                });
            }
            if (t.use_tile) {
                lds_barrier();                // splat of this view complete
                // flush and clear, the tile's entries spread over the
                // workgroup's threads; the next view splats into the other
                // tile.  Entry e is (row, col) = (e / tw, e % tw), the row by
                // a multiply-high with ceil(2^32 / tw) (exact for e * tw < 2^32).
                const uint32_t n = (uint32_t)(t.tw * t.th);
                const uint32_t magic = __builtin_amdgcn_readfirstlane(bv.magic);
                AccT *base = t.vsum + (uint32_t)t.v0 * t.det_width + (uint32_t)t.u0;
                for (uint32_t e = threadIdx.x; e < n; e += (uint32_t)kThreads) {
                    const TileT v = t.tl[e];
                    if (v != 0) {
                        const uint32_t row = magic ? __umulhi(e, magic) : e;
                        const uint32_t col = e - row * (uint32_t)t.tw;
                        global_add(base + (row * t.det_width + col), sum_term<AccT>(v));
                        t.tl[e] = 0;
                    }
                }
                buf ^= 1;
            }
        }
    }
}

// One brick of the forward (b < s.nbricks, uniform over the workgroup).  lds
// holds two tiles and one sink slot per lane ([2*kTileCap + kThreads]); the
// tiles are zero on entry and left zero.
//
// Lane order per view.  Voxels along x project to neighbouring detector
// columns when the ray is across x, but onto one or two columns when the ray
// runs along x (views near 90 and 270 degrees): 64 lanes of one x-row would
// then hit the same LDS entries in one atomic.  Each wave therefore holds 1024
// voxels of the brick twice, each time as 256 columns of the 4 layers: in
// x-lane order wave w takes depths 4w..4w+3 with x along the lanes; in
// depth-lane order it takes x = w + 4m (m = 16 xg + 4 xs, xs = 0..3) with the
// 16 depths along lanes and 4 x-groups xg.  A view uses the order whose lanes
// spread more across columns.
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
    const int32_t dk = lane & 15, xg = lane >> 4;     // depth-lane order
    // 64-bit and float sums take Q.22 values of 2^32 and more (the direct
    // path) and tile sums of 2^32 and more (64-bit tiles); 32-bit sums wrap
    // like the reference's
    constexpr bool kWideSums = !std::is_same<AccT, uint32_t>::value;

    // Q.22 voxel values in both lane orders (32 bits suffice unless some
    // value is >= 1024), qx[d][l] and qd[xs][l].  Raw buffer loads from each
    // layer's brick corner (x0, l0 + l, k0): out-of-slab voxels read 0 without
    // a branch, so all 32 loads are in flight together.
    uint32_t qx[kD][kWaves], qd[kD][kWaves];
    bool big = false;
    {
        constexpr uint32_t esz = (uint32_t)sizeof(typename VS::T);
        auto q22 = [&](float v, bool track) {
            const float sc = v * kQ22One;               // FLOAT_TO_FIXED (constants.cl:22)
            const bool pos = sc > 0.0f, over = sc >= 4294967296.0f;
            if (kWideSums && track) big |= pos & over;
            return pos ? (over ? 0xffffffffu : (uint32_t)sc) : 0u;
        };
        float vx_[kD][kWaves], vd_[kD][kWaves];
#pragma unroll
        for (int l = 0; l < kWaves; ++l) {
            const int32_t ll = l0 + l;
            const int64_t base = ((int64_t)ll * s.nk + k0) * s.nx + x0;
            const int64_t left = (int64_t)s.nl * s.nk * s.nx - base;  // voxels from the corner on
            const uint32_t bytes = ll < s.nl ? (uint32_t)min<int64_t>(left * esz, 0x7fffffff) : 0u;
            const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(vol + (ll < s.nl ? base : 0)), 0, (int)bytes, kBufferWord3);
#pragma unroll
            for (int d = 0; d < kD; ++d) {
                const int32_t kk = wave * kD + d;
                const bool in = (x0 + lane < s.nx) & (k0 + kk < s.nk);
                const uint32_t off = (uint32_t)(kk * s.nx + lane) * esz;
                vx_[d][l] = load_or_zero<VS>(rsrc, in ? off : kOutOfRange);
            }
#pragma unroll
            for (int xs = 0; xs < kD; ++xs) {
                const int32_t xx = wave + 16 * xg + 4 * xs;
                const bool in = (x0 + xx < s.nx) & (k0 + dk < s.nk);
                const uint32_t off = (uint32_t)(dk * s.nx + xx) * esz;
                vd_[xs][l] = load_or_zero<VS>(rsrc, in ? off : kOutOfRange);
            }
        }
#pragma unroll
        for (int d = 0; d < kD; ++d)
#pragma unroll
            for (int l = 0; l < kWaves; ++l) {
                qx[d][l] = q22(vx_[d][l], true);
                qd[d][l] = q22(vd_[d][l], false);
            }
    }
    if (kWideSums && __syncthreads_or(big)) {
        forward_direct_u64<VS, AccT, CROP>(vol, sums, pd, a, s, x0, l0, k0);
        return;
    }
    // A tile entry sums at most the brick's 4096 voxel-views of one view, so
    // 32-bit entries hold every sum exactly when all Q.22 values are below
    // 2^20, and always when the detector sums themselves are 32-bit (the same
    // sums modulo 2^32): half the LDS atomic traffic.
    bool wide = false;
    if (kWideSums) {
#pragma unroll
        for (int d = 0; d < kD; ++d)
#pragma unroll
            for (int l = 0; l < kWaves; ++l) wide |= qx[d][l] >= (1u << 20);
    }
    if (kWideSums && __syncthreads_or(wide))
        forward_views<VS, AccT, CROP, unsigned long long>(sums, pd, a, s, x0, l0, k0, qx, qd, lds, bins);
    else
        forward_views<VS, AccT, CROP, uint32_t>(sums, pd, a, s, x0, l0, k0, qx, qd, lds, bins);
}

// Workgroups per CU the forward is built for: 4 (128 VGPRs) in general; 5
// (96 VGPRs) with 32-bit sums, whose kernel holds no 64-bit tile path
// (0.504 against 0.544 ms per 64-layer slab of the 768^3 workload).
template <typename AccT>
constexpr int forward_wg_per_cu() { return std::is_same<AccT, uint32_t>::value ? 5 : 4; }

template <typename VS, typename AccT, bool CROP>
__global__ __launch_bounds__(kThreads, forward_wg_per_cu<AccT>()) void k_forward_bin(
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
