// HIP kernels (gfx950) and the C-ABI of the iterative cone-beam hot path.
//
// Canonical arithmetic shared with the CPU oracle (the reference compiles its
// OpenCL kernels with -cl-fast-relaxed-math -cl-mad-enable, gpu/kernel.hpp:26-31,
// so its own output is platform dependent; this build fixes one legal lowering):
//   position  p_a = fmaf((float)idx_a, voxel_a, origin_a)       KernelForward.cl:26
//   dot(c, p) = fmaf(c.x, p.x, fmaf(c.y, p.y, fmaf(c.z, p.z, c.w)))   utils.cl:26
//   (u, v)    = (X * r, Y * r - line_offset), r = 1.0f / W (IEEE)     utils.cl:26
//   pixel     = roundf (half away from zero), half-open viewport     KernelForward.cl:34-35
// Built with -ffp-contract=off, so only the explicit fmaf calls fuse.
//
// Layouts in HBM:
//   volume slab : [layer][depth][x] (to3D, utils.cl:39-45), x fastest
//   sums / err  : [view][row][col], row stride det_width, view stride
//                 det_width*line_rows (IMAGE_OFFSET_2D, utils.cl:47)
//   ProjData    : 20 floats per (view, module): glm column-major 4x4, then
//                 viewport (vpl, vpr, vpb, vpt) (constants.cl:25-28)

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>

#include "../../include/xpr_hip.h"
#include "batch.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define XPR_HIP_CHECK(expr)                                                          \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(XPR_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define XPR_LAUNCH_CHECK()                                                           \
    do {                                                                             \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess)                                                        \
            return fail(XPR_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr float kQ22One = 4194304.0f;        // FIXED_FRAC_ONE (constants.cl:19)
constexpr float kQ22Eps = 1.0f / 4194304.0f; // EPSILON (constants.cl:23)

// ---------------------------------------------------------------------------
// storage types
// ---------------------------------------------------------------------------
// Volume, image and error buffers are global memory; accesses go through
// global-address-space pointers so they stay global instructions (not generic
// ones, which the LDS wait counter also covers) whatever the pointer's origin.
template <typename T> using Glob = __attribute__((address_space(1))) T;

struct F32 {
    using T = float;
    static constexpr int code = XPR_F32;
    static __device__ __forceinline__ float load(const T *p, uint64_t i) { return ((const Glob<T> *)p)[i]; }
    static __device__ __forceinline__ void store(T *p, uint64_t i, float v) { ((Glob<T> *)p)[i] = v; }
};
struct F16 {
    using T = uint16_t;
    static constexpr int code = XPR_F16;
    static __device__ __forceinline__ float load(const T *p, uint64_t i)
    {
        return (float)__builtin_bit_cast(_Float16, ((const Glob<T> *)p)[i]);
    }
    static __device__ __forceinline__ void store(T *p, uint64_t i, float v)
    {
        ((Glob<T> *)p)[i] = __builtin_bit_cast(uint16_t, (_Float16)v);
    }
};
struct BF16 {
    using T = uint16_t;
    static constexpr int code = XPR_BF16;
    static __device__ __forceinline__ float load(const T *p, uint64_t i)
    {
        return __builtin_bit_cast(float, (uint32_t)((const Glob<T> *)p)[i] << 16);
    }
    static __device__ __forceinline__ void store(T *p, uint64_t i, float v)
    {
        uint32_t b = __builtin_bit_cast(uint32_t, v);
        if ((b & 0x7fffffffu) > 0x7f800000u) {
            ((Glob<T> *)p)[i] = (uint16_t)((b >> 16) | 0x40u);                 // quiet NaN
        } else {
            b += 0x7fffu + ((b >> 16) & 1u);                                   // nearest even
            ((Glob<T> *)p)[i] = (uint16_t)(b >> 16);
        }
    }
};

// ---------------------------------------------------------------------------
// Exact fast forms of the canonical projection.  Each helper returns exactly
// what the canonical expression returns; the two hardware identities they
// rely on were checked over every float32 input on gfx950:
//   * v_rcp_f32 followed by one FMA Newton step is the IEEE quotient 1/w for
//     every normal |w| in [1e-6, 1e6] (w outside that range takes 1.0f/w);
//   * v_cvt_rpi_i32_f32 (floor(u + 0.5)) equals (int)roundf(u) for every u in
//     (-0.5, 2^23), which holds for every in-bounds coordinate.
// In-bounds tests compare the unrounded coordinate with half-integer
// thresholds: for an integer bound L >= 1, roundf(u) >= L <=> u >= L - 0.5,
// roundf(u) >= 0 <=> u > -0.5 (ties round away from zero), and for R >= 1
// roundf(u) < R <=> u < R - 0.5.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float recip(float w)
{
    const float aw = fabsf(w);
    if (__builtin_expect(aw >= 1e-6f && aw <= 1e6f, 1)) {
        const float r0 = __builtin_amdgcn_rcpf(w);
        const float e = __builtin_fmaf(-w, r0, 1.0f);
        return __builtin_fmaf(e, r0, r0);
    }
    return 1.0f / w;
}

__device__ __forceinline__ int32_t pixel_of(float u)
{
    int32_t r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(u));
    return r;
}

// thresholds on the unrounded coordinate for roundf(u) in [lo, hi) with real
// bounds lo >= 0 (an integer n satisfies n >= lo <=> n >= ceil(lo), n < hi <=>
// n < ceil(hi))
__device__ __forceinline__ float lo_threshold(float lo)
{
    const float L = ceilf(lo);
    return L >= 1.0f ? L - 0.5f : -0.49999997f;       // the float just above -0.5
}
__device__ __forceinline__ float hi_threshold(float hi) { return ceilf(hi) - 0.5f; }

// per (row, view) constants
struct ViewConsts {
    float c0x, c1x, c3x;       // x coefficients of glm columns 0, 1, 3
    float k0, k1, k3;          // the (y, z, w) part of the three dot products for (py, pz)
    float ulo, uhi, vlo, vhi;  // thresholds of the half-open column and row bounds
};

// ProjData is read-only for every kernel: reading it through the constant
// address space lets uniform records go through scalar loads even where the
// compiler cannot prove that the kernel's own stores leave it unchanged.
using ConstF = const __attribute__((address_space(4))) float;
__device__ __forceinline__ ConstF *pd_const(const float *p) { return (ConstF *)p; }

// view bounds, clamped to the buffers the call holds
__device__ __forceinline__ void view_thresholds(const float *__restrict__ pdg,
                                                const xpr_proj_args &a, ViewConsts &c)
{
    ConstF *pd = pd_const(pdg);
    c.ulo = lo_threshold(fmaxf(pd[16], 0.0f));
    c.uhi = hi_threshold(fminf(pd[17], (float)a.det_width));
    c.vlo = lo_threshold(fmaxf(pd[18], 0.0f));
    c.vhi = hi_threshold(fminf(pd[19], (float)a.line_rows));
}

__device__ __forceinline__ void row_terms(const float *__restrict__ pdg, float py, float pz,
                                          ViewConsts &c)
{
    ConstF *pd = pd_const(pdg);
    c.c0x = pd[0];
    c.c1x = pd[4];
    c.c3x = pd[12];
    c.k0 = __builtin_fmaf(pd[1], py, __builtin_fmaf(pd[2], pz, pd[3]));
    c.k1 = __builtin_fmaf(pd[5], py, __builtin_fmaf(pd[6], pz, pd[7]));
    c.k3 = __builtin_fmaf(pd[13], py, __builtin_fmaf(pd[14], pz, pd[15]));
}

// The exact W of a voxel is affine in its position, so over a box of voxels
// it lies between its values at the box corners; the three FMAs that
// evaluate it err by at most T * 2^-22, T = |c3x px| + |c3y py| + |c3z pz| + |c3w|.
// Corner values in [1e-2, 1e5] with T <= 1e4 therefore keep every W of the
// box inside [1e-6, 1e6], where the fast reciprocal is exact.
__device__ __forceinline__ bool fast_rcp_box(float w, float t)
{
    return (w >= 1e-2f) & (w <= 1e5f) & (t <= 1e4f);
}

// Projects the voxel at x-position px: true when in bounds, with its pixel
// column iu and row iv (meaningful when in bounds).  FAST: the caller has
// checked (row_setup) that W is in the fast reciprocal's range.
template <bool CROP, bool FAST>
__device__ __forceinline__ bool project(const ViewConsts &c, float px, float line_off,
                                        uint32_t &iu, uint32_t &iv)
{
    const float X = __builtin_fmaf(c.c0x, px, c.k0);
    const float Y = __builtin_fmaf(c.c1x, px, c.k1);
    const float W = __builtin_fmaf(c.c3x, px, c.k3);
    float r;
    if (FAST) {
        const float r0 = __builtin_amdgcn_rcpf(W);
        r = __builtin_fmaf(__builtin_fmaf(-W, r0, 1.0f), r0, r0);
    } else {
        r = recip(W);
    }
    const float u = X * r;
    const float v = CROP ? Y * r - line_off : Y * r;   // subtracting 0 is exact
    const bool inb = (u >= c.ulo) & (u < c.uhi) & (v >= c.vlo) & (v < c.vhi);
    iu = (uint32_t)pixel_of(u);
    iv = (uint32_t)pixel_of(v);
    return inb;
}

// detector view of launch view `view` (uniform: every caller passes a uniform view)
__device__ __forceinline__ uint32_t view_of(const xpr_proj_args &a, uint32_t view)
{
    return a.view_index ? __builtin_amdgcn_readfirstlane(a.view_index[view]) : view;
}

// raw buffer over one detector view: out-of-range offsets read as 0
constexpr int kBufferWord3 = 0x00020000;     // gfx9 raw buffer, no format conversion
constexpr uint32_t kOutOfRange = 0x80000000u;

template <typename T>
__device__ __forceinline__ float load_or_zero(__amdgpu_buffer_rsrc_t r, uint32_t byte_off);
template <>
__device__ __forceinline__ float load_or_zero<F32>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ float load_or_zero<F16>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    const uint16_t h = __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
    return (float)__builtin_bit_cast(_Float16, h);
}
template <>
__device__ __forceinline__ float load_or_zero<BF16>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    const uint16_t h = __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Indexed (structured) buffer loads: the buffer's stride is the sample size
// and num_records its sample count, so the hardware scales the index, and an
// index at or beyond num_records reads 0 (kNoSample is beyond any view).  The
// compiler offers raw buffer builtins only, so the intrinsic is bound by name.
constexpr uint32_t kNoSample = 0xffffffffu;
__device__ float struct_buffer_load_f32(__amdgpu_buffer_rsrc_t rsrc, uint32_t index, uint32_t voffset,
                                        uint32_t soffset, uint32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.f32");
__device__ uint16_t struct_buffer_load_u16(__amdgpu_buffer_rsrc_t rsrc, uint32_t index, uint32_t voffset,
                                           uint32_t soffset, uint32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.i16");

template <typename T>
__device__ __forceinline__ float sample_or_zero(__amdgpu_buffer_rsrc_t r, uint32_t index);
template <>
__device__ __forceinline__ float sample_or_zero<F32>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return struct_buffer_load_f32(r, index, 0, 0, 0);
}
template <>
__device__ __forceinline__ float sample_or_zero<F16>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return (float)__builtin_bit_cast(_Float16, struct_buffer_load_u16(r, index, 0, 0, 0));
}
template <>
__device__ __forceinline__ float sample_or_zero<BF16>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return __builtin_bit_cast(float, (uint32_t)struct_buffer_load_u16(r, index, 0, 0, 0) << 16);
}

// ---------------------------------------------------------------------------
// Backward: one wave per (layer, depth) row, lane owns J voxels at
// x = seg*64*J + lane + 64*j.  Per view the (y, z, w) part of the three dot
// products is computed once per row; a voxel-view then costs 3 FMA, the
// reciprocal, 2 MUL, 4 compares, 2 conversions, the index and one gather.
// The gather reads through an indexed buffer over the view, so out-of-bounds
// voxels read 0 without a branch and all J loads stay in flight.
// ---------------------------------------------------------------------------
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kBwdJ = 4;          // voxels per lane of the backward (default shape)

struct SlabShape {
    int32_t nx, nl, nk;
    int32_t nseg;          // x segments per row
    int64_t row_base;      // first row of this launch
    int64_t nrows;         // rows in the slab
    int32_t inside = 1;    // 1: backward views wholly in bounds skip the per-voxel test
};

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

// FLOAT_TO_FIXED (constants.cl:22), saturating like the GPU conversion.
__device__ __forceinline__ uint32_t q22_u32(float v)
{
    const float s = v * kQ22One;
    if (!(s > 0.0f)) return 0u;
    if (s >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)s;
}
__device__ __forceinline__ uint64_t q22_u64(float v)
{
    const float s = v * kQ22One;
    if (!(s > 0.0f)) return 0ull;
    if (s >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)s;
}

// ---------------------------------------------------------------------------
// Forward with LDS binning.  A workgroup owns a brick of 64 (x) x 4 (layers,
// one per wave) x 16 (depths).  Per view the brick's 8 corners bound its
// detector footprint (one pixel of margin covers float rounding); voxel-views
// accumulate there with LDS u64 atomics and the non-zero footprint pixels go
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
// The footprint, bounds thresholds and order of a chunk of 32 views are
// computed once per workgroup (8 corner lanes per view) into LDS.
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
    s.nrows = (int64_t)s.nl * s.nk;
    return s;
}

__global__ __launch_bounds__(kThreads) void k_batch_reset(uint32_t *counters)
{
    if (threadIdx.x < (unsigned)xpr_batch::kPhases * kXcds) counters[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(kThreads, 4) void k_batch_forward(const Table *__restrict__ t,
                                                               uint32_t *counters)
{
    __shared__ unsigned long long lds[2 * kTileCap + kThreads];
    __shared__ ViewBin bins[kViewChunk];
    __shared__ uint32_t slot;
    const uint32_t ph = xpr_batch::kForward;
    const uint32_t total = uniform_u32(&t->total[ph]);
    if (total == 0) return;
    for (int i = threadIdx.x; i < 2 * kTileCap; i += kThreads) lds[i] = 0ull;
    for (;;) {
        const uint32_t u = next_unit(counters, ph, &slot);
        if (u >= total) break;                              // uniform: every workgroup ends here
        const Item it = uniform_item(&t->items[ph][item_of(t, ph, u)]);
        const BrickShape s = batch_brick_shape(it.a);
        const int64_t b = (int64_t)(u - it.first);
        uint64_t *sums = (uint64_t *)it.sums;
        const void *vol = it.dst;
        switch (it.dt_dst) {
        case XPR_F32:
            forward_brick<F32, uint64_t, false>((const float *)vol, sums, it.pd, it.a, s, b, lds, bins);
            break;
        case XPR_F16:
            forward_brick<F16, uint64_t, false>((const uint16_t *)vol, sums, it.pd, it.a, s, b, lds, bins);
            break;
        default:
            forward_brick<BF16, uint64_t, false>((const uint16_t *)vol, sums, it.pd, it.a, s, b, lds, bins);
            break;
        }
    }
}

// Backward units are the single-job launch's workgroups: segment unit % nseg
// of the kWaves adjacent rows of row block unit / nseg, one row per wave, run
// together so the four rows (which gather nearly the same detector pixels)
// share the CU's L1.
__global__ __launch_bounds__(kThreads, 7) void k_batch_backward(const Table *__restrict__ t,
                                                                uint32_t *counters)
{
    __shared__ uint32_t slot;
    __shared__ float4 rec[kWaves * 128];              // per-wave view-record tables
    float4 *wrec = rec + (threadIdx.x >> 6) * 128;
    const uint32_t ph = xpr_batch::kBackward;
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
        switch (it.dt_dst) {                                // volume and error share the dtype
        case XPR_F32:
            backward_row<F32, F32, kBwdJ, false>((float *)it.dst, (const float *)it.err, it.pd, it.a, s,
                                                 row, seg, nullptr, wrec);
            break;
        case XPR_F16:
            backward_row<F16, F16, kBwdJ, false>((uint16_t *)it.dst, (const uint16_t *)it.err, it.pd,
                                                 it.a, s, row, seg, nullptr, wrec);
            break;
        default:
            backward_row<BF16, BF16, kBwdJ, false>((uint16_t *)it.dst, (const uint16_t *)it.err, it.pd,
                                                   it.a, s, row, seg, nullptr, wrec);
            break;
        }
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

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
inline unsigned grid_for(uint64_t n, unsigned block = 256)
{
    uint64_t g = (n + block - 1) / block;
    const uint64_t cap = 256ull * 256ull;     // grid-stride beyond this
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

int check_dtype(int dt)
{
    if (dt != XPR_F32 && dt != XPR_F16 && dt != XPR_BF16)
        return fail(XPR_E_UNSUPPORTED, "unknown dtype " + std::to_string(dt));
    return XPR_OK;
}

int check_args(const xpr_proj_args *a, SlabShape &s)
{
    if (a == nullptr) return fail(XPR_E_INVALID, "null xpr_proj_args");
    const int64_t nx = (int64_t)a->vmax[0] - a->vmin[0];
    const int64_t nl = (int64_t)a->vmax[1] - a->vmin[1];
    const int64_t nk = (int64_t)a->vmax[2] - a->vmin[2];
    if (nx < 0 || nl < 0 || nk < 0) return fail(XPR_E_INVALID, "volumeMax < volumeMin");
    if (a->modules == 0) return fail(XPR_E_INVALID, "modules must be >= 1");
    if (a->line_rows < 0) return fail(XPR_E_INVALID, "line_rows < 0");
    if (nx > (1 << 24) || nk > (1 << 24) || nl > (1 << 24))
        return fail(XPR_E_INVALID, "slab edge above 2^24 voxels");
    if ((uint64_t)a->det_width * (uint64_t)a->line_rows > 0xffffffffull)
        return fail(XPR_E_INVALID, "detector view above 2^32 pixels");
    s.nx = (int32_t)nx;
    s.nl = (int32_t)nl;
    s.nk = (int32_t)nk;
    s.nrows = nl * nk;
    s.row_base = 0;
    return XPR_OK;
}

template <typename Launch>
int launch_rows(SlabShape s, int J, Launch &&launch)
{
    if (s.nx == 0 || s.nrows == 0) return XPR_OK;
    s.nseg = (s.nx + 64 * J - 1) / (64 * J);
    // keep grid * block below 2^31 per launch
    const int64_t max_blocks = (int64_t(1) << 31) / kThreads;
    const int64_t rows_per_launch = (max_blocks / s.nseg) * kWaves;
    for (int64_t r0 = 0; r0 < s.nrows; r0 += rows_per_launch) {
        SlabShape t = s;
        t.row_base = r0;
        const int64_t rows = std::min<int64_t>(rows_per_launch, s.nrows - r0);
        const int64_t blocks = ((rows + kWaves - 1) / kWaves) * s.nseg;
        launch((unsigned)blocks, t);
        XPR_LAUNCH_CHECK();
    }
    return XPR_OK;
}

// ---------------------------------------------------------------------------
// Tile-shape table: the backward's voxels per lane (J) and the forward's tile
// capacity per launch shape, set by xpr_tune_* from timings on the caller's
// buffers.  A shape with no entry uses the defaults below.  The choice moves
// work between lanes and paths only: results are the same integers and the
// same per-voxel arithmetic whatever is chosen.
// ---------------------------------------------------------------------------
enum TuneKind { kTuneForward = 0, kTuneBackward = 1 };
using TuneKey = std::tuple<int, int, int, int, int32_t, int32_t, uint32_t, uint32_t, uint32_t,
                           int32_t, int>;    // kind, device, dtypes (2), nx, nk, views, modules,
                                             // det_width, line_rows, crop
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;

TuneKey tune_key(int kind, int d0, int d1, const xpr_proj_args &a, const SlabShape &s)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    return TuneKey(kind, dev, d0, d1, s.nx, s.nk, a.n_views, a.modules, a.det_width, a.line_rows,
                   a.line_offset != 0 ? 1 : 0);
}

int tune_lookup(const TuneKey &k)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(k);
    return it == g_tune.end() ? -1 : it->second;
}

void tune_store(const TuneKey &k, int v)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[k] = v;
}

// median-free timing: the best of `reps` launches after one warm-up launch
template <typename F>
int time_best(hipStream_t st, int reps, F &&launch, float *best)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XPR_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(XPR_E_HIP, "event");
    }
    int rc = launch();
    *best = 3.4e38f;
    for (int r = 0; rc == XPR_OK && r < reps; ++r) {
        (void)hipEventRecord(e0, st);
        rc = launch();
        (void)hipEventRecord(e1, st);
        float ms = 0.0f;
        if (rc == XPR_OK && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
            *best = std::min(*best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

constexpr int kBwdJs[] = {2, 4, 8, 16};

// default voxels per lane (J = 4 measured best at 768^3 / 36 views among 2, 4,
// 8, 16); narrow slabs take J = 2 so one 64*J segment leaves few lanes idle
int backward_default_j(int32_t nx) { return nx <= 128 ? 2 : kBwdJ; }

template <typename VS, typename ES, int J>
int backward_launch(hipStream_t st, void *vol, const void *err, const float *pd,
                    const xpr_proj_args &a, SlabShape s, uint64_t *empty)
{
    const bool crop = a.line_offset != 0;
    return launch_rows(s, J, [&](unsigned blocks, const SlabShape &t) {
        if (crop)
            hipLaunchKernelGGL((k_backward<VS, ES, J, true>), dim3(blocks), dim3(kThreads), 0,
                               st, (typename VS::T *)vol, (const typename ES::T *)err, pd, a, t,
                               (unsigned long long *)empty);
        else
            hipLaunchKernelGGL((k_backward<VS, ES, J, false>), dim3(blocks), dim3(kThreads), 0,
                               st, (typename VS::T *)vol, (const typename ES::T *)err, pd, a, t,
                               (unsigned long long *)empty);
    });
}

// J < 0: the table's choice for this shape, else the default
template <typename VS, typename ES>
int backward_impl(hipStream_t st, void *vol, const void *err, const float *pd,
                  const xpr_proj_args &a, SlabShape s, uint64_t *empty, int J = -1)
{
    if ((uint64_t)a.det_width * (uint64_t)a.line_rows * sizeof(typename ES::T) >= 0x80000000ull)
        return fail(XPR_E_INVALID, "a detector view above 2 GiB");
    // tuning knob: XPR_BACKWARD_INSIDE=0 keeps the canonical bounds test in every view
    if (const char *e = std::getenv("XPR_BACKWARD_INSIDE")) s.inside = std::atoi(e) != 0;
    if constexpr (std::is_same<VS, ES>::value) {       // the schedules' dtype pairs: all J compiled
        if (J < 0) {
            // tuning knob: force the voxels per lane
            if (const char *e = std::getenv("XPR_BACKWARD_J")) J = std::atoi(e);
            else J = tune_lookup(tune_key(kTuneBackward, VS::code, ES::code, a, s));
        }
        if (J < 0) J = backward_default_j(s.nx);
        switch (J) {
        case 2: return backward_launch<VS, ES, 2>(st, vol, err, pd, a, s, empty);
        case 4: return backward_launch<VS, ES, 4>(st, vol, err, pd, a, s, empty);
        case 8: return backward_launch<VS, ES, 8>(st, vol, err, pd, a, s, empty);
        case 16: return backward_launch<VS, ES, 16>(st, vol, err, pd, a, s, empty);
        default: return backward_launch<VS, ES, kBwdJ>(st, vol, err, pd, a, s, empty);
        }
    } else {
        (void)J;
        return backward_launch<VS, ES, kBwdJ>(st, vol, err, pd, a, s, empty);
    }
}

int forward_tile_cap()
{
    // tuning knob: tile entries in use; 0 sends every voxel-view to the direct path
    const char *e = std::getenv("XPR_FORWARD_TILE_CAP");
    const int c = e ? std::atoi(e) : kTileCap;
    return c < 0 ? 0 : (c > kTileCap ? kTileCap : c);
}

constexpr int kTileCaps[] = {384, 768, 1536};

// cap < 0: the table's choice for this shape, else the default (XPR_FORWARD_TILE_CAP or kTileCap)
template <typename VS, typename AccT>
int forward_impl(hipStream_t st, const void *vol, void *sums, const float *pd,
                 const xpr_proj_args &a, SlabShape s, int cap = -1)
{
    BrickShape b;
    b.nx = s.nx;
    b.nl = s.nl;
    b.nk = s.nk;
    b.nbx = (s.nx + 63) / 64;
    b.nbk = (s.nk + kBK - 1) / kBK;
    const int64_t nbl = (s.nl + kWaves - 1) / kWaves;
    b.nbricks = (int64_t)b.nbx * b.nbk * nbl;
    if (cap < 0 && !std::getenv("XPR_FORWARD_TILE_CAP"))
        cap = tune_lookup(tune_key(kTuneForward, VS::code, (int)sizeof(AccT), a, s));
    b.cap = cap < 0 ? forward_tile_cap() : std::min(cap, kTileCap);
    // tuning knob: XPR_FORWARD_EXACT=0 keeps the canonical bounds test in every view
    const char *ex = std::getenv("XPR_FORWARD_EXACT");
    b.exact = (ex && std::atoi(ex) == 0) ? 0 : 1;
    const bool crop = a.line_offset != 0;
    const int64_t max_blocks = (int64_t(1) << 31) / kThreads;
    for (int64_t b0 = 0; b0 < b.nbricks; b0 += max_blocks) {
        BrickShape t = b;
        t.brick_base = b0;
        const int64_t blocks = std::min<int64_t>(max_blocks, b.nbricks - b0);
        if (crop)
            hipLaunchKernelGGL((k_forward_bin<VS, AccT, true>), dim3((unsigned)blocks),
                               dim3(kThreads), 0, st, (const typename VS::T *)vol, (AccT *)sums,
                               pd, a, t);
        else
            hipLaunchKernelGGL((k_forward_bin<VS, AccT, false>), dim3((unsigned)blocks),
                               dim3(kThreads), 0, st, (const typename VS::T *)vol, (AccT *)sums,
                               pd, a, t);
        XPR_LAUNCH_CHECK();
    }
    return XPR_OK;
}

template <typename IS, typename AccT>
int error_impl2(hipStream_t st, const void *img, const void *sums, void *err, int err_dtype,
                uint64_t n, float w)
{
    const unsigned g = grid_for(n);
    switch (err_dtype) {
    case XPR_F32:
        hipLaunchKernelGGL((k_error<IS, AccT, F32>), dim3(g), dim3(256), 0, st,
                           (const typename IS::T *)img, (const AccT *)sums, (float *)err, n, w);
        break;
    case XPR_F16:
        hipLaunchKernelGGL((k_error<IS, AccT, F16>), dim3(g), dim3(256), 0, st,
                           (const typename IS::T *)img, (const AccT *)sums, (uint16_t *)err, n, w);
        break;
    default:
        hipLaunchKernelGGL((k_error<IS, AccT, BF16>), dim3(g), dim3(256), 0, st,
                           (const typename IS::T *)img, (const AccT *)sums, (uint16_t *)err, n, w);
        break;
    }
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

template <typename IS>
int error_impl(hipStream_t st, const void *img, const void *sums, int acc_bits, void *err,
               int err_dtype, uint64_t n, float w)
{
    if (acc_bits == XPR_ACC_U32)
        return error_impl2<IS, uint32_t>(st, img, sums, err, err_dtype, n, w);
    return error_impl2<IS, uint64_t>(st, img, sums, err, err_dtype, n, w);
}

template <typename DS>
int convert_from(hipStream_t st, void *dst, const void *src, int src_dtype, uint64_t n)
{
    const unsigned g = grid_for(n);
    switch (src_dtype) {
    case XPR_F32:
        hipLaunchKernelGGL((k_convert<DS, F32>), dim3(g), dim3(256), 0, st,
                           (typename DS::T *)dst, (const float *)src, n);
        break;
    case XPR_F16:
        hipLaunchKernelGGL((k_convert<DS, F16>), dim3(g), dim3(256), 0, st,
                           (typename DS::T *)dst, (const uint16_t *)src, n);
        break;
    default:
        hipLaunchKernelGGL((k_convert<DS, BF16>), dim3(g), dim3(256), 0, st,
                           (typename DS::T *)dst, (const uint16_t *)src, n);
        break;
    }
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

} // namespace

int xprint_fail(int code, const char *msg) { return fail(code, msg); }

// ===========================================================================
// Batched step engine (library-internal, used by the job server)
// ===========================================================================
namespace xpr_batch {

uint32_t forward_units(const xpr_proj_args &a)
{
    const int64_t nx = a.vmax[0] - a.vmin[0], nl = a.vmax[1] - a.vmin[1], nk = a.vmax[2] - a.vmin[2];
    return (uint32_t)(((nx + 63) / 64) * ((nk + kBK - 1) / kBK) * ((nl + kWaves - 1) / kWaves));
}

uint32_t backward_units(const xpr_proj_args &a)
{
    const int64_t nx = a.vmax[0] - a.vmin[0], nl = a.vmax[1] - a.vmin[1], nk = a.vmax[2] - a.vmin[2];
    return (uint32_t)(((nl * nk + kWaves - 1) / kWaves) * ((nx + 64 * kBwdJ - 1) / (64 * kBwdJ)));
}

struct Engine {
    hipStream_t stream = nullptr;
    Table *host[2] = {nullptr, nullptr};
    Table *dev[2] = {nullptr, nullptr};
    uint32_t *counters[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipEvent_t start[2] = {nullptr, nullptr}, stop[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
};

namespace {

int capture(Engine &e, int k, int cus)
{
    XPR_HIP_CHECK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    const dim3 blk(kThreads);
    // grids are multiples of kXcds, so every XCD residue class has workgroups.
    // Workgroups per CU: the projectors fit 4 (forward: 128 VGPRs, 28 KB LDS)
    // and 7 (backward) workgroups of 4 waves per CU; the defaults leave one
    // slot per CU free so the small launches of jobs running beside the steps
    // start at once (XPR_SERVE_SLOTS="forward,backward,elementwise" sets them).
    int per_cu[3] = {3, 6, 2};
    if (const char *v = std::getenv("XPR_SERVE_SLOTS"))
        std::sscanf(v, "%d,%d,%d", &per_cu[0], &per_cu[1], &per_cu[2]);
    for (int &q : per_cu) q = std::max(1, std::min(q, 16));
    auto grid = [&](int per) { return dim3((unsigned)(((cus * per + 7) / 8) * 8)); };
    (void)hipMemcpyAsync(e.dev[k], e.host[k], sizeof(Table), hipMemcpyHostToDevice, e.stream);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), blk, 0, e.stream, e.counters[k]);
    hipLaunchKernelGGL(k_batch_elementwise, grid(per_cu[2]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kConvert);
    hipLaunchKernelGGL(k_batch_forward, grid(per_cu[0]), blk, 0, e.stream, e.dev[k], e.counters[k]);
    hipLaunchKernelGGL(k_batch_elementwise, grid(per_cu[2]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kError);
    hipLaunchKernelGGL(k_batch_backward, grid(per_cu[1]), blk, 0, e.stream, e.dev[k], e.counters[k]);
    hipLaunchKernelGGL(k_batch_elementwise, grid(per_cu[2]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kFill);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(e.stream, &e.graph[k]);
    if (le != hipSuccess) return fail(XPR_E_HIP, std::string("batch capture: ") + hipGetErrorString(le));
    if (ce != hipSuccess) return fail(XPR_E_HIP, std::string("batch capture: ") + hipGetErrorString(ce));
    XPR_HIP_CHECK(hipGraphInstantiate(&e.exec[k], e.graph[k], nullptr, nullptr, 0));
    return XPR_OK;
}

} // namespace

int engine_create(Engine **out)
{
    auto *e = new Engine();
    auto bad = [&](int rc) { engine_destroy(e); return rc; };
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return bad(fail(XPR_E_HIP, "batch engine: device query"));
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(fail(XPR_E_HIP, "batch engine: stream"));
    for (int k = 0; k < 2; ++k) {
        if (hipHostMalloc((void **)&e->host[k], sizeof(Table), hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void **)&e->dev[k], sizeof(Table)) != hipSuccess ||
            hipMalloc((void **)&e->counters[k], sizeof(uint32_t) * kPhases * kXcds) != hipSuccess ||
            hipEventCreate(&e->start[k]) != hipSuccess || hipEventCreate(&e->stop[k]) != hipSuccess)
            return bad(fail(XPR_E_NOMEM, "batch engine: allocation"));
        std::memset(e->host[k], 0, sizeof(Table));
        const int rc = capture(*e, k, prop.multiProcessorCount);
        if (rc != XPR_OK) return bad(rc);
    }
    *out = e;
    return XPR_OK;
}

Table *engine_table(Engine *e, int slot) { return e->host[slot & 1]; }

int engine_launch(Engine *e, int slot)
{
    const int k = slot & 1;
    XPR_HIP_CHECK(hipEventRecord(e->start[k], e->stream));
    XPR_HIP_CHECK(hipGraphLaunch(e->exec[k], e->stream));
    XPR_HIP_CHECK(hipEventRecord(e->stop[k], e->stream));
    e->pending[k] = true;
    return XPR_OK;
}

int engine_wait(Engine *e, int slot, float *ms)
{
    const int k = slot & 1;
    if (ms) *ms = 0.0f;
    if (!e->pending[k]) return XPR_OK;
    XPR_HIP_CHECK(hipEventSynchronize(e->stop[k]));
    e->pending[k] = false;
    if (ms) XPR_HIP_CHECK(hipEventElapsedTime(ms, e->start[k], e->stop[k]));
    return XPR_OK;
}

void engine_destroy(Engine *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (int k = 0; k < 2; ++k) {
        if (e->exec[k]) (void)hipGraphExecDestroy(e->exec[k]);
        if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
        if (e->host[k]) (void)hipHostFree(e->host[k]);
        if (e->dev[k]) (void)hipFree(e->dev[k]);
        if (e->counters[k]) (void)hipFree(e->counters[k]);
        if (e->start[k]) (void)hipEventDestroy(e->start[k]);
        if (e->stop[k]) (void)hipEventDestroy(e->stop[k]);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

} // namespace xpr_batch

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int xpr_abi_version(void) { return XPR_ABI_VERSION; }

const char *xpr_last_error(void) { return g_last_error.c_str(); }

int xpr_set_device(int device)
{
    XPR_HIP_CHECK(hipSetDevice(device));
    return XPR_OK;
}

int xpr_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return fail(XPR_E_INVALID, "null output");
    XPR_HIP_CHECK(hipMemGetInfo(free_bytes, total_bytes));
    return XPR_OK;
}

int xpr_alloc(void **dptr, size_t bytes)
{
    if (!dptr) return fail(XPR_E_INVALID, "null output");
    *dptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(XPR_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " +
                                     hipGetErrorString(e));
    }
    return XPR_OK;
}

int xpr_free(void *dptr)
{
    XPR_HIP_CHECK(hipFree(dptr));
    return XPR_OK;
}

int xpr_host_alloc(void **hptr, size_t bytes)
{
    if (!hptr) return fail(XPR_E_INVALID, "null output");
    hipError_t e = hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(XPR_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    return XPR_OK;
}

int xpr_alloc_async(void **dptr, size_t bytes, xpr_stream s)
{
    if (!dptr) return fail(XPR_E_INVALID, "null output");
    *dptr = nullptr;
    hipError_t e = hipMallocAsync(dptr, bytes ? bytes : 1, (hipStream_t)s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(XPR_E_NOMEM, "hipMallocAsync(" + std::to_string(bytes) + "): " +
                                     hipGetErrorString(e));
    }
    return XPR_OK;
}

int xpr_free_async(void *dptr, xpr_stream s)
{
    XPR_HIP_CHECK(hipFreeAsync(dptr, (hipStream_t)s));
    return XPR_OK;
}

int xpr_host_free(void *hptr)
{
    XPR_HIP_CHECK(hipHostFree(hptr));
    return XPR_OK;
}

int xpr_upload(xpr_stream s, void *dst, const void *src, size_t bytes)
{
    XPR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s));
    return XPR_OK;
}

int xpr_download(xpr_stream s, void *dst, const void *src, size_t bytes)
{
    XPR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s));
    return XPR_OK;
}

int xpr_stream_create(xpr_stream *s)
{
    if (!s) return fail(XPR_E_INVALID, "null output");
    hipStream_t h;
    XPR_HIP_CHECK(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = (xpr_stream)h;
    return XPR_OK;
}

int xpr_stream_destroy(xpr_stream s)
{
    XPR_HIP_CHECK(hipStreamDestroy((hipStream_t)s));
    return XPR_OK;
}

int xpr_stream_sync(xpr_stream s)
{
    XPR_HIP_CHECK(hipStreamSynchronize((hipStream_t)s));
    return XPR_OK;
}

int xpr_event_create(xpr_event *e)
{
    if (!e) return fail(XPR_E_INVALID, "null output");
    hipEvent_t h;
    XPR_HIP_CHECK(hipEventCreate(&h));
    *e = (xpr_event)h;
    return XPR_OK;
}

int xpr_event_destroy(xpr_event e)
{
    XPR_HIP_CHECK(hipEventDestroy((hipEvent_t)e));
    return XPR_OK;
}

int xpr_event_record(xpr_event e, xpr_stream s)
{
    XPR_HIP_CHECK(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
    return XPR_OK;
}

int xpr_stream_wait_event(xpr_stream s, xpr_event e)
{
    XPR_HIP_CHECK(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)e, 0));
    return XPR_OK;
}

int xpr_event_sync(xpr_event e)
{
    XPR_HIP_CHECK(hipEventSynchronize((hipEvent_t)e));
    return XPR_OK;
}

int xpr_event_elapsed_ms(xpr_event start, xpr_event stop, float *ms)
{
    if (!ms) return fail(XPR_E_INVALID, "null output");
    XPR_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return XPR_OK;
}

int xpr_fill_u32(xpr_stream s, uint32_t *dst, uint32_t value, uint64_t n)
{
    if (n == 0) return XPR_OK;
    if (!dst) return fail(XPR_E_INVALID, "null buffer");
    hipLaunchKernelGGL(k_fill_raw<uint32_t>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)s,
                       dst, value, n);
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

int xpr_fill_u64(xpr_stream s, uint64_t *dst, uint64_t value, uint64_t n)
{
    if (n == 0) return XPR_OK;
    if (!dst) return fail(XPR_E_INVALID, "null buffer");
    hipLaunchKernelGGL(k_fill_raw<uint64_t>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)s,
                       dst, value, n);
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

int xpr_fill(xpr_stream s, void *dst, int dtype, float value, uint64_t n)
{
    if (int rc = check_dtype(dtype)) return rc;
    if (n == 0) return XPR_OK;
    if (!dst) return fail(XPR_E_INVALID, "null buffer");
    if (dtype == XPR_F32) {
        hipLaunchKernelGGL(k_fill_raw<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)s,
                           (float *)dst, value, n);
    } else {
        uint16_t bits;
        if (dtype == XPR_F16) {
            bits = __builtin_bit_cast(uint16_t, (_Float16)value);
        } else {
            uint32_t b;
            std::memcpy(&b, &value, 4);
            if ((b & 0x7fffffffu) > 0x7f800000u)
                bits = (uint16_t)((b >> 16) | 0x40u);
            else
                bits = (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
        }
        hipLaunchKernelGGL(k_fill_raw<uint16_t>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)s, (uint16_t *)dst, bits, n);
    }
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

int xpr_convert(xpr_stream s, void *dst, int dst_dtype, const void *src, int src_dtype,
                uint64_t n)
{
    if (int rc = check_dtype(dst_dtype)) return rc;
    if (int rc = check_dtype(src_dtype)) return rc;
    if (n == 0) return XPR_OK;
    if (!dst || !src) return fail(XPR_E_INVALID, "null buffer");
    hipStream_t st = (hipStream_t)s;
    switch (dst_dtype) {
    case XPR_F32: return convert_from<F32>(st, dst, src, src_dtype, n);
    case XPR_F16: return convert_from<F16>(st, dst, src, src_dtype, n);
    default: return convert_from<BF16>(st, dst, src, src_dtype, n);
    }
}

int xpr_forward(xpr_stream s, const void *volume, int vol_dtype, void *sums, int acc_bits,
                const float *projdata, const xpr_proj_args *a)
{
    SlabShape sh;
    if (int rc = check_args(a, sh)) return rc;
    if (int rc = check_dtype(vol_dtype)) return rc;
    if (acc_bits != XPR_ACC_U32 && acc_bits != XPR_ACC_U64)
        return fail(XPR_E_INVALID, "acc_bits must be 32 or 64");
    if (a->n_views == 0 || sh.nx == 0 || sh.nrows == 0) return XPR_OK;
    if (!volume || !sums || !projdata) return fail(XPR_E_INVALID, "null buffer");
    hipStream_t st = (hipStream_t)s;
    if (acc_bits == XPR_ACC_U32) {
        switch (vol_dtype) {
        case XPR_F32: return forward_impl<F32, uint32_t>(st, volume, sums, projdata, *a, sh);
        case XPR_F16: return forward_impl<F16, uint32_t>(st, volume, sums, projdata, *a, sh);
        default: return forward_impl<BF16, uint32_t>(st, volume, sums, projdata, *a, sh);
        }
    }
    switch (vol_dtype) {
    case XPR_F32: return forward_impl<F32, uint64_t>(st, volume, sums, projdata, *a, sh);
    case XPR_F16: return forward_impl<F16, uint64_t>(st, volume, sums, projdata, *a, sh);
    default: return forward_impl<BF16, uint64_t>(st, volume, sums, projdata, *a, sh);
    }
}

int xpr_error(xpr_stream s, const void *images, int img_dtype, const void *sums, int acc_bits,
              void *err, int err_dtype, uint64_t n, float weight)
{
    if (int rc = check_dtype(img_dtype)) return rc;
    if (int rc = check_dtype(err_dtype)) return rc;
    if (acc_bits != XPR_ACC_U32 && acc_bits != XPR_ACC_U64)
        return fail(XPR_E_INVALID, "acc_bits must be 32 or 64");
    if (n == 0) return XPR_OK;
    if (!images || !sums || !err) return fail(XPR_E_INVALID, "null buffer");
    if (err == sums && !(acc_bits == XPR_ACC_U32 && err_dtype == XPR_F32))
        return fail(XPR_E_INVALID, "in-place error needs 32-bit sums and f32 error");
    hipStream_t st = (hipStream_t)s;
    switch (img_dtype) {
    case XPR_F32: return error_impl<F32>(st, images, sums, acc_bits, err, err_dtype, n, weight);
    case XPR_F16: return error_impl<F16>(st, images, sums, acc_bits, err, err_dtype, n, weight);
    default: return error_impl<BF16>(st, images, sums, acc_bits, err, err_dtype, n, weight);
    }
}

int xpr_backward(xpr_stream s, void *volume, int vol_dtype, const void *err, int err_dtype,
                 const float *projdata, const xpr_proj_args *a, uint64_t *empty_count)
{
    SlabShape sh;
    if (int rc = check_args(a, sh)) return rc;
    if (int rc = check_dtype(vol_dtype)) return rc;
    if (int rc = check_dtype(err_dtype)) return rc;
    if (sh.nx == 0 || sh.nrows == 0) return XPR_OK;
    if (!volume || !err || !projdata) return fail(XPR_E_INVALID, "null buffer");
    hipStream_t st = (hipStream_t)s;
#define XPR_BW(VS, ES) backward_impl<VS, ES>(st, volume, err, projdata, *a, sh, empty_count)
    switch (vol_dtype * 3 + err_dtype) {
    case XPR_F32 * 3 + XPR_F32: return XPR_BW(F32, F32);
    case XPR_F32 * 3 + XPR_F16: return XPR_BW(F32, F16);
    case XPR_F32 * 3 + XPR_BF16: return XPR_BW(F32, BF16);
    case XPR_F16 * 3 + XPR_F32: return XPR_BW(F16, F32);
    case XPR_F16 * 3 + XPR_F16: return XPR_BW(F16, F16);
    case XPR_F16 * 3 + XPR_BF16: return XPR_BW(F16, BF16);
    case XPR_BF16 * 3 + XPR_F32: return XPR_BW(BF16, F32);
    case XPR_BF16 * 3 + XPR_F16: return XPR_BW(BF16, F16);
    default: return XPR_BW(BF16, BF16);
    }
#undef XPR_BW
}

// ---- tile-shape tuning --------------------------------------------------------
int xpr_tune_forward(xpr_stream s, const void *volume, int vol_dtype, void *sums_scratch,
                     int acc_bits, const float *projdata, const xpr_proj_args *a, int *choice)
{
    SlabShape sh;
    if (int rc = check_args(a, sh)) return rc;
    if (int rc = check_dtype(vol_dtype)) return rc;
    if (acc_bits != XPR_ACC_U32 && acc_bits != XPR_ACC_U64)
        return fail(XPR_E_INVALID, "acc_bits must be 32 or 64");
    if (!volume || !sums_scratch || !projdata || !choice) return fail(XPR_E_INVALID, "null buffer");
    if (a->n_views == 0 || sh.nx == 0 || sh.nrows == 0) return fail(XPR_E_INVALID, "empty launch");
    hipStream_t st = (hipStream_t)s;
    float best = 3.4e38f;
    int pick = kTileCap;
    for (int cap : kTileCaps) {
        auto launch = [&]() {
            if (acc_bits == XPR_ACC_U32) {
                switch (vol_dtype) {
                case XPR_F32: return forward_impl<F32, uint32_t>(st, volume, sums_scratch, projdata, *a, sh, cap);
                case XPR_F16: return forward_impl<F16, uint32_t>(st, volume, sums_scratch, projdata, *a, sh, cap);
                default: return forward_impl<BF16, uint32_t>(st, volume, sums_scratch, projdata, *a, sh, cap);
                }
            }
            switch (vol_dtype) {
            case XPR_F32: return forward_impl<F32, uint64_t>(st, volume, sums_scratch, projdata, *a, sh, cap);
            case XPR_F16: return forward_impl<F16, uint64_t>(st, volume, sums_scratch, projdata, *a, sh, cap);
            default: return forward_impl<BF16, uint64_t>(st, volume, sums_scratch, projdata, *a, sh, cap);
            }
        };
        float ms = 0.0f;
        if (int rc = time_best(st, 2, launch, &ms)) return rc;
        if (ms < best) {
            best = ms;
            pick = cap;
        }
    }
    const int acc_bytes = acc_bits / 8;
    tune_store(tune_key(kTuneForward, vol_dtype, acc_bytes, *a, sh), pick);
    *choice = pick;
    return XPR_OK;
}

int xpr_tune_backward(xpr_stream s, void *volume_scratch, int vol_dtype, const void *err,
                      int err_dtype, const float *projdata, const xpr_proj_args *a, int *choice)
{
    SlabShape sh;
    if (int rc = check_args(a, sh)) return rc;
    if (int rc = check_dtype(vol_dtype)) return rc;
    if (int rc = check_dtype(err_dtype)) return rc;
    if (!volume_scratch || !err || !projdata || !choice) return fail(XPR_E_INVALID, "null buffer");
    if (sh.nx == 0 || sh.nrows == 0) return fail(XPR_E_INVALID, "empty launch");
    if (vol_dtype != err_dtype) {                // one shape compiled for mixed storage pairs
        *choice = kBwdJ;
        return XPR_OK;
    }
    hipStream_t st = (hipStream_t)s;
    float best = 3.4e38f;
    int pick = kBwdJ;
    for (int J : kBwdJs) {
        auto launch = [&]() {
            switch (vol_dtype) {
            case XPR_F32: return backward_impl<F32, F32>(st, volume_scratch, err, projdata, *a, sh, nullptr, J);
            case XPR_F16: return backward_impl<F16, F16>(st, volume_scratch, err, projdata, *a, sh, nullptr, J);
            default: return backward_impl<BF16, BF16>(st, volume_scratch, err, projdata, *a, sh, nullptr, J);
            }
        };
        float ms = 0.0f;
        if (int rc = time_best(st, 2, launch, &ms)) return rc;
        if (ms < best) {
            best = ms;
            pick = J;
        }
    }
    tune_store(tune_key(kTuneBackward, vol_dtype, err_dtype, *a, sh), pick);
    *choice = pick;
    return XPR_OK;
}

int xpr_tune_reset(void)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune.clear();
    return XPR_OK;
}

int xpr_q22_to_float(xpr_stream s, const void *sums, int acc_bits, float *out, uint64_t n)
{
    if (acc_bits != XPR_ACC_U32 && acc_bits != XPR_ACC_U64)
        return fail(XPR_E_INVALID, "acc_bits must be 32 or 64");
    if (n == 0) return XPR_OK;
    if (!sums || !out) return fail(XPR_E_INVALID, "null buffer");
    if (acc_bits == XPR_ACC_U32)
        hipLaunchKernelGGL(k_q22_to_float<uint32_t>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)s, (const uint32_t *)sums, out, n);
    else
        hipLaunchKernelGGL(k_q22_to_float<uint64_t>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)s, (const uint64_t *)sums, out, n);
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

int xpr_minmax_init(xpr_stream s, int32_t *minmax)
{
    if (!minmax) return fail(XPR_E_INVALID, "null buffer");
    // running values start at FLT_MAX / FLT_MIN (data_loader.hpp:287-288)
    const uint32_t init[2] = {0x7f7fffffu, 0x00800000u};
    hipLaunchKernelGGL(k_fill_raw<uint32_t>, dim3(1), dim3(64), 0, (hipStream_t)s,
                       (uint32_t *)minmax, init[0], (uint64_t)1);
    hipLaunchKernelGGL(k_fill_raw<uint32_t>, dim3(1), dim3(64), 0, (hipStream_t)s,
                       (uint32_t *)minmax + 1, init[1], (uint64_t)1);
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

int xpr_minmax_normal(xpr_stream s, const float *data, uint64_t n, int32_t *minmax)
{
    if (n == 0) return XPR_OK;
    if (!data || !minmax) return fail(XPR_E_INVALID, "null buffer");
    hipLaunchKernelGGL(k_minmax, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)s, data, n,
                       minmax);
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

int xpr_minmax_read(const int32_t *minmax_host, float *mn, float *mx)
{
    if (!minmax_host || !mn || !mx) return fail(XPR_E_INVALID, "null pointer");
    for (int i = 0; i < 2; ++i) {
        int32_t k = minmax_host[i];
        int32_t b = k < 0 ? (k ^ 0x7fffffff) : k;
        float f;
        std::memcpy(&f, &b, 4);
        (i == 0 ? *mn : *mx) = f;
    }
    return XPR_OK;
}

int xpr_preprocess(xpr_stream s, const float *in, void *out, int out_dtype, uint64_t n,
                   int normalize, int mlog, float mn, float mx)
{
    if (int rc = check_dtype(out_dtype)) return rc;
    if (n == 0) return XPR_OK;
    if (!in || !out) return fail(XPR_E_INVALID, "null buffer");
    hipStream_t st = (hipStream_t)s;
    const unsigned g = grid_for(n);
    switch (out_dtype) {
    case XPR_F32:
        hipLaunchKernelGGL(k_preprocess<F32>, dim3(g), dim3(256), 0, st, in, (float *)out, n,
                           normalize, mlog, mn, mx);
        break;
    case XPR_F16:
        hipLaunchKernelGGL(k_preprocess<F16>, dim3(g), dim3(256), 0, st, in, (uint16_t *)out,
                           n, normalize, mlog, mn, mx);
        break;
    default:
        hipLaunchKernelGGL(k_preprocess<BF16>, dim3(g), dim3(256), 0, st, in, (uint16_t *)out,
                           n, normalize, mlog, mn, mx);
        break;
    }
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

} // extern "C"
