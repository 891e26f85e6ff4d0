// Device building blocks: storage types, the canonical projection and its
// exact fast forms, buffer loads, launch shapes, Q.22 conversion.
// (library-internal; every header here is included once, by xpr_kernels.hip,
// into one translation unit)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/xpr_hip.h"

namespace {

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

// lane `lane`'s x, uniform (v_readlane)
__device__ __forceinline__ float lane_float(float x, uint32_t lane)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
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

// The backward keeps a 16-bit sample in its storage form until it is added:
// a half sample joins the sum in one v_fma_mix_f32 (h * 1 + sum rounds once,
// exactly like the float add of the widened sample), instead of a conversion
// and an add.
template <typename T> struct SampleOf { using type = float; };
template <> struct SampleOf<F16> { using type = uint16_t; };
template <> struct SampleOf<BF16> { using type = uint16_t; };

template <typename T>
__device__ __forceinline__ typename SampleOf<T>::type sample_raw(__amdgpu_buffer_rsrc_t r,
                                                                 uint32_t index);
template <>
__device__ __forceinline__ float sample_raw<F32>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return struct_buffer_load_f32(r, index, 0, 0, 0);
}
template <>
__device__ __forceinline__ uint16_t sample_raw<F16>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return struct_buffer_load_u16(r, index, 0, 0, 0);
}
template <>
__device__ __forceinline__ uint16_t sample_raw<BF16>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return struct_buffer_load_u16(r, index, 0, 0, 0);
}

// an indexed sample widened to float
template <typename T>
__device__ __forceinline__ float sample_f32(__amdgpu_buffer_rsrc_t r, uint32_t index);
template <>
__device__ __forceinline__ float sample_f32<F32>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return struct_buffer_load_f32(r, index, 0, 0, 0);
}
template <>
__device__ __forceinline__ float sample_f32<F16>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return (float)__builtin_bit_cast(_Float16, struct_buffer_load_u16(r, index, 0, 0, 0));
}
template <>
__device__ __forceinline__ float sample_f32<BF16>(__amdgpu_buffer_rsrc_t r, uint32_t index)
{
    return __builtin_bit_cast(float, (uint32_t)struct_buffer_load_u16(r, index, 0, 0, 0) << 16);
}

template <typename T>
__device__ __forceinline__ float add_sample(float sum, typename SampleOf<T>::type s);
template <>
__device__ __forceinline__ float add_sample<F32>(float sum, float s)
{
    return sum + s;
}
template <>
__device__ __forceinline__ float add_sample<F16>(float sum, uint16_t s)
{
    // 1.0 through a scalar register the compiler cannot see into: a literal
    // 1.0 lets it rewrite the FMA as an add of the converted sample
    uint32_t one;
    asm("s_mov_b32 %0, 0x3f800000" : "=s"(one));
    return __builtin_fmaf((float)__builtin_bit_cast(_Float16, s), __builtin_bit_cast(float, one), sum);
}
template <>
__device__ __forceinline__ float add_sample<BF16>(float sum, uint16_t s)
{
    return sum + __builtin_bit_cast(float, (uint32_t)s << 16);
}

// ---------------------------------------------------------------------------
// launch shapes shared by the projectors and the batch engine
// ---------------------------------------------------------------------------
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kBwdJ = 2;          // backward default shape: x positions per lane
constexpr int kBwdL = 4;          // and layers per wave row

struct SlabShape {
    int32_t nx, nl, nk;
    int32_t nseg;          // x segments per row
    int64_t row_base;      // first row of this launch
    int64_t nrows;         // wave rows in the slab (backward: layer groups x depths)
    int32_t inside = 1;    // 1: backward views wholly in bounds skip the per-voxel test
};

// The value of a detector sum: FIXED_TO_FLOAT of a Q.22 integer
// (constants.cl:21), or the float sum itself (XPR_ACC_F32).
template <typename AccT>
__device__ __forceinline__ float sum_value(AccT s) { return (float)s / kQ22One; }
template <>
__device__ __forceinline__ float sum_value<float>(float s) { return s; }

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

} // namespace
