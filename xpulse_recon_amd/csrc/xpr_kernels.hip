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

#include <algorithm>

#include <cstddef>
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

} // namespace

#include "kernels/common.hpp"
#include "kernels/backward.hpp"
#include "kernels/forward.hpp"
#include "kernels/elementwise.hpp"
#include "kernels/batch.hpp"
#include "kernels/calibrate.hpp"

namespace {

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

// the projectors take xpr_proj_args by value: interp sits in what was the
// padding before view_index, so the argument block kept its size and offsets
static_assert(sizeof(xpr_proj_args) == 80 && offsetof(xpr_proj_args, interp) == 68 &&
                  offsetof(xpr_proj_args, view_index) == 72,
              "xpr_proj_args layout");

int check_acc(int acc_bits)
{
    if (acc_bits != XPR_ACC_U32 && acc_bits != XPR_ACC_U64 && acc_bits != XPR_ACC_F32)
        return fail(XPR_E_INVALID, "acc_bits must be XPR_ACC_U32, XPR_ACC_U64 or XPR_ACC_F32");
    return XPR_OK;
}

// bytes per detector sum
uint64_t acc_bytes(int acc_bits) { return acc_bits == XPR_ACC_U64 ? 8u : 4u; }

// the acc_bits value of a sums type (tuning-table key)
template <typename AccT>
constexpr int acc_code()
{
    return std::is_same<AccT, float>::value ? XPR_ACC_F32 : (int)(8 * sizeof(AccT));
}

int check_args(const xpr_proj_args *a, SlabShape &s)
{
    if (a == nullptr) return fail(XPR_E_INVALID, "null xpr_proj_args");
    if (a->interp != XPR_INTERP_NEAREST && a->interp != XPR_INTERP_BILINEAR)
        return fail(XPR_E_INVALID, "interp must be XPR_INTERP_NEAREST or XPR_INTERP_BILINEAR");
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

// backward launches: wave rows of L layers, segments of 64*J voxels along x
template <typename Launch>
int launch_rows(SlabShape s, int J, int L, Launch &&launch)
{
    if (s.nx == 0 || s.nrows == 0) return XPR_OK;
    s.nseg = (s.nx + 64 * J - 1) / (64 * J);
    s.nrows = backward_rows(s.nl, s.nk, L);
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

// backward shapes by voxels per lane, J x positions times L layers:
// 4 = 1 x 4, 8 = 2 x 4, 16 = 4 x 4
constexpr int kBwdJs[] = {4, 8, 16};

// The tuned shape: the default unless another is more than 3 % faster, so
// timing noise on the tuning launch cannot lock in a shape that is no better.
template <size_t N>
int pick_shape(const int (&cands)[N], const float (&ms)[N], int fallback)
{
    float t_def = 3.4e38f;
    for (size_t i = 0; i < N; ++i)
        if (cands[i] == fallback) t_def = ms[i];
    int pick = fallback;
    float best = 0.97f * t_def;
    for (size_t i = 0; i < N; ++i)
        if (ms[i] < best) {
            best = ms[i];
            pick = cands[i];
        }
    return pick;
}

// default voxels per lane (kBwdJ x kBwdL); narrow slabs take one x position
// per lane so one 64*J segment leaves few lanes idle
int backward_default_j(int32_t nx) { return nx <= 128 ? kBwdL : kBwdJ * kBwdL; }

template <typename VS, typename ES, int J, int L, bool BILIN = false>
int backward_launch(hipStream_t st, void *vol, const void *err, const float *pd,
                    const xpr_proj_args &a, SlabShape s, uint64_t *empty, uint32_t *max_bits)
{
    const bool crop = a.line_offset != 0;
    return launch_rows(s, J, L, [&](unsigned blocks, const SlabShape &t) {
        if (crop)
            hipLaunchKernelGGL((k_backward<VS, ES, J, L, true, BILIN>), dim3(blocks),
                               dim3(kThreads), 0, st, (typename VS::T *)vol,
                               (const typename ES::T *)err, pd, a, t, (unsigned long long *)empty,
                               max_bits);
        else
            hipLaunchKernelGGL((k_backward<VS, ES, J, L, false, BILIN>), dim3(blocks),
                               dim3(kThreads), 0, st, (typename VS::T *)vol,
                               (const typename ES::T *)err, pd, a, t, (unsigned long long *)empty,
                               max_bits);
    });
}

// J < 0: the table's choice for this shape, else the default
template <typename VS, typename ES>
int backward_impl(hipStream_t st, void *vol, const void *err, const float *pd,
                  const xpr_proj_args &a, SlabShape s, uint64_t *empty, int J = -1,
                  uint32_t *max_bits = nullptr)
{
    if ((uint64_t)a.det_width * (uint64_t)a.line_rows * sizeof(typename ES::T) >= 0x80000000ull)
        return fail(XPR_E_INVALID, "a detector view above 2 GiB");
    // tuning knob: XPR_BACKWARD_INSIDE=0 keeps the canonical bounds test in every view
    if (const char *e = std::getenv("XPR_BACKWARD_INSIDE")) s.inside = std::atoi(e) != 0;
    // bilinear sampling (opt-in): the default shape only
    if (a.interp == XPR_INTERP_BILINEAR)
        return backward_launch<VS, ES, kBwdJ, kBwdL, true>(st, vol, err, pd, a, s, empty, max_bits);
    if constexpr (std::is_same<VS, ES>::value) {       // the schedules' dtype pairs: all J compiled
        if (J < 0) {
            // tuning knob: force the voxels per lane
            if (const char *e = std::getenv("XPR_BACKWARD_J")) J = std::atoi(e);
            else J = tune_lookup(tune_key(kTuneBackward, VS::code, ES::code, a, s));
        }
        if (J < 0) J = backward_default_j(s.nx);
        switch (J) {
        case 4: return backward_launch<VS, ES, 1, 4>(st, vol, err, pd, a, s, empty, max_bits);
        case 8: return backward_launch<VS, ES, 2, 4>(st, vol, err, pd, a, s, empty, max_bits);
        case 16: return backward_launch<VS, ES, 4, 4>(st, vol, err, pd, a, s, empty, max_bits);
        default: return backward_launch<VS, ES, kBwdJ, kBwdL>(st, vol, err, pd, a, s, empty, max_bits);
        }
    } else {
        (void)J;
        return backward_launch<VS, ES, kBwdJ, kBwdL>(st, vol, err, pd, a, s, empty, max_bits);
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
        cap = tune_lookup(tune_key(kTuneForward, VS::code, acc_code<AccT>(), a, s));
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
    if (acc_bits == XPR_ACC_F32)
        return error_impl2<IS, float>(st, img, sums, err, err_dtype, n, w);
    return error_impl2<IS, uint64_t>(st, img, sums, err, err_dtype, n, w);
}

// forward launch for the sums type acc_bits names (cap as forward_impl)
template <typename AccT>
int forward_dtype(hipStream_t st, const void *vol, int vol_dtype, void *sums, const float *pd,
                  const xpr_proj_args &a, const SlabShape &s, int cap)
{
    switch (vol_dtype) {
    case XPR_F32: return forward_impl<F32, AccT>(st, vol, sums, pd, a, s, cap);
    case XPR_F16: return forward_impl<F16, AccT>(st, vol, sums, pd, a, s, cap);
    default: return forward_impl<BF16, AccT>(st, vol, sums, pd, a, s, cap);
    }
}

int forward_dispatch(hipStream_t st, const void *vol, int vol_dtype, void *sums, int acc_bits,
                     const float *pd, const xpr_proj_args &a, const SlabShape &s, int cap = -1)
{
    if (acc_bits == XPR_ACC_U32) return forward_dtype<uint32_t>(st, vol, vol_dtype, sums, pd, a, s, cap);
    if (acc_bits == XPR_ACC_F32) return forward_dtype<float>(st, vol, vol_dtype, sums, pd, a, s, cap);
    return forward_dtype<uint64_t>(st, vol, vol_dtype, sums, pd, a, s, cap);
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

// ---------------------------------------------------------------------------
// Scratch of xpr_subiteration_static: one array per GPU, held from module
// load until the process ends.  Host code never passes `scratch` itself (on
// the host it is a stand-in symbol, not a device address); it uses the
// device address hipGetSymbolAddress returns, looked up once per device.
// ---------------------------------------------------------------------------
namespace {

static __device__ __align__(256) unsigned char scratch[XPR_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int static_scratch(void **out)
{
    int dev = 0;
    XPR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(XPR_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(XPR_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                       hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return XPR_OK;
}

} // namespace

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
    const int64_t rows = backward_rows((int32_t)nl, (int32_t)nk, kBwdL);
    return (uint32_t)(((rows + kWaves - 1) / kWaves) * ((nx + 64 * kBwdJ - 1) / (64 * kBwdJ)));
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
    hipLaunchKernelGGL(k_batch_forward<F32>, grid(per_cu[0]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kForward);
    hipLaunchKernelGGL(k_batch_forward<F16>, grid(per_cu[0]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kForwardF16);
    hipLaunchKernelGGL(k_batch_forward<BF16>, grid(per_cu[0]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kForwardBF16);
    hipLaunchKernelGGL(k_batch_elementwise, grid(per_cu[2]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kError);
    hipLaunchKernelGGL(k_batch_backward<F32>, grid(per_cu[1]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kBackward);
    hipLaunchKernelGGL(k_batch_backward<F16>, grid(per_cu[1]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kBackwardF16);
    hipLaunchKernelGGL(k_batch_backward<BF16>, grid(per_cu[1]), blk, 0, e.stream, e.dev[k], e.counters[k],
                       (uint32_t)kBackwardBF16);
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

void engine_sync(Engine *e)
{
    if (e && e->stream) (void)hipStreamSynchronize(e->stream);
    if (e) e->pending[0] = e->pending[1] = false;
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
    if (int rc = check_acc(acc_bits)) return rc;
    if (a->n_views == 0 || sh.nx == 0 || sh.nrows == 0) return XPR_OK;
    if (!volume || !sums || !projdata) return fail(XPR_E_INVALID, "null buffer");
    return forward_dispatch((hipStream_t)s, volume, vol_dtype, sums, acc_bits, projdata, *a, sh);
}

int xpr_error(xpr_stream s, const void *images, int img_dtype, const void *sums, int acc_bits,
              void *err, int err_dtype, uint64_t n, float weight)
{
    if (int rc = check_dtype(img_dtype)) return rc;
    if (int rc = check_dtype(err_dtype)) return rc;
    if (int rc = check_acc(acc_bits)) return rc;
    if (n == 0) return XPR_OK;
    if (!images || !sums || !err) return fail(XPR_E_INVALID, "null buffer");
    if (err == sums && !(acc_bits != XPR_ACC_U64 && err_dtype == XPR_F32))
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
    return xpr_backward_max(s, volume, vol_dtype, err, err_dtype, projdata, a, empty_count, nullptr);
}

int xpr_backward_max(xpr_stream s, void *volume, int vol_dtype, const void *err, int err_dtype,
                     const float *projdata, const xpr_proj_args *a, uint64_t *empty_count,
                     uint32_t *max_bits)
{
    SlabShape sh;
    if (int rc = check_args(a, sh)) return rc;
    if (int rc = check_dtype(vol_dtype)) return rc;
    if (int rc = check_dtype(err_dtype)) return rc;
    if (sh.nx == 0 || sh.nrows == 0) return XPR_OK;
    if (!volume || !err || !projdata) return fail(XPR_E_INVALID, "null buffer");
    hipStream_t st = (hipStream_t)s;
#define XPR_BW(VS, ES) backward_impl<VS, ES>(st, volume, err, projdata, *a, sh, empty_count, -1, max_bits)
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

int xpr_subiteration(xpr_stream s, void *volume, int vol_dtype, const void *images,
                     int img_dtype, const float *projdata, const xpr_proj_args *a, float weight,
                     int first_pass, void *sums, int acc_bits, void *err, int err_dtype,
                     uint64_t *empty_count)
{
    SlabShape sh;
    if (int rc = check_args(a, sh)) return rc;
    if (int rc = check_dtype(vol_dtype)) return rc;
    if (int rc = check_dtype(img_dtype)) return rc;
    if (int rc = check_dtype(err_dtype)) return rc;
    if (int rc = check_acc(acc_bits)) return rc;
    if (a->view_index)
        return fail(XPR_E_INVALID, "xpr_subiteration: indexed views need xpr_forward / xpr_backward");
    if (!volume || !images || !projdata || !sums || !err) return fail(XPR_E_INVALID, "null buffer");
    if (err == sums && !(acc_bits != XPR_ACC_U64 && err_dtype == XPR_F32))
        return fail(XPR_E_INVALID, "in-place error needs 32-bit sums and f32 error");
    const uint64_t n = (uint64_t)a->n_views * (uint64_t)a->line_rows * (uint64_t)a->det_width;
    // sums start at 0 before a forward, at 1.0 (Q.22) on the first pass
    // (reconstruction_n.hpp:60,75): set here, so the workspace may hold anything
    const uint64_t init = first_pass ? (1ull << 22) : 0ull;
    int rc = acc_bits == XPR_ACC_U32   ? xpr_fill_u32(s, (uint32_t *)sums, (uint32_t)init, n)
             : acc_bits == XPR_ACC_F32 ? xpr_fill(s, sums, XPR_F32, first_pass ? 1.0f : 0.0f, n)
                                       : xpr_fill_u64(s, (uint64_t *)sums, init, n);
    if (rc == XPR_OK && !first_pass)
        rc = xpr_forward(s, volume, vol_dtype, sums, acc_bits, projdata, a);
    if (rc == XPR_OK)
        rc = xpr_error(s, images, img_dtype, sums, acc_bits, err, err_dtype, n, weight);
    if (rc == XPR_OK)
        rc = xpr_backward(s, volume, vol_dtype, err, err_dtype, projdata, a, empty_count);
    return rc;
}

int xpr_subiteration_static(xpr_stream s, void *volume, int vol_dtype, const void *images,
                            int img_dtype, const float *projdata, const xpr_proj_args *a,
                            float weight, int first_pass, int acc_bits, int err_dtype,
                            uint64_t *empty_count)
{
    if (a == nullptr) return fail(XPR_E_INVALID, "null xpr_proj_args");
    if (int rc = check_dtype(err_dtype)) return rc;
    if (int rc = check_acc(acc_bits)) return rc;
    if (a->line_rows < 0) return fail(XPR_E_INVALID, "line_rows < 0");
    const uint64_t per_view = (uint64_t)a->line_rows * (uint64_t)a->det_width;
    const uint64_t n = per_view * (uint64_t)a->n_views;
    const uint64_t pixel_bytes = acc_bytes(acc_bits) + (err_dtype == XPR_F32 ? 4u : 2u);
    // per_view bounded before the products, so they cannot wrap
    const bool fits = a->n_views == 0 ||
                      (per_view <= XPR_STATIC_SCRATCH_BYTES &&
                       n * pixel_bytes + 256 <= XPR_STATIC_SCRATCH_BYTES);
    if (!fits)
        return fail(XPR_E_INVALID,
                    "xpr_subiteration_static: " + std::to_string(a->n_views) + " views of " +
                        std::to_string(per_view) + " pixels need more than the " +
                        std::to_string((unsigned long long)XPR_STATIC_SCRATCH_BYTES) +
                        " bytes of static scratch (use xpr_subiteration with a workspace)");
    const uint64_t sums_bytes = (n * acc_bytes(acc_bits) + 255) & ~255ull;
    void *base = nullptr;
    if (int rc = static_scratch(&base)) return rc;
    unsigned char *p = (unsigned char *)base;
    return xpr_subiteration(s, volume, vol_dtype, images, img_dtype, projdata, a, weight,
                            first_pass, p, acc_bits, p + sums_bytes, err_dtype, empty_count);
}

// ---- tile-shape tuning --------------------------------------------------------
int xpr_tune_forward(xpr_stream s, const void *volume, int vol_dtype, void *sums_scratch,
                     int acc_bits, const float *projdata, const xpr_proj_args *a, int *choice)
{
    SlabShape sh;
    if (int rc = check_args(a, sh)) return rc;
    if (int rc = check_dtype(vol_dtype)) return rc;
    if (int rc = check_acc(acc_bits)) return rc;
    if (!volume || !sums_scratch || !projdata || !choice) return fail(XPR_E_INVALID, "null buffer");
    if (a->n_views == 0 || sh.nx == 0 || sh.nrows == 0) return fail(XPR_E_INVALID, "empty launch");
    hipStream_t st = (hipStream_t)s;
    float times[sizeof(kTileCaps) / sizeof(kTileCaps[0])];
    int ti = 0;
    for (int cap : kTileCaps) {
        auto launch = [&]() {
            return forward_dispatch(st, volume, vol_dtype, sums_scratch, acc_bits, projdata, *a, sh, cap);
        };
        if (int rc = time_best(st, 3, launch, &times[ti++])) return rc;
    }
    const int pick = pick_shape(kTileCaps, times, kTileCap);
    tune_store(tune_key(kTuneForward, vol_dtype, acc_bits, *a, sh), pick);
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
        *choice = kBwdJ * kBwdL;
        return XPR_OK;
    }
    hipStream_t st = (hipStream_t)s;
    float times[sizeof(kBwdJs) / sizeof(kBwdJs[0])];
    int ti = 0;
    for (int J : kBwdJs) {
        auto launch = [&]() {
            switch (vol_dtype) {
            case XPR_F32: return backward_impl<F32, F32>(st, volume_scratch, err, projdata, *a, sh, nullptr, J);
            case XPR_F16: return backward_impl<F16, F16>(st, volume_scratch, err, projdata, *a, sh, nullptr, J);
            default: return backward_impl<BF16, BF16>(st, volume_scratch, err, projdata, *a, sh, nullptr, J);
            }
        };
        if (int rc = time_best(st, 3, launch, &times[ti++])) return rc;
    }
    const int pick = pick_shape(kBwdJs, times, backward_default_j(sh.nx));
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

int xpr_check_fast_paths(xpr_stream s, uint64_t mismatches[3])
{
    if (!mismatches) return fail(XPR_E_INVALID, "null mismatches");
    hipStream_t st = (hipStream_t)s;
    unsigned long long *d = nullptr;
    XPR_HIP_CHECK(hipMallocAsync((void **)&d, 3 * sizeof(unsigned long long), st));
    int rc = XPR_OK;
    if (hipMemsetAsync(d, 0, 3 * sizeof(unsigned long long), st) != hipSuccess) {
        rc = fail(XPR_E_HIP, "memset");
    } else {
        hipLaunchKernelGGL(k_check_fast_paths, dim3(8192), dim3(256), 0, st, d);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(mismatches, d, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = fail(XPR_E_HIP, "fast-path check");
    }
    (void)hipFreeAsync(d, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

int xpr_valu_rates(xpr_stream s, double *rates, int n_rates)
{
    if (!rates || n_rates < 1) return fail(XPR_E_INVALID, "null output");
    hipStream_t st = (hipStream_t)s;
    int dev = 0;
    hipDeviceProp_t prop;
    XPR_HIP_CHECK(hipGetDevice(&dev));
    XPR_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    const unsigned blocks = (unsigned)prop.multiProcessorCount * 8u;
    const uint32_t iters = 4096;
    float *out = nullptr;
    XPR_HIP_CHECK(hipMallocAsync((void **)&out, (size_t)blocks * 256 * sizeof(float), st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = XPR_OK;
    const int kinds = std::min(n_rates, kValuKinds);
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(XPR_E_HIP, "event");
    const double lane_instr = (double)blocks * 256.0 * (double)iters * (double)kValuChains;
    for (int k = 0; k < kinds && rc == XPR_OK; ++k) {
        float best = 3.4e38f;
        // one warm-up, then the best of five
        for (int r = 0; r < 6 && rc == XPR_OK; ++r) {
            (void)hipEventRecord(e0, st);
            switch (k) {
            case 0: hipLaunchKernelGGL(k_valu_ceiling<0>, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f); break;
            case 1: hipLaunchKernelGGL(k_valu_ceiling<1>, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f); break;
            case 2: hipLaunchKernelGGL(k_valu_ceiling<2>, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f); break;
            case 3: hipLaunchKernelGGL(k_valu_ceiling<3>, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f); break;
            case 4: hipLaunchKernelGGL(k_valu_ceiling<4>, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f); break;
            default: hipLaunchKernelGGL(k_valu_ceiling<5>, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f); break;
            }
            (void)hipEventRecord(e1, st);
            float ms = 0.0f;
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                rc = fail(XPR_E_HIP, "VALU ceiling launch");
            else if (r != 0 && ms < best)
                best = ms;
        }
        if (rc == XPR_OK) rates[k] = lane_instr / ((double)best * 1e-3);
    }
    for (int k = kinds; k < n_rates; ++k) rates[k] = 0.0;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFreeAsync(out, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

int xpr_valu_ceiling(xpr_stream s, double *lane_instr_per_s)
{
    if (!lane_instr_per_s) return fail(XPR_E_INVALID, "null output");
    double rates[kValuKinds];
    if (int rc = xpr_valu_rates(s, rates, kValuKinds)) return rc;
    *lane_instr_per_s = *std::max_element(rates, rates + kValuKinds);
    return XPR_OK;
}

int xpr_q22_to_float(xpr_stream s, const void *sums, int acc_bits, float *out, uint64_t n)
{
    if (int rc = check_acc(acc_bits)) return rc;
    if (n == 0) return XPR_OK;
    if (!sums || !out) return fail(XPR_E_INVALID, "null buffer");
    if (acc_bits == XPR_ACC_U32)
        hipLaunchKernelGGL(k_q22_to_float<uint32_t>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)s, (const uint32_t *)sums, out, n);
    else if (acc_bits == XPR_ACC_F32)
        hipLaunchKernelGGL(k_q22_to_float<float>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)s, (const float *)sums, out, n);
    else
        hipLaunchKernelGGL(k_q22_to_float<uint64_t>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)s, (const uint64_t *)sums, out, n);
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

int xpr_scale(xpr_stream s, float *x, uint64_t n, float factor)
{
    if (n == 0) return XPR_OK;
    if (!x) return fail(XPR_E_INVALID, "null buffer");
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)s, x, n, factor);
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

// n_views views of n pixels at in + j * in_stride, four pixels per access,
// when the rows allow it (n and in_stride multiples of 4, in 16-byte
// aligned); false otherwise (nothing launched)
static bool preprocess_views4(hipStream_t st, const float *in, uint64_t in_stride, void *out,
                              int out_dtype, uint64_t n, uint32_t n_views, int normalize, int mlog, float mn, float mx)
{
    if (n % 4 || in_stride % 4 || (uintptr_t)in % 16) return false;
    const uint64_t n4 = n / 4;
    const float4 *in4 = (const float4 *)in;
    // x: half the float4s of a view (two per thread), y: the views (each row
    // of the grid strides over them)
    const unsigned gx = grid_for((n4 + 1) / 2);
    const dim3 g(gx, std::min<uint32_t>(n_views, std::max<uint32_t>(1u, 65536u / gx)));
    switch (out_dtype) {
    case XPR_F32:
        hipLaunchKernelGGL(k_preprocess_views4<F32>, g, dim3(256), 0, st, in4, in_stride / 4,
                           (float *)out, n4, n_views, normalize, mlog, mn, mx);
        break;
    case XPR_F16:
        hipLaunchKernelGGL(k_preprocess_views4<F16>, g, dim3(256), 0, st, in4, in_stride / 4,
                           (uint16_t *)out, n4, n_views, normalize, mlog, mn, mx);
        break;
    default:
        hipLaunchKernelGGL(k_preprocess_views4<BF16>, g, dim3(256), 0, st, in4, in_stride / 4,
                           (uint16_t *)out, n4, n_views, normalize, mlog, mn, mx);
        break;
    }
    return true;
}

int xpr_preprocess(xpr_stream s, const float *in, void *out, int out_dtype, uint64_t n,
                   int normalize, int mlog, float mn, float mx)
{
    if (int rc = check_dtype(out_dtype)) return rc;
    if (n == 0) return XPR_OK;
    if (!in || !out) return fail(XPR_E_INVALID, "null buffer");
    hipStream_t st = (hipStream_t)s;
    if (preprocess_views4(st, in, n, out, out_dtype, n, 1, normalize, mlog, mn, mx)) {
        XPR_LAUNCH_CHECK();
        return XPR_OK;
    }
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

int xpr_preprocess_views(xpr_stream s, const float *in, uint64_t in_view_stride, void *out,
                         int out_dtype, uint64_t n, uint32_t n_views, int normalize, int mlog,
                         float mn, float mx)
{
    if (int rc = check_dtype(out_dtype)) return rc;
    if (n == 0 || n_views == 0) return XPR_OK;
    if (!in || !out) return fail(XPR_E_INVALID, "null buffer");
    if (in_view_stride < n) return fail(XPR_E_INVALID, "views overlap: in_view_stride < n");
    hipStream_t st = (hipStream_t)s;
    if (preprocess_views4(st, in, in_view_stride, out, out_dtype, n, n_views, normalize, mlog, mn, mx)) {
        XPR_LAUNCH_CHECK();
        return XPR_OK;
    }
    // x covers a view, y the views (at most 64 rows of the grid; each row
    // strides over the views)
    const unsigned gx = grid_for(n);
    const dim3 g(gx, std::min<uint32_t>(n_views, std::max<uint32_t>(1u, 65536u / gx)));
    switch (out_dtype) {
    case XPR_F32:
        hipLaunchKernelGGL(k_preprocess_views<F32>, g, dim3(256), 0, st, in, in_view_stride,
                           (float *)out, n, n_views, normalize, mlog, mn, mx);
        break;
    case XPR_F16:
        hipLaunchKernelGGL(k_preprocess_views<F16>, g, dim3(256), 0, st, in, in_view_stride,
                           (uint16_t *)out, n, n_views, normalize, mlog, mn, mx);
        break;
    default:
        hipLaunchKernelGGL(k_preprocess_views<BF16>, g, dim3(256), 0, st, in, in_view_stride,
                           (uint16_t *)out, n, n_views, normalize, mlog, mn, mx);
        break;
    }
    XPR_LAUNCH_CHECK();
    return XPR_OK;
}

} // extern "C"
