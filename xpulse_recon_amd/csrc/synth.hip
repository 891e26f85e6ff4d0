// Synthetic projections: FP64 analytic line integrals of an ellipsoid phantom
// (test and benchmark inputs; the reference reads measured TIFFs instead).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/xpr_hip.h"

int xprint_fail(int code, const char *msg);

namespace {

constexpr int kMaxEllipsoids = 512;

__device__ __forceinline__ double chord(const xpr_ellipsoid &e, const double o[3],
                                        const double d[3])
{
    const double q0[3] = {o[0] - e.c[0], o[1] - e.c[1], o[2] - e.c[2]};
    double q[3], k[3];
    for (int r = 0; r < 3; ++r) {
        q[r] = (e.rot[3 * r] * q0[0] + e.rot[3 * r + 1] * q0[1] + e.rot[3 * r + 2] * q0[2]) / e.a[r];
        k[r] = (e.rot[3 * r] * d[0] + e.rot[3 * r + 1] * d[1] + e.rot[3 * r + 2] * d[2]) / e.a[r];
    }
    const double A = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    const double B = 2.0 * (q[0] * k[0] + q[1] * k[1] + q[2] * k[2]);
    const double C = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] - 1.0;
    const double disc = B * B - 4.0 * A * C;
    if (disc <= 0.0) return 0.0;
    return sqrt(disc) / A * sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

__global__ __launch_bounds__(256) void k_line_integrals(
    const double *__restrict__ frames, uint32_t n_views, uint32_t modules, uint32_t width,
    uint32_t height, const xpr_ellipsoid *__restrict__ ell, int ne, double *__restrict__ out)
{
    __shared__ xpr_ellipsoid s_ell[64];
    const uint64_t npix = (uint64_t)width * height;
    const uint64_t total = npix * n_views;
    const uint32_t mw = width / modules;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < total;
         base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        const bool live = i < total;
        double acc = 0.0;
        double o[3] = {0, 0, 0}, d[3] = {0, 0, 0};
        if (live) {
            const uint32_t view = (uint32_t)(i / npix);
            const uint32_t rem = (uint32_t)(i % npix);
            const uint32_t row = rem / width, col = rem % width;
            uint32_t m = col / mw;
            if (m >= modules) m = modules - 1;
            const double *f = frames + 14ull * (view * modules + m);
            const double du = (double)col - f[12], dv = (double)row - f[13];
            for (int r = 0; r < 3; ++r) {
                o[r] = f[r];
                const double hit = f[3 + r] + f[6 + r] * du + f[9 + r] * dv;
                d[r] = hit - o[r];
            }
        }
        for (int e0 = 0; e0 < ne; e0 += 64) {
            const int cnt = min(64, ne - e0);
            __syncthreads();
            for (int t = threadIdx.x; t < cnt * 16; t += blockDim.x)
                ((double *)s_ell)[t] = ((const double *)(ell + e0))[t];
            __syncthreads();
            if (live)
                for (int j = 0; j < cnt; ++j) acc += s_ell[j].mu * chord(s_ell[j], o, d);
        }
        if (live) out[i] = acc;
    }
}

// Sum of the attenuations of the ellipsoids containing each voxel centre
// (voxel (i, l, k) centred at orig + (i, l, k) + 0.5 voxels), float32 result.
__global__ __launch_bounds__(256) void k_phantom(double vx, int64_t vwidth, int64_t vheight,
                                                 double ox, double oy, double oz,
                                                 const xpr_ellipsoid *__restrict__ ell, int ne,
                                                 float *__restrict__ out)
{
    const uint64_t total = (uint64_t)vwidth * vwidth * vheight;
    for (uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; id < total;
         id += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = id % vwidth, k = (id / vwidth) % vwidth, l = id / ((uint64_t)vwidth * vwidth);
        const double p[3] = {ox + ((double)i + 0.5) * vx, oy + ((double)l + 0.5) * vx,
                             oz + ((double)k + 0.5) * vx};
        double s = 0.0;
        for (int j = 0; j < ne; ++j) {
            const xpr_ellipsoid &e = ell[j];
            const double q0[3] = {p[0] - e.c[0], p[1] - e.c[1], p[2] - e.c[2]};
            double r2 = 0.0;
            for (int r = 0; r < 3; ++r) {
                const double b = (e.rot[3 * r] * q0[0] + e.rot[3 * r + 1] * q0[1] +
                                  e.rot[3 * r + 2] * q0[2]) / e.a[r];
                r2 = r2 + b * b;
            }
            if (r2 <= 1.0) s += e.mu;
        }
        out[id] = (float)s;
    }
}

} // namespace

extern "C" int xpr_synth_phantom(xpr_stream s, double vx, int64_t vwidth, int64_t vheight,
                                 const double orig[3], const xpr_ellipsoid *ellipsoids,
                                 int n_ellipsoids, float *out)
{
    if (!out || !orig || (n_ellipsoids > 0 && !ellipsoids))
        return xprint_fail(XPR_E_INVALID, "null buffer");
    if (vwidth <= 0 || vheight <= 0 || n_ellipsoids < 0 || n_ellipsoids > kMaxEllipsoids)
        return xprint_fail(XPR_E_INVALID, "synth: bad grid or ellipsoid count");
    const uint64_t total = (uint64_t)vwidth * vwidth * vheight;
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_phantom, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s, vx, vwidth,
                       vheight, orig[0], orig[1], orig[2], ellipsoids, n_ellipsoids, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return xprint_fail(XPR_E_HIP, (std::string("synth launch: ") + hipGetErrorString(e)).c_str());
    return XPR_OK;
}

extern "C" int xpr_synth_line_integrals(xpr_stream s, const double *frames, uint32_t n_views,
                                        uint32_t modules, uint32_t det_width,
                                        uint32_t det_height, const xpr_ellipsoid *ellipsoids,
                                        int n_ellipsoids, double *out)
{
    if (!frames || !out || (n_ellipsoids > 0 && !ellipsoids))
        return xprint_fail(XPR_E_INVALID, "null buffer");
    if (modules == 0 || n_ellipsoids < 0 || n_ellipsoids > kMaxEllipsoids)
        return xprint_fail(XPR_E_INVALID, "synth: bad module or ellipsoid count");
    const uint64_t total = (uint64_t)det_width * det_height * n_views;
    if (total == 0) return XPR_OK;
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_line_integrals, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s,
                       frames, n_views, modules, det_width, det_height, ellipsoids, n_ellipsoids,
                       out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return xprint_fail(XPR_E_HIP, (std::string("synth launch: ") + hipGetErrorString(e)).c_str());
    return XPR_OK;
}
