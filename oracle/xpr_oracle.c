/*
 * CPU oracle for the iterative cone-beam (XPRecon) hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, smoke() and bench.py's
 * cpu_baseline leg load this library; the product never does.
 *
 * Canonical arithmetic (the reference builds its kernels with
 * -cl-fast-relaxed-math -cl-mad-enable, gpu/kernel.hpp:26-31, so any of the
 * following is a legal lowering of its kernel text; this file fixes one):
 *   position  p.a = fmaf((float)idx_a, voxel_a, origin_a)
 *   dot(c, p) = fmaf(c.x, p.x, fmaf(c.y, p.y, fmaf(c.z, p.z, c.w)))   (p.w = 1)
 *   (u, v)    = (X * r, Y * r - line_offset),  r = 1.0f / W  (IEEE)
 *   pixel     = roundf (half away from zero)
 * Compile with -ffp-contract=off: every fused operation is an explicit fmaf.
 */
#include "xpr_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define Q22_ONE 4194304.0f          /* FIXED_FRAC_ONE, constants.cl:19 */
#define Q22_EPS (1.0f / 4194304.0f) /* EPSILON = FIXED_TO_FLOAT(1), constants.cl:23 */

/* ------------------------------------------------------------------------ */
/* Projection of one voxel position by one ProjData (KernelForward.cl:30-35,
 * utils.cl:26).  Returns 1 when in bounds and writes the detector index.   */
/* ------------------------------------------------------------------------ */
static inline int project_f32(const float *pd, float px, float py, float pz,
                              const orc_geom *g, uint32_t view, uint64_t *index)
{
    /* pd[0..3] = glm column 0, pd[4..7] = column 1, pd[12..15] = column 3;
     * the kernel reads them as mvp[0], mvp[1], mvp[3] (utils.cl:26). */
    const float X = fmaf(pd[0], px, fmaf(pd[1], py, fmaf(pd[2], pz, pd[3])));
    const float Y = fmaf(pd[4], px, fmaf(pd[5], py, fmaf(pd[6], pz, pd[7])));
    const float W = fmaf(pd[12], px, fmaf(pd[13], py, fmaf(pd[14], pz, pd[15])));
    const float r = 1.0f / W;
    const float u = X * r;
    const float v = Y * r - (float)g->line_offset;       /* KernelForward.cl:32 */
    const float cu = roundf(u), cv = roundf(v);           /* :34 */
    /* vp.even = (vpl, vpb), vp.odd = (vpr, vpt), half-open (:35).  Rows are
     * also bounded by the rows held (line_rows) so a cropped band never
     * spills into the next view. */
    if (!(cu >= pd[16] && cu < pd[17] && cv >= pd[18] && cv < pd[19] &&
          cv >= 0.0f && cv < (float)g->line_rows))
        return 0;
    /* IMAGE_OFFSET_2D (utils.cl:47) */
    *index = (uint64_t)view * ((uint64_t)g->det_width * (uint64_t)g->line_rows) +
             (uint64_t)(int64_t)cv * g->det_width + (uint64_t)(int64_t)cu;
    return 1;
}

static inline uint64_t slab_voxels(const orc_geom *g, int64_t *nx, int64_t *nl,
                                   int64_t *nk)
{
    *nx = g->vmax[0] - g->vmin[0];
    *nl = g->vmax[1] - g->vmin[1];
    *nk = g->vmax[2] - g->vmin[2];
    if (*nx <= 0 || *nl <= 0 || *nk <= 0) return 0;
    return (uint64_t)(*nx) * (uint64_t)(*nl) * (uint64_t)(*nk);
}

/* FLOAT_TO_FIXED (constants.cl:22): truncating conversion of v*2^22.  The
 * GPU conversion the reference runs on saturates, so values outside the
 * range clamp instead of being undefined. */
static inline uint64_t to_q22(float v, int acc_bits)
{
    const float s = v * Q22_ONE;
    if (!(s > 0.0f)) return 0;
    if (acc_bits == 32) {
        if (s >= 4294967296.0f) return 0xFFFFFFFFull;
        return (uint64_t)(uint32_t)s;
    }
    if (s >= 18446744073709551616.0f) return 0xFFFFFFFFFFFFFFFFull;
    return (uint64_t)s;
}

/* Forward (KernelForward.cl:22-37). */
uint64_t orc_forward_f32(const float *vol, void *sums, int acc_bits,
                         const float *projdata, const orc_geom *g, int threads)
{
    int64_t nx, nl, nk;
    const uint64_t nvox = slab_voxels(g, &nx, &nl, &nk);
    const uint32_t nm = g->n_views * g->modules;
    uint64_t hits = 0;
    if (threads <= 0) threads = 1;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : hits) num_threads(threads)
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;   /* to3D: z = layer, y = depth */
        const float py = fmaf((float)(l + g->vmin[1]), g->voxel[1], g->origin[1]);
        const float pz = fmaf((float)(k + g->vmin[2]), g->voxel[2], g->origin[2]);
        for (int64_t i = 0; i < nx; ++i) {
            const float px = fmaf((float)(i + g->vmin[0]), g->voxel[0], g->origin[0]);
            const uint64_t id = (uint64_t)row * nx + i;
            const uint64_t q = to_q22(vol[id], acc_bits);
            for (uint32_t m = 0; m < nm; ++m) {
                uint64_t idx;
                if (!project_f32(projdata + 20 * (size_t)m, px, py, pz, g,
                                 m / g->modules, &idx))
                    continue;
                ++hits;
                if (acc_bits == 32) {
                    uint32_t *s = (uint32_t *)sums;
#pragma omp atomic
                    s[idx] += (uint32_t)q;
                } else {
                    uint64_t *s = (uint64_t *)sums;
#pragma omp atomic
                    s[idx] += q;
                }
            }
        }
    }
    (void)nvox;
    return hits;
}

/* Error (KernelError.cl:14-23). */
void orc_error_f32(const float *images, const void *sums, int acc_bits,
                   float *err, uint64_t n, float weight)
{
    for (uint64_t id = 0; id < n; ++id) {
        const float ref = images[id];
        float got;
        if (acc_bits == 32)
            got = (float)((const uint32_t *)sums)[id] / Q22_ONE;   /* FIXED_TO_FLOAT */
        else
            got = (float)((const uint64_t *)sums)[id] / Q22_ONE;
        err[id] = logf(fmaxf(ref / fmaxf(got, Q22_EPS), Q22_EPS)) * weight;
    }
}

/* Backward + update (KernelBackward.cl:20-40). */
uint64_t orc_backward_f32(float *vol, const float *err, const float *projdata,
                          const orc_geom *g, int threads)
{
    int64_t nx, nl, nk;
    slab_voxels(g, &nx, &nl, &nk);
    const uint32_t nm = g->n_views * g->modules;
    uint64_t empty = 0;
    if (threads <= 0) threads = 1;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : empty) num_threads(threads)
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const float py = fmaf((float)(l + g->vmin[1]), g->voxel[1], g->origin[1]);
        const float pz = fmaf((float)(k + g->vmin[2]), g->voxel[2], g->origin[2]);
        for (int64_t i = 0; i < nx; ++i) {
            const float px = fmaf((float)(i + g->vmin[0]), g->voxel[0], g->origin[0]);
            const uint64_t id = (uint64_t)row * nx + i;
            float sum = 0.0f;
            uint32_t n = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                uint64_t idx;
                if (!project_f32(projdata + 20 * (size_t)m, px, py, pz, g,
                                 m / g->modules, &idx))
                    continue;
                sum = sum + err[idx];
                ++n;
            }
            if (n == 0) {           /* reference: exp(0/0); kept unchanged here */
                ++empty;
                continue;
            }
            vol[id] = vol[id] * expf(sum / (float)n);
        }
    }
    return empty;
}

void orc_hits_f32(uint32_t *counts, const float *projdata, const orc_geom *g)
{
    int64_t nx, nl, nk;
    slab_voxels(g, &nx, &nl, &nk);
    const uint32_t nm = g->n_views * g->modules;
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const float py = fmaf((float)(l + g->vmin[1]), g->voxel[1], g->origin[1]);
        const float pz = fmaf((float)(k + g->vmin[2]), g->voxel[2], g->origin[2]);
        for (int64_t i = 0; i < nx; ++i) {
            const float px = fmaf((float)(i + g->vmin[0]), g->voxel[0], g->origin[0]);
            uint32_t n = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                uint64_t idx;
                n += (uint32_t)project_f32(projdata + 20 * (size_t)m, px, py, pz,
                                           g, m / g->modules, &idx);
            }
            counts[(uint64_t)row * nx + i] = n;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 model of the same NN operators (no fixed point).                    */
/* ------------------------------------------------------------------------ */
static inline int project_f64(const double *pd, double px, double py, double pz,
                              const orc_geom64 *g, uint32_t view, uint64_t *index)
{
    const double X = pd[0] * px + pd[1] * py + pd[2] * pz + pd[3];
    const double Y = pd[4] * px + pd[5] * py + pd[6] * pz + pd[7];
    const double W = pd[12] * px + pd[13] * py + pd[14] * pz + pd[15];
    const double u = X / W;
    const double v = Y / W - (double)g->line_offset;
    const double cu = round(u), cv = round(v);
    if (!(cu >= pd[16] && cu < pd[17] && cv >= pd[18] && cv < pd[19] &&
          cv >= 0.0 && cv < (double)g->line_rows))
        return 0;
    *index = (uint64_t)view * ((uint64_t)g->det_width * (uint64_t)g->line_rows) +
             (uint64_t)(int64_t)cv * g->det_width + (uint64_t)(int64_t)cu;
    return 1;
}

void orc_forward_f64(const double *vol, double *sums, const double *projdata,
                     const orc_geom64 *g)
{
    const int64_t nx = g->vmax[0] - g->vmin[0], nl = g->vmax[1] - g->vmin[1],
                  nk = g->vmax[2] - g->vmin[2];
    const uint32_t nm = g->n_views * g->modules;
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const double py = (double)(l + g->vmin[1]) * g->voxel[1] + g->origin[1];
        const double pz = (double)(k + g->vmin[2]) * g->voxel[2] + g->origin[2];
        for (int64_t i = 0; i < nx; ++i) {
            const double px = (double)(i + g->vmin[0]) * g->voxel[0] + g->origin[0];
            const double val = vol[(uint64_t)row * nx + i];
            for (uint32_t m = 0; m < nm; ++m) {
                uint64_t idx;
                if (project_f64(projdata + 20 * (size_t)m, px, py, pz, g,
                                m / g->modules, &idx))
                    sums[idx] += val;
            }
        }
    }
}

void orc_error_f64(const double *images, const double *sums, double *err,
                   uint64_t n, double weight)
{
    const double eps = 1.0 / 4194304.0;
    for (uint64_t id = 0; id < n; ++id)
        err[id] = log(fmax(images[id] / fmax(sums[id], eps), eps)) * weight;
}

uint64_t orc_backward_f64(double *vol, const double *err, const double *projdata,
                          const orc_geom64 *g)
{
    const int64_t nx = g->vmax[0] - g->vmin[0], nl = g->vmax[1] - g->vmin[1],
                  nk = g->vmax[2] - g->vmin[2];
    const uint32_t nm = g->n_views * g->modules;
    uint64_t empty = 0;
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const double py = (double)(l + g->vmin[1]) * g->voxel[1] + g->origin[1];
        const double pz = (double)(k + g->vmin[2]) * g->voxel[2] + g->origin[2];
        for (int64_t i = 0; i < nx; ++i) {
            const double px = (double)(i + g->vmin[0]) * g->voxel[0] + g->origin[0];
            double sum = 0.0;
            uint32_t n = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                uint64_t idx;
                if (project_f64(projdata + 20 * (size_t)m, px, py, pz, g,
                                m / g->modules, &idx)) {
                    sum += err[idx];
                    ++n;
                }
            }
            if (n == 0) { ++empty; continue; }
            vol[(uint64_t)row * nx + i] *= exp(sum / (double)n);
        }
    }
    return empty;
}

/* ------------------------------------------------------------------------ */
/* FP64 values on the canonical float32 pixel choice: the voxel-views and     */
/* pixels of the FP32 kernels (project_f32 on float32 ProjData and            */
/* positions), with the sums, the error and the update in double.  Against   */
/* it only value arithmetic differs from the FP32 path.  q22 != 0 truncates  */
/* each forward contribution to a whole multiple of 2^-22 first, as          */
/* FLOAT_TO_FIXED does (constants.cl:22), so that the Q.22 quantisation is   */
/* modelled and only float rounding remains between the two.                 */
/* ------------------------------------------------------------------------ */
void orc_forward_f64nn(const double *vol, double *sums, const float *projdata,
                       const orc_geom *g, int threads, int q22)
{
    int64_t nx, nl, nk;
    slab_voxels(g, &nx, &nl, &nk);
    const uint32_t nm = g->n_views * g->modules;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const float py = fmaf((float)(l + g->vmin[1]), g->voxel[1], g->origin[1]);
        const float pz = fmaf((float)(k + g->vmin[2]), g->voxel[2], g->origin[2]);
        for (int64_t i = 0; i < nx; ++i) {
            const float px = fmaf((float)(i + g->vmin[0]), g->voxel[0], g->origin[0]);
            double val = vol[(uint64_t)row * nx + i];
            if (q22) val = val > 0.0 ? floor(val * 4194304.0) / 4194304.0 : 0.0;
            for (uint32_t m = 0; m < nm; ++m) {
                uint64_t idx;
                if (!project_f32(projdata + 20 * (size_t)m, px, py, pz, g, m / g->modules, &idx))
                    continue;
#pragma omp atomic
                sums[idx] += val;
            }
        }
    }
}

uint64_t orc_backward_f64nn(double *vol, const double *err, const float *projdata,
                            const orc_geom *g, int threads)
{
    int64_t nx, nl, nk;
    slab_voxels(g, &nx, &nl, &nk);
    const uint32_t nm = g->n_views * g->modules;
    uint64_t empty = 0;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : empty) num_threads(threads)
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const float py = fmaf((float)(l + g->vmin[1]), g->voxel[1], g->origin[1]);
        const float pz = fmaf((float)(k + g->vmin[2]), g->voxel[2], g->origin[2]);
        for (int64_t i = 0; i < nx; ++i) {
            const float px = fmaf((float)(i + g->vmin[0]), g->voxel[0], g->origin[0]);
            double sum = 0.0;
            uint32_t n = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                uint64_t idx;
                if (project_f32(projdata + 20 * (size_t)m, px, py, pz, g, m / g->modules, &idx)) {
                    sum += err[idx];
                    ++n;
                }
            }
            if (n == 0) { ++empty; continue; }
            vol[(uint64_t)row * nx + i] *= exp(sum / (double)n);
        }
    }
    return empty;
}

/* ------------------------------------------------------------------------ */
/* Opt-in bilinear backprojector.  The reference's built kernels sample the  */
/* nearest pixel only; KernelForward.cl:38-53 holds an unused bilinear       */
/* branch (INTERP_BILIN, never defined) whose intent this follows.  The      */
/* voxel-views counted are those the nearest-neighbour test accepts          */
/* (KernelForward.cl:34-35, with the viewport clamped to the detector and    */
/* the rows held); each samples the error image at the continuous (u, v) of  */
/* the canonical projection, pixel k centred at coordinate k, from columns   */
/* floor u, floor u + 1 and rows floor v, floor v + 1, each clamped to the   */
/* accepted pixels [cl, cr - 1] x [rl, rr - 1] (clamp to edge: an edge       */
/* pixel's value holds over the half pixel beyond its centre), with          */
/*   cl = ceil(max(vpl, 0)), cr = ceil(min(vpr, det_width)),                 */
/*   rl = ceil(max(vpb, 0)), rr = ceil(min(vpt, line_rows)).                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    float u, v;            /* continuous coordinate                         */
    float fu, fv;          /* floor u, floor v                              */
    uint64_t c0, c1, r0, r1;  /* clamped tap columns and rows               */
} bil_taps;

static inline int project_bilinear_f32(const float *pd, float px, float py, float pz,
                                       const orc_geom *g, bil_taps *t)
{
    const float X = fmaf(pd[0], px, fmaf(pd[1], py, fmaf(pd[2], pz, pd[3])));
    const float Y = fmaf(pd[4], px, fmaf(pd[5], py, fmaf(pd[6], pz, pd[7])));
    const float W = fmaf(pd[12], px, fmaf(pd[13], py, fmaf(pd[14], pz, pd[15])));
    const float r = 1.0f / W;
    t->u = X * r;
    t->v = Y * r - (float)g->line_offset;
    const float cl = ceilf(fmaxf(pd[16], 0.0f)), cr = ceilf(fminf(pd[17], (float)g->det_width));
    const float rl = ceilf(fmaxf(pd[18], 0.0f)), rr = ceilf(fminf(pd[19], (float)g->line_rows));
    const float cu = roundf(t->u), cv = roundf(t->v);
    if (!(cu >= cl && cu < cr && cv >= rl && cv < rr)) return 0;
    t->fu = floorf(t->u);
    t->fv = floorf(t->v);
    t->c0 = (uint64_t)fmaxf(t->fu, cl);
    t->c1 = (uint64_t)fminf(t->fu + 1.0f, cr - 1.0f);
    t->r0 = (uint64_t)fmaxf(t->fv, rl);
    t->r1 = (uint64_t)fminf(t->fv + 1.0f, rr - 1.0f);
    return 1;
}

/* FP32 canonical form: weights fx = u - floor u, fy = v - floor v and three
 * lerps over the clamped taps,
 *   top = fmaf(fx, E[r0][c1] - E[r0][c0], E[r0][c0]),
 *   bot = fmaf(fx, E[r1][c1] - E[r1][c0], E[r1][c0]),
 *   sample = fmaf(fy, bot - top, top),
 * summed in view order; update v * expf(sum / n) as Backward. */
uint64_t orc_backward_bilinear_f32(float *vol, const float *err, const float *projdata,
                                   const orc_geom *g, int threads)
{
    int64_t nx, nl, nk;
    slab_voxels(g, &nx, &nl, &nk);
    const uint32_t nm = g->n_views * g->modules;
    const uint64_t W = g->det_width, view_elems = W * (uint64_t)g->line_rows;
    uint64_t empty = 0;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : empty) num_threads(threads)
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const float py = fmaf((float)(l + g->vmin[1]), g->voxel[1], g->origin[1]);
        const float pz = fmaf((float)(k + g->vmin[2]), g->voxel[2], g->origin[2]);
        for (int64_t i = 0; i < nx; ++i) {
            const float px = fmaf((float)(i + g->vmin[0]), g->voxel[0], g->origin[0]);
            float sum = 0.0f;
            uint32_t n = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                bil_taps t;
                if (!project_bilinear_f32(projdata + 20 * (size_t)m, px, py, pz, g, &t)) continue;
                const float *E = err + (uint64_t)(m / g->modules) * view_elems;
                const float fx = t.u - t.fu, fy = t.v - t.fv;
                const float e00 = E[t.r0 * W + t.c0], e01 = E[t.r0 * W + t.c1];
                const float e10 = E[t.r1 * W + t.c0], e11 = E[t.r1 * W + t.c1];
                const float top = fmaf(fx, e01 - e00, e00);
                const float bot = fmaf(fx, e11 - e10, e10);
                sum = sum + fmaf(fy, bot - top, top);
                ++n;
            }
            if (n == 0) { ++empty; continue; }
            const uint64_t id = (uint64_t)row * nx + i;
            vol[id] = vol[id] * expf(sum / (float)n);
        }
    }
    return empty;
}

/* FP64 model of the same operator on the float32 accept set: which
 * voxel-views count and their clamped taps come from project_bilinear_f32
 * (as on the device); the weights use u and v evaluated in double from the
 * same float32 positions and ProjData, fx = u - floor u (the float floor),
 * and the sample
 *   (1-fy) ((1-fx) E[r0][c0] + fx E[r0][c1]) + fy ((1-fx) E[r1][c0] + fx E[r1][c1]),
 * the sum and the update are double. */
uint64_t orc_backward_bilinear_f64nn(double *vol, const double *err, const float *projdata,
                                     const orc_geom *g, int threads)
{
    int64_t nx, nl, nk;
    slab_voxels(g, &nx, &nl, &nk);
    const uint32_t nm = g->n_views * g->modules;
    const uint64_t W = g->det_width, view_elems = W * (uint64_t)g->line_rows;
    uint64_t empty = 0;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : empty) num_threads(threads)
    for (int64_t row = 0; row < nl * nk; ++row) {
        const int64_t l = row / nk, k = row % nk;
        const float py = fmaf((float)(l + g->vmin[1]), g->voxel[1], g->origin[1]);
        const float pz = fmaf((float)(k + g->vmin[2]), g->voxel[2], g->origin[2]);
        for (int64_t i = 0; i < nx; ++i) {
            const float px = fmaf((float)(i + g->vmin[0]), g->voxel[0], g->origin[0]);
            double sum = 0.0;
            uint32_t n = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                const float *pd = projdata + 20 * (size_t)m;
                bil_taps t;
                if (!project_bilinear_f32(pd, px, py, pz, g, &t)) continue;
                const double X = (double)pd[0] * px + (double)pd[1] * py + (double)pd[2] * pz + pd[3];
                const double Y = (double)pd[4] * px + (double)pd[5] * py + (double)pd[6] * pz + pd[7];
                const double Wd = (double)pd[12] * px + (double)pd[13] * py + (double)pd[14] * pz + pd[15];
                const double fx = X / Wd - t.fu, fy = Y / Wd - g->line_offset - t.fv;
                const double *E = err + (uint64_t)(m / g->modules) * view_elems;
                sum += (1.0 - fy) * ((1.0 - fx) * E[t.r0 * W + t.c0] + fx * E[t.r0 * W + t.c1]) +
                       fy * ((1.0 - fx) * E[t.r1 * W + t.c0] + fx * E[t.r1 * W + t.c1]);
                ++n;
            }
            if (n == 0) { ++empty; continue; }
            vol[(uint64_t)row * nx + i] *= exp(sum / (double)n);
        }
    }
    return empty;
}

/* ------------------------------------------------------------------------ */
/* Preprocessing (data_loader.hpp:55-86).                                   */
/* ------------------------------------------------------------------------ */
void orc_minmax_normal(const float *data, uint64_t n, float *mn, float *mx)
{
    float lo = *mn, hi = *mx;
    for (uint64_t i = 0; i < n; ++i) {
        const float v = data[i];
        if (!isnormal(v)) continue;          /* remove_if(!isnormal), :58-61 */
        if (v < lo) lo = v;
        if (v > hi) hi = v;
    }
    *mn = lo;
    *mx = hi;
}

void orc_preprocess(float *data, uint64_t n, int normalize, int mlog, float mn,
                    float mx)
{
    const float hi_clamp = 1.0f - Q22_EPS;
    for (uint64_t i = 0; i < n; ++i) {
        float v = data[i];
        if (normalize)                                         /* :80-82 */
            v = isnormal(v) ? (v - mn) / (mx - mn) : (mx + mn) / 2.0f;
        if (mlog) {                                            /* :84-86 */
            float t = (hi_clamp < v) ? hi_clamp : v;           /* std::min  */
            t = (t < Q22_EPS) ? Q22_EPS : t;                   /* std::max  */
            v = -logf(t);
        }
        data[i] = v;
    }
}

/* ------------------------------------------------------------------------ */
/* Geometry (geometry.hpp).  Two arithmetics share one body through a       */
/* macro-free pair of small vector types.                                   */
/* ------------------------------------------------------------------------ */
typedef struct { double x, y, z; } d3;
typedef struct { float x, y, z; } f3;

static d3 d3_sub(d3 a, d3 b) { d3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static d3 d3_add(d3 a, d3 b) { d3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static d3 d3_mul(d3 a, double s) { d3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static double d3_dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 d3_cross(d3 a, d3 b)
{
    d3 r = {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
    return r;
}
static d3 d3_norm(d3 a) { return d3_mul(a, 1.0 / sqrt(d3_dot(a, a))); }
static d3 d3_rot_y(d3 v, double rad)   /* rotation about +Y (glm rotateY) */
{
    const double c = cos(rad), s = sin(rad);
    d3 r = {v.x * c + v.z * s, v.y, -v.x * s + v.z * c};
    return r;
}

static f3 f3_sub(f3 a, f3 b) { f3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static f3 f3_add(f3 a, f3 b) { f3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static f3 f3_mul(f3 a, float s) { f3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static float f3_dot(f3 a, f3 b)
{
    /* product vector, then x + y + z left to right */
    const float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
    return (tx + ty) + tz;
}
static f3 f3_cross(f3 a, f3 b)
{
    f3 r = {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
    return r;
}
static f3 f3_norm(f3 a) { return f3_mul(a, 1.0f / sqrtf(f3_dot(a, a))); }
static float f3_len(f3 a) { return sqrtf(f3_dot(a, a)); }
static f3 f3_rot_x(f3 v, float rad)
{
    const float c = cosf(rad), s = sinf(rad);
    f3 r = {v.x, v.y * c - v.z * s, v.y * s + v.z * c};
    return r;
}
static f3 f3_rot_y(f3 v, float rad)
{
    const float c = cosf(rad), s = sinf(rad);
    f3 r = {v.x * c + v.z * s, v.y, -v.x * s + v.z * c};
    return r;
}
static f3 f3_rot_z(f3 v, float rad)
{
    const float c = cosf(rad), s = sinf(rad);
    f3 r = {v.x * c - v.y * s, v.x * s + v.y * c, v.z};
    return r;
}
#define F_RADIANS 0.01745329251994329576923690768489f
#define F_DEGREES 57.295779513082320876798154814105f

/* Geometry::rotate(vec, origin, 0, y, 0) (geometry.hpp:88-95). */
static f3 f3_rotate_about(f3 vec, f3 origin, float deg_y)
{
    f3 v = f3_sub(vec, origin);
    v = f3_rot_x(v, 0.0f * F_RADIANS);
    v = f3_rot_y(v, deg_y * F_RADIANS);
    v = f3_rot_z(v, 0.0f * F_RADIANS);
    return f3_add(v, origin);
}

/* 4x4 matrices as glm stores them: m[col][row]. */
typedef struct { double m[4][4]; } d44;
typedef struct { float m[4][4]; } f44;

static d44 d44_mul(const d44 *a, const d44 *b)
{
    d44 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r.m[c][row] = a->m[0][row] * b->m[c][0] + a->m[1][row] * b->m[c][1] +
                          a->m[2][row] * b->m[c][2] + a->m[3][row] * b->m[c][3];
    return r;
}
static f44 f44_mul(const f44 *a, const f44 *b)
{
    /* column c of the product = a[0]*b[c][0] + a[1]*b[c][1] + ... left to right */
    f44 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            float acc = a->m[0][row] * b->m[c][0];
            acc = acc + a->m[1][row] * b->m[c][1];
            acc = acc + a->m[2][row] * b->m[c][2];
            acc = acc + a->m[3][row] * b->m[c][3];
            r.m[c][row] = acc;
        }
    return r;
}
/* glm::mat4(16 scalars) fills column by column. */
static void d44_cols(d44 *r, const double v[16])
{
    for (int i = 0; i < 16; ++i) r->m[i / 4][i % 4] = v[i];
}
static void f44_cols(f44 *r, const float v[16])
{
    for (int i = 0; i < 16; ++i) r->m[i / 4][i % 4] = v[i];
}

/* computeGeneralizedProjection (geometry.hpp:182-234), FP64. */
static d44 gen_proj_d(d3 pe, d3 pa, d3 pb, d3 pc, double vpl, double vpr,
                      double vpb, double vpt, double sd)
{
    const d3 vr = d3_norm(d3_sub(pb, pa));
    const d3 vu = d3_norm(d3_sub(pc, pa));
    const d3 vn = d3_norm(d3_cross(vu, vr));
    const d3 va = d3_sub(pa, pe), vb = d3_sub(pb, pe), vc = d3_sub(pc, pe);
    const double d = -d3_dot(va, vn), n = 1.0, f = sd * 2.0;
    const double l = d3_dot(vr, va) * (n / d), r = d3_dot(vr, vb) * (n / d);
    const double b = d3_dot(vu, va) * (n / d), t = d3_dot(vu, vc) * (n / d);
    const double P[16] = {(2 * n) / (r - l), 0, (r + l) / (r - l), 0,
                          0, (2 * n) / (t - b), (t + b) / (t - b), 0,
                          0, 0, -(f + n) / (f - n), -(2 * f * n) / (f - n),
                          0, 0, -1, 0};
    const double Mt[16] = {vr.x, vr.y, vr.z, 0, vu.x, vu.y, vu.z, 0,
                           vn.x, vn.y, vn.z, 0, 0, 0, 0, 1};
    const double T[16] = {1, 0, 0, -pe.x, 0, 1, 0, -pe.y, 0, 0, 1, -pe.z, 0, 0, 0, 1};
    const double S[16] = {(vpr - vpl) * 0.5, 0, 0, (vpr + vpl) * 0.5,
                          0, (vpt - vpb) * 0.5, 0, (vpt + vpb) * 0.5,
                          0, 0, 0.5, 0.5, 0, 0, 0, 1};
    d44 mP, mMt, mT, mS;
    d44_cols(&mP, P); d44_cols(&mMt, Mt); d44_cols(&mT, T); d44_cols(&mS, S);
    d44 a = d44_mul(&mT, &mMt);
    a = d44_mul(&a, &mP);
    return d44_mul(&a, &mS);
}

/* computeGeneralizedProjection (geometry.hpp:182-234), float32 as glm. */
static f44 gen_proj_f(f3 pe, f3 pa, f3 pb, f3 pc, float vpl, float vpr,
                      float vpb, float vpt, float sd)
{
    const f3 vr = f3_norm(f3_sub(pb, pa));
    const f3 vu = f3_norm(f3_sub(pc, pa));
    const f3 vn = f3_norm(f3_cross(vu, vr));
    const f3 va = f3_sub(pa, pe), vb = f3_sub(pb, pe), vc = f3_sub(pc, pe);
    const float d = -f3_dot(va, vn), n = 1.0f, f = sd * 2.0f;
    const float l = f3_dot(vr, va) * (n / d), r = f3_dot(vr, vb) * (n / d);
    const float b = f3_dot(vu, va) * (n / d), t = f3_dot(vu, vc) * (n / d);
    const float P[16] = {(2 * n) / (r - l), 0.0f, (r + l) / (r - l), 0.0f,
                         0.0f, (2 * n) / (t - b), (t + b) / (t - b), 0.0f,
                         0.0f, 0.0f, -(f + n) / (f - n), -(2 * f * n) / (f - n),
                         0.0f, 0.0f, -1.0f, 0.0f};
    const float Mt[16] = {vr.x, vr.y, vr.z, 0.0f, vu.x, vu.y, vu.z, 0.0f,
                          vn.x, vn.y, vn.z, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    const float T[16] = {1.0f, 0.0f, 0.0f, -pe.x, 0.0f, 1.0f, 0.0f, -pe.y,
                         0.0f, 0.0f, 1.0f, -pe.z, 0.0f, 0.0f, 0.0f, 1.0f};
    const float S[16] = {(vpr - vpl) * 0.5f, 0.0f, 0.0f, (vpr + vpl) * 0.5f,
                         0.0f, (vpt - vpb) * 0.5f, 0.0f, (vpt + vpb) * 0.5f,
                         0.0f, 0.0f, 0.5f, 0.5f, 0.0f, 0.0f, 0.0f, 1.0f};
    f44 mP, mMt, mT, mS;
    f44_cols(&mP, P); f44_cols(&mMt, Mt); f44_cols(&mT, T); f44_cols(&mS, S);
    f44 a = f44_mul(&mT, &mMt);
    a = f44_mul(&a, &mP);
    return f44_mul(&a, &mS);
}

/* Rotation matrix (row-major) by angle t about the unit axis k:
 * cos t I + sin t [k]x + (1 - cos t) k k^T. */
static void axis_matrix(double M[9], d3 k, double t)
{
    const double c = cos(t), s = sin(t), q = 1.0 - c;
    M[0] = c + q * k.x * k.x;       M[1] = q * k.x * k.y - s * k.z; M[2] = q * k.x * k.z + s * k.y;
    M[3] = q * k.y * k.x + s * k.z; M[4] = c + q * k.y * k.y;       M[5] = q * k.y * k.z - s * k.x;
    M[6] = q * k.z * k.x - s * k.y; M[7] = q * k.z * k.y + s * k.x; M[8] = c + q * k.z * k.z;
}
static void mat3_mul(double R[9], const double A[9], const double B[9])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            R[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
static d3 mat3_apply(const double M[9], d3 v)
{
    d3 r = {M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z,
            M[6] * v.x + M[7] * v.y + M[8] * v.z};
    return r;
}

static double angle_deg_d(const orc_geom_params *p, int64_t i)
{
    if (p->angle_list) return p->angle_list[i];
    return (double)i * (p->angle / (double)p->projections);     /* :41-44 */
}

/* Source/detector frame in FP64 (geometry.hpp:58-68, 244-292). */
typedef struct { d3 pe, pa, pb, pc; double vpl, vpr, vpb, vpt; } d_frame;

static void frame_d(const orc_geom_params *p, int64_t a, int64_t mod, d_frame *fr,
                    orc_grid *grid)
{
    const double vx = p->px * (p->so / p->sd);
    const double rx = p->rx_px * p->px;                       /* parameters.hpp:39 */
    d3 src = {0, 0, 0}, at = {0, 0, p->sd};
    d3 axis = d3_add(at, (d3){-p->sx + rx, -p->sy, 0});
    d3 oc = d3_mul(d3_norm(d3_sub(axis, src)), p->so);
    src = d3_sub(src, oc);
    at = d3_sub(at, oc);
    if (grid) {
        grid->vx = vx;
        grid->vwidth = p->dwidth;
        grid->vheight = p->dheight + (int64_t)fabs(p->heli_offset / vx);
        grid->orig[0] = -0.5 * (double)grid->vwidth * vx;
        grid->orig[1] = -0.5 * (double)grid->vheight * vx;
        grid->orig[2] = -0.5 * (double)grid->vwidth * vx;
    }
    const double heli = (p->heli_offset / (double)p->projections) * (double)a -
                        p->heli_offset * 0.5;
    const double rad = angle_deg_d(p, a) * (M_PI / 180.0);
    d3 pos = d3_add(d3_rot_y(src, rad), (d3){0, heli, 0});
    d3 tgt = d3_add(d3_rot_y(at, rad), (d3){0, heli, 0});
    /* detectorFromSource (:244-265) */
    const d3 vdir = {0, 1, 0};
    const d3 u = d3_norm(d3_cross(vdir, d3_sub(tgt, pos)));
    const int64_t width = p->dwidth / p->module_number;
    const double height = (double)p->dheight;
    const d3 det = d3_sub(d3_sub(tgt, d3_mul(u, p->sx)), d3_mul(vdir, p->sy));
    d3 ul = d3_sub(d3_sub(det, d3_mul(u, 0.5 * (double)width * p->px)),
                   d3_mul(vdir, 0.5 * height * p->px));
    d3 ur = d3_add(ul, d3_mul(u, (double)width * p->px));
    d3 bl = d3_add(ul, d3_mul(vdir, height * p->px));
    /* module rotation about the detector centre (the build's extension of the
     * formula commented out at geometry.hpp:260-262): the corners' offsets
     * from the centre turned by yaw about u, then pitch about v, then roll
     * about the beam axis n, all fixed axes: R = R_n(roll) R_v(pitch) R_u(yaw) */
    if (p->roll != 0.0 || p->pitch != 0.0 || p->yaw != 0.0) {
        const d3 nb = d3_norm(d3_sub(tgt, pos));
        double Ru[9], Rv[9], Rn[9], T1[9], R[9];
        axis_matrix(Ru, u, p->yaw * (M_PI / 180.0));
        axis_matrix(Rv, vdir, p->pitch * (M_PI / 180.0));
        axis_matrix(Rn, nb, p->roll * (M_PI / 180.0));
        mat3_mul(T1, Rv, Ru);
        mat3_mul(R, Rn, T1);
        ul = d3_add(det, mat3_apply(R, d3_sub(ul, det)));
        ur = d3_add(det, mat3_apply(R, d3_sub(ur, det)));
        bl = d3_add(det, mat3_apply(R, d3_sub(bl, det)));
    }
    /* modules (:277-284) */
    const d3 mc = d3_add(det, d3_mul(d3_norm(d3_sub(pos, tgt)),
                                     p->sd - p->module_center_offset_z));
    double mang;
    if (p->has_module_angle)
        mang = p->module_angle_offset;
    else {
        const d3 dd = d3_sub(det, mc);
        mang = (2.0 * atan2((double)width * 0.5 * p->px, sqrt(d3_dot(dd, dd)))) *
               (180.0 / M_PI);
    }
    const double cur = mang * ((double)mod - floor((double)p->module_number / 2.0));
    const double crad = cur * (M_PI / 180.0);
    ul = d3_add(d3_rot_y(d3_sub(ul, mc), crad), mc);
    ur = d3_add(d3_rot_y(d3_sub(ur, mc), crad), mc);
    bl = d3_add(d3_rot_y(d3_sub(bl, mc), crad), mc);
    fr->pe = pos; fr->pa = ul; fr->pb = ur; fr->pc = bl;
    fr->vpl = (double)(width * mod);
    fr->vpr = (double)(width * (mod + 1));
    fr->vpb = 0.0;
    fr->vpt = (double)p->dheight;
}

orc_grid orc_geometry(const orc_geom_params *p, int fp64, double *out64,
                      float *out32)
{
    orc_grid grid;
    d_frame fr;
    frame_d(p, 0, 0, &fr, &grid);
    const int64_t nmod = p->module_number;
    if (fp64) {
        for (int64_t a = 0; a < p->projections; ++a)
            for (int64_t m = 0; m < nmod; ++m) {
                frame_d(p, a, m, &fr, NULL);
                const d44 M = gen_proj_d(fr.pe, fr.pa, fr.pb, fr.pc, fr.vpl, fr.vpr,
                                         fr.vpb, fr.vpt, p->sd);
                double *o = out64 + 20 * (a * nmod + m);
                for (int i = 0; i < 16; ++i) o[i] = M.m[i / 4][i % 4];
                o[16] = fr.vpl; o[17] = fr.vpr; o[18] = fr.vpb; o[19] = fr.vpt;
            }
        return grid;
    }
    /* float32 path, op for op (geometry.hpp:35-76, 244-292, 298-328); the
     * module rotation extension has an FP64 construction only. */
    if (p->roll != 0.0 || p->pitch != 0.0 || p->yaw != 0.0) {
        grid.vwidth = -1;
        return grid;
    }
    const float so = (float)p->so, sd = (float)p->sd, px = (float)p->px;
    const float sx = (float)p->sx, sy = (float)p->sy;
    const float rx = (float)p->rx_px * px;                     /* parameters.hpp:39 */
    const float vx = px * (so / sd);                           /* :50 */
    const int64_t vwidth = p->dwidth;
    const int64_t vheight = p->dheight + (int64_t)fabsf((float)p->heli_offset / vx);
    f3 src = {0.0f, 0.0f, 0.0f};
    f3 at = f3_add(src, (f3){0.0f, 0.0f, sd});
    f3 axis = f3_add(at, (f3){-sx + rx, -sy, 0.0f});
    f3 oc = f3_add(src, f3_mul(f3_norm(f3_sub(axis, src)), so));
    src = f3_sub(src, oc);
    at = f3_sub(at, oc);
    const f3 zero = {0.0f, 0.0f, 0.0f};
    grid.vx = vx;
    grid.vwidth = vwidth;
    grid.vheight = vheight;
    grid.orig[0] = -0.5f * (float)vwidth * vx;
    grid.orig[1] = -0.5f * (float)vheight * vx;
    grid.orig[2] = -0.5f * (float)vwidth * vx;
    const float heli_off = (float)p->heli_offset;
    const float angle = (float)p->angle;
    for (int64_t a = 0; a < p->projections; ++a) {
        const float adeg = p->angle_list ? (float)p->angle_list[a]
                                         : (float)a * (angle / (float)p->projections);
        const float heli = (heli_off / (float)p->projections) * (float)a - heli_off * 0.5f;
        const f3 pos = f3_add(f3_rotate_about(src, zero, adeg), (f3){0.0f, heli, 0.0f});
        const f3 tgt = f3_add(f3_rotate_about(at, zero, adeg), (f3){0.0f, heli, 0.0f});
        for (int64_t m = 0; m < nmod; ++m) {
            /* detectorFromSource (:244-265) */
            const f3 vdir = {0.0f, 1.0f, 0.0f};
            const f3 u = f3_norm(f3_cross(vdir, f3_sub(tgt, pos)));
            const uint64_t width = (uint64_t)p->dwidth / (uint64_t)nmod;
            const uint64_t height = (uint64_t)p->dheight;
            const f3 det = f3_sub(f3_sub(tgt, f3_mul(u, sx)), f3_mul(vdir, sy));
            f3 ul = f3_sub(f3_sub(det, f3_mul(u, 0.5f * (float)width * px)),
                           f3_mul(vdir, 0.5f * (float)height * px));
            f3 ur = f3_add(ul, f3_mul(u, (float)width * px));
            f3 bl = f3_add(ul, f3_mul(vdir, (float)height * px));
            /* computeDetectorProjection (:277-291) */
            const f3 mc = f3_add(det, f3_mul(f3_norm(f3_sub(pos, tgt)),
                                             sd - (float)p->module_center_offset_z));
            float mang;
            if (p->has_module_angle)
                mang = (float)p->module_angle_offset;
            else
                mang = (2.0f * atan2f((float)width * 0.5f * px, f3_len(f3_sub(det, mc)))) *
                       F_DEGREES;
            const float cur = mang * ((float)m - floorf((float)nmod / 2.0f));
            ul = f3_rotate_about(ul, mc, cur);
            ur = f3_rotate_about(ur, mc, cur);
            bl = f3_rotate_about(bl, mc, cur);
            const float vpl = (float)(width * (uint64_t)m);
            const float vpr = (float)(width * (uint64_t)(m + 1));
            const float vpt = (float)p->dheight;
            const f44 M = gen_proj_f(pos, ul, ur, bl, vpl, vpr, 0.0f, vpt, sd);
            float *o = out32 + 20 * (a * nmod + m);
            for (int i = 0; i < 16; ++i) o[i] = M.m[i / 4][i % 4];
            o[16] = vpl; o[17] = vpr; o[18] = 0.0f; o[19] = vpt;
        }
    }
    return grid;
}

void orc_project_physical(const orc_geom_params *p, int64_t a, int64_t m, double x,
                          double y, double z, double *u, double *v)
{
    d_frame fr;
    frame_d(p, a, m, &fr, NULL);
    const d3 P = {x, y, z};
    const d3 er = d3_sub(fr.pb, fr.pa), eu = d3_sub(fr.pc, fr.pa);
    const d3 nrm = d3_cross(er, eu);
    const d3 dir = d3_sub(P, fr.pe);
    const double t = d3_dot(d3_sub(fr.pa, fr.pe), nrm) / d3_dot(dir, nrm);
    const d3 hit = d3_add(fr.pe, d3_mul(dir, t));
    const d3 rel = d3_sub(hit, fr.pa);
    *u = fr.vpl + d3_dot(rel, er) / d3_dot(er, er) * (fr.vpr - fr.vpl);
    *v = fr.vpb + d3_dot(rel, eu) / d3_dot(eu, eu) * (fr.vpt - fr.vpb);
}

/* ------------------------------------------------------------------------ */
/* Synthetic data.                                                          */
/* ------------------------------------------------------------------------ */
static double chord(const orc_ellipsoid *e, d3 o, d3 d)
{
    const double *R = e->rot;
    const d3 q0 = d3_sub(o, (d3){e->c[0], e->c[1], e->c[2]});
    const d3 q = {(R[0] * q0.x + R[1] * q0.y + R[2] * q0.z) / e->a[0],
                  (R[3] * q0.x + R[4] * q0.y + R[5] * q0.z) / e->a[1],
                  (R[6] * q0.x + R[7] * q0.y + R[8] * q0.z) / e->a[2]};
    const d3 k = {(R[0] * d.x + R[1] * d.y + R[2] * d.z) / e->a[0],
                  (R[3] * d.x + R[4] * d.y + R[5] * d.z) / e->a[1],
                  (R[6] * d.x + R[7] * d.y + R[8] * d.z) / e->a[2]};
    const double A = d3_dot(k, k), B = 2.0 * d3_dot(q, k), C = d3_dot(q, q) - 1.0;
    const double disc = B * B - 4.0 * A * C;
    if (disc <= 0.0) return 0.0;
    /* t-interval length times |d| */
    return sqrt(disc) / A * sqrt(d3_dot(d, d));
}

void orc_line_integrals(const orc_geom_params *p, int64_t a, const orc_ellipsoid *e,
                        int ne, double *out)
{
    const int64_t nmod = p->module_number;
    const int64_t width = p->dwidth / nmod;
    for (int64_t m = 0; m < nmod; ++m) {
        d_frame fr;
        frame_d(p, a, m, &fr, NULL);
        const d3 er = d3_mul(d3_sub(fr.pb, fr.pa), 1.0 / (fr.vpr - fr.vpl));
        const d3 eu = d3_mul(d3_sub(fr.pc, fr.pa), 1.0 / (fr.vpt - fr.vpb));
        for (int64_t row = 0; row < p->dheight; ++row)
            for (int64_t c = m * width; c < (m + 1) * width; ++c) {
                /* pixel (c, row) is centred at detector coordinate (c, row) */
                const d3 hit = d3_add(d3_add(fr.pa, d3_mul(er, (double)c - fr.vpl)),
                                      d3_mul(eu, (double)row - fr.vpb));
                const d3 dir = d3_sub(hit, fr.pe);
                double s = 0.0;
                for (int j = 0; j < ne; ++j) s += e[j].mu * chord(&e[j], fr.pe, dir);
                out[row * p->dwidth + c] = s;
            }
    }
}

void orc_phantom_voxels(const orc_grid *g, const orc_ellipsoid *e, int ne, double *out)
{
    for (int64_t l = 0; l < g->vheight; ++l)
        for (int64_t k = 0; k < g->vwidth; ++k)
            for (int64_t i = 0; i < g->vwidth; ++i) {
                const d3 P = {g->orig[0] + ((double)i + 0.5) * g->vx,
                              g->orig[1] + ((double)l + 0.5) * g->vx,
                              g->orig[2] + ((double)k + 0.5) * g->vx};
                double s = 0.0;
                for (int j = 0; j < ne; ++j) {
                    const double *R = e[j].rot;
                    const d3 q0 = d3_sub(P, (d3){e[j].c[0], e[j].c[1], e[j].c[2]});
                    const double bx = (R[0] * q0.x + R[1] * q0.y + R[2] * q0.z) / e[j].a[0];
                    const double by = (R[3] * q0.x + R[4] * q0.y + R[5] * q0.z) / e[j].a[1];
                    const double bz = (R[6] * q0.x + R[7] * q0.y + R[8] * q0.z) / e[j].a[2];
                    if (bx * bx + by * by + bz * bz <= 1.0) s += e[j].mu;
                }
                out[(l * g->vwidth + k) * g->vwidth + i] = s;
            }
}
