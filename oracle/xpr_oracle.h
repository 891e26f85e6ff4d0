/*
 * CPU oracle for the iterative cone-beam (XPRecon) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (xpulse_recon_amd/) never links, imports or calls it.
 *
 * It restates, in plain C, the reference's OpenCL kernels and host-side
 * arithmetic under one declared canonical arithmetic (explicit fmaf, IEEE
 * division, roundf half-away-from-zero, no contraction).  Every function
 * cites the reference file:line it follows (paths relative to the reference
 * checkout root).
 *
 * Layouts (shared with the HIP path):
 *   volume slab : float [layer][depth][x], x fastest (to3D, utils.cl:39-45)
 *   ProjData    : 20 floats per (view, module): 16 = glm column-major 4x4,
 *                 then viewport (vpl, vpr, vpb, vpt)  (constants.cl:25-28,
 *                 geometry.hpp:313-323)
 *   detector    : [view][row][col], row stride det_width, view stride
 *                 det_width*line_rows  (IMAGE_OFFSET_2D, utils.cl:47)
 */
#ifndef XPR_ORACLE_H
#define XPR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel arguments of Forward/Backward (KernelForward.cl:10-21). */
typedef struct {
    uint32_t det_width;    /* size.x                                    */
    uint32_t n_views;      /* angleNumber                               */
    uint32_t modules;      /* modulePerAngle                            */
    int32_t  vmin[3];      /* volumeMin (x, layer, depth)               */
    int32_t  vmax[3];      /* volumeMax                                 */
    int32_t  line_offset;  /* lineOffsetY.x : first detector row held   */
    int32_t  line_rows;    /* lineOffsetY.y : detector rows held        */
    float    origin[3];    /* origin (mm), already jittered             */
    float    voxel[3];     /* voxelSize (mm)                            */
} orc_geom;

/* ---- FP32 canonical restatement of the three kernels ------------------ */

/* Forward (KernelForward.cl:22-37).  acc_bits = 32: sums are uint32 Q10.22
 * (wrap mod 2^32, value saturated to [0, 2^32-1]); 64: uint64 Q.22.
 * Returns the number of in-bounds voxel-views. */
uint64_t orc_forward_f32(const float *vol, void *sums, int acc_bits,
                         const float *projdata, const orc_geom *g, int threads);

/* Error (KernelError.cl:14-23).  sums read as uint32 or uint64 Q.22; E is
 * written to err (may alias sums when acc_bits == 32, like the reference). */
void orc_error_f32(const float *images, const void *sums, int acc_bits,
                   float *err, uint64_t n, float weight);

/* Backward + multiplicative update (KernelBackward.cl:20-40).  Voxels with
 * no in-bounds view keep their value; their count is returned. */
uint64_t orc_backward_f32(float *vol, const float *err, const float *projdata,
                          const orc_geom *g, int threads);

/* Hit map: number of in-bounds views per voxel (brute force, same canonical
 * coordinates).  counts has one entry per voxel of the slab. */
void orc_hits_f32(uint32_t *counts, const float *projdata, const orc_geom *g);

/* ---- FP64 model of the same voxel-driven NN operators ------------------ */

typedef struct {
    uint32_t det_width, n_views, modules;
    int32_t  vmin[3], vmax[3];
    int32_t  line_offset, line_rows;
    double   origin[3];
    double   voxel[3];
} orc_geom64;

void orc_forward_f64(const double *vol, double *sums, const double *projdata,
                     const orc_geom64 *g);
void orc_error_f64(const double *images, const double *sums, double *err,
                   uint64_t n, double weight);
uint64_t orc_backward_f64(double *vol, const double *err, const double *projdata,
                          const orc_geom64 *g);

/* FP64 values on the canonical float32 pixel choice (project_f32): the
 * voxel-views of the FP32 kernels with double sums, error and update, so
 * that only value arithmetic differs from the FP32 path.  q22 != 0 truncates
 * each forward contribution to a multiple of 2^-22 (FLOAT_TO_FIXED). */
void orc_forward_f64nn(const double *vol, double *sums, const float *projdata,
                       const orc_geom *g, int threads, int q22);
uint64_t orc_backward_f64nn(double *vol, const double *err, const float *projdata,
                            const orc_geom *g, int threads);

/* ---- opt-in bilinear backprojector (not the reference's arithmetic; the
 *      intent of its unused INTERP_BILIN branch, KernelForward.cl:38-53) -- */

/* FP32 canonical restatement: the voxel-views the nearest-neighbour test
 * accepts, taps (floor u, floor v) .. (+1, +1) clamped to the accepted
 * pixels, three fmaf lerps, update v * expf(sum/n). */
uint64_t orc_backward_bilinear_f32(float *vol, const float *err, const float *projdata,
                                   const orc_geom *g, int threads);
/* FP64 model on the float32 accept set and clamped taps: double weights,
 * sample, sum and update. */
uint64_t orc_backward_bilinear_f64nn(double *vol, const double *err, const float *projdata,
                                     const orc_geom *g, int threads);

/* ---- preprocessing (data_loader.hpp:55-86) ----------------------------- */

/* Global min/max over isnormal() samples (data_loader.hpp:55-74).
 * *mn / *mx must hold the running values (start FLT_MAX / FLT_MIN). */
void orc_minmax_normal(const float *data, uint64_t n, float *mn, float *mx);
/* In-place normalise (if normalize) then -log clamp (if mlog)
 * (data_loader.hpp:80-86). */
void orc_preprocess(float *data, uint64_t n, int normalize, int mlog,
                    float mn, float mx);

/* ---- geometry (geometry.hpp:35-76, 182-328) ---------------------------- */

typedef struct {
    double so, sd, angle;          /* geometry: so, sd, angle (deg)          */
    double heli_offset;            /* geometry.heli_offset (mm)              */
    double px, rx_px, sx, sy;      /* detector: px, rx (pixels), sx, sy (mm) */
    int64_t module_number;
    int32_t has_module_angle;      /* module_angle_offset assigned?          */
    double module_angle_offset;    /* deg                                    */
    double module_center_offset_z;
    int64_t dwidth, dheight, projections;
    const double *angle_list;      /* NULL: i*angle/projections              */
    double roll, pitch, yaw;       /* deg, module rotation about the detector
                                      centre (FP64 construction only)        */
} orc_geom_params;

typedef struct {
    double vx;
    int64_t vwidth, vheight;
    double orig[3];
} orc_grid;

/* ProjData for every (angle, module) in angle order: out has
 * projections*module_number*20 entries.  fp64 != 0: double arithmetic
 * (out64 written); else float32 arithmetic following glm op for op (out32
 * written).  Returns the grid (vx, vwidth, vheight, orig). */
orc_grid orc_geometry(const orc_geom_params *p, int fp64, double *out64,
                      float *out32);

/* Physical ray/plane projection used to pin the matrices: pixel coordinate
 * (u, v) of point (x, y, z) for angle index a, module m, computed from the
 * detector corners directly (no matrix). */
void orc_project_physical(const orc_geom_params *p, int64_t a, int64_t m,
                          double x, double y, double z, double *u, double *v);

/* ---- synthetic data: ellipsoid phantom and analytic line integrals ---- */

typedef struct {
    double c[3];      /* centre (mm, object frame)            */
    double a[3];      /* semi-axes (mm)                        */
    double rot[9];    /* row-major rotation: body = rot*(x-c)  */
    double mu;        /* attenuation per mm                    */
} orc_ellipsoid;

/* Line integral of the phantom along source->pixel-centre rays for every
 * pixel of angle index a (all modules): out[row*dwidth + col]. */
void orc_line_integrals(const orc_geom_params *p, int64_t a,
                        const orc_ellipsoid *e, int ne, double *out);

/* Phantom sampled at voxel corners of the reconstruction grid (same points
 * the projectors use), out [layer][depth][x]. */
void orc_phantom_voxels(const orc_grid *g, const orc_ellipsoid *e, int ne,
                        double *out);

#ifdef __cplusplus
}
#endif
#endif
