/*
 * xpr_hip.h - C-ABI of the MI355X (gfx950) iterative cone-beam hot path.
 *
 * Plain C: pointers, sizes and POD structs only.  Every device pointer is a
 * HIP device address owned by the caller; the library keeps no pointer past a
 * call.  Every call is stateless (all kernel arguments are passed per call),
 * enqueues on the given stream (NULL = the null stream) and returns an int
 * status: XPR_OK or a negative XPR_E_* code, with a message in
 * xpr_last_error() (thread-local).  No C++ exception crosses this boundary.
 *
 * Each entry point replaces one piece of the reference's OpenCL operator
 * API; the citation names the reference interface it stands in for
 * (paths relative to the reference checkout root).
 */
#ifndef XPR_HIP_H
#define XPR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPR_ABI_VERSION 2

/* status codes (stand in for CHECK -> std::runtime_error, gpu/error.hpp:17-27,
 * and std::bad_alloc, gpu/context.hpp:329-336) */
#define XPR_OK            0
#define XPR_E_INVALID    -1   /* bad argument / shape                        */
#define XPR_E_HIP        -2   /* HIP runtime error                           */
#define XPR_E_NOMEM      -3   /* device allocation failed / over budget      */
#define XPR_E_UNSUPPORTED -4  /* dtype or mode not built                     */

/* storage types of volume, projection and error buffers */
#define XPR_F32  0
#define XPR_F16  1
#define XPR_BF16 2

/* Accumulators of the detector sums buffer (the acc_bits arguments).  32
 * reproduces the reference's uint32 Q10.22 (wraps mod 2^32,
 * constants.cl:17-22); 64 is the same sum without wrap-around.  Both are
 * integers: the sums do not depend on launch order. */
#define XPR_ACC_U32 32
#define XPR_ACC_U64 64
/* Opt-in, not the reference's arithmetic: float32 sums.  The forward bins the
 * same Q.22 values exactly in integer LDS tiles and adds each binned sum
 * (times 2^-22, rounded once to float) to the detector pixel with a global
 * float atomic.  The order of those additions changes from launch to launch,
 * so the sums do, within float rounding of the exact Q.22 sum. */
#define XPR_ACC_F32 (-32)

/* Detector sampling of the backprojector (xpr_proj_args.interp). */
#define XPR_INTERP_NEAREST  0  /* the reference's: the pixel roundf picks (KernelBackward.cl:34) */
/* Opt-in, not the reference's arithmetic: the voxel-views the
 * nearest-neighbour test accepts sample the error images bilinearly at the
 * continuous coordinate (u, v) (pixel k centred at k), taps (floor u, floor
 * v) .. (+1, +1) clamped to the accepted pixels (the reference's unused
 * INTERP_BILIN branch, KernelForward.cl:38-53, states the intent; it never
 * compiled). */
#define XPR_INTERP_BILINEAR 1

typedef void *xpr_stream;   /* a hipStream_t */

/* Kernel arguments of Forward and Backward, passed by value instead of
 * positional OpenCL arguments (constants.cl:30-54; KernelForward.cl:10-21;
 * set by Reconstruction::setImageParameters / setVolumeParameters /
 * setAngleNumber / setOrigin / setImageOffset, src/reconstruction.hpp:208-281). */
typedef struct {
    uint32_t det_width;    /* size.x: detector columns (= dwidth)                    */
    uint32_t n_views;      /* angleNumber: views in this launch                      */
    uint32_t modules;      /* modulePerAngle: ProjData records per view              */
    int32_t  vmin[3];      /* volumeMin (x, layer, depth) of the slab, voxels        */
    int32_t  vmax[3];      /* volumeMax (exclusive)                                  */
    int32_t  line_offset;  /* lineOffsetY.x: first detector row held in the buffers  */
    int32_t  line_rows;    /* lineOffsetY.y: detector rows held per view             */
    float    origin[3];    /* origin (mm) of voxel (0,0,0), jitter included          */
    float    voxel[3];     /* voxelSize (mm)                                         */
    /* XPR_INTERP_NEAREST (0, the reference's) or XPR_INTERP_BILINEAR: how
     * xpr_backward samples the error images.  The forward always uses the
     * nearest pixel.  (It fills what would be padding before view_index.) */
    int32_t  interp;
    /* Optional device array of n_views detector view indices: view j of this
     * call reads ProjData record j and detector view view_index[j] (the
     * per-layer view lists of ForwardH/BackwardH, KernelForwardH.cl:8-37,
     * geometry.hpp:141-171).  NULL: view j is detector view j. */
    const uint32_t *view_index;
} xpr_proj_args;

/* ---- library / device --------------------------------------------------- */

/* Version of this ABI (XPR_ABI_VERSION). */
int xpr_abi_version(void);
/* Last error message of the calling thread ("" if none). */
const char *xpr_last_error(void);
/* Select the device for the calling thread; replaces gpu::Context's device
 * pick (gpu/context.hpp:36-71). */
int xpr_set_device(int device);
/* Free and total device memory in bytes (gpu/context.hpp:87-110). */
int xpr_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- buffers (gpu::Context::createBuffer / setBuffer / getBuffer,
 *      gpu/context.hpp:152-300) -------------------------------------------- */
int xpr_alloc(void **dptr, size_t bytes);
int xpr_free(void *dptr);
int xpr_host_alloc(void **hptr, size_t bytes);   /* pinned host memory */
/* Stream-ordered allocation (no device-wide synchronisation), for work that
 * runs concurrently on several streams. */
int xpr_alloc_async(void **dptr, size_t bytes, xpr_stream s);
int xpr_free_async(void *dptr, xpr_stream s);
int xpr_host_free(void *hptr);
int xpr_upload(xpr_stream s, void *dst, const void *src, size_t bytes);   /* async */
int xpr_download(xpr_stream s, void *dst, const void *src, size_t bytes); /* async */
int xpr_stream_create(xpr_stream *s);
int xpr_stream_destroy(xpr_stream s);
int xpr_stream_sync(xpr_stream s);
/* Events: ordering between a copy stream and the compute stream, and device
 * timing (the reference enables queue profiling, gpu/context.hpp:80-84). */
typedef void *xpr_event;    /* a hipEvent_t */
int xpr_event_create(xpr_event *e);
int xpr_event_destroy(xpr_event e);
int xpr_event_record(xpr_event e, xpr_stream s);
int xpr_stream_wait_event(xpr_stream s, xpr_event e);
int xpr_event_sync(xpr_event e);
int xpr_event_elapsed_ms(xpr_event start, xpr_event stop, float *ms);
/* Fill kernels (enqueueFillBuffer, gpu/context.hpp:222-239): n elements. */
int xpr_fill_u32(xpr_stream s, uint32_t *dst, uint32_t value, uint64_t n);
int xpr_fill_u64(xpr_stream s, uint64_t *dst, uint64_t value, uint64_t n);
/* Fill a buffer of dtype XPR_F32/F16/BF16 with value (rounded to nearest even). */
int xpr_fill(xpr_stream s, void *dst, int dtype, float value, uint64_t n);
/* Convert n elements between storage types (round to nearest even). */
int xpr_convert(xpr_stream s, void *dst, int dst_dtype, const void *src, int src_dtype,
                uint64_t n);

/* ---- the hot path -------------------------------------------------------- */

/* Forward projection (Reconstruction::forward, src/reconstruction.hpp:294-296;
 * kernel Forward, KernelForward.cl:10-55).  Adds trunc(v*2^22) of every voxel
 * of the slab into the detector pixel each view rounds it to:
 *   volume   : slab [layer][depth][x] of (vmax-vmin) voxels, dtype vol_dtype
 *   sums     : uint32 or uint64 Q.22 [n_views][line_rows][det_width]
 *              (acc_bits 32 / 64), or float32 (acc_bits XPR_ACC_F32, opt-in:
 *              the same Q.22 values binned exactly per brick, the bins added
 *              as trunc(v*2^22)*2^-22 sums with float atomics)
 *   projdata : device float[n_views*modules][20] (ProjData, constants.cl:25-28)
 * Integer accumulation: the result does not depend on launch order. */
int xpr_forward(xpr_stream s, const void *volume, int vol_dtype, void *sums, int acc_bits,
                const float *projdata, const xpr_proj_args *a);

/* Error (Reconstruction::error, src/reconstruction.hpp:302-304; kernel Error,
 * KernelError.cl:8-24): err = log(max(img / max(sum*2^-22, eps), eps)) * weight,
 * eps = 2^-22, for n pixels (float sums, XPR_ACC_F32: sum itself in place of
 * sum*2^-22).  err may alias sums when acc_bits == 32 and err_dtype ==
 * XPR_F32 (the reference's in-place form). */
int xpr_error(xpr_stream s, const void *images, int img_dtype, const void *sums, int acc_bits,
              void *err, int err_dtype, uint64_t n, float weight);

/* Backward projection and multiplicative update (Reconstruction::backward,
 * src/reconstruction.hpp:310-312; kernel Backward, KernelBackward.cl:8-41):
 *   v <- v * exp(sum_j err_j(pixel_j(v)) / n_in_bounds)
 * Voxels with no in-bounds view keep their value (the reference computes
 * exp(0/0)); if empty_count (device uint64) is non-NULL their number is added
 * to it.  a->interp == XPR_INTERP_BILINEAR (opt-in) samples err bilinearly
 * at (u, v) instead, over the same in-bounds voxel-views: with
 * fx = u - floor(u), fy = v - floor(v), columns i0 = max(floor u, cl),
 * i1 = min(floor u + 1, cr) and rows j0, j1 likewise, [cl, cr] the accepted
 * columns ceil(max(vpl, 0)) .. ceil(min(vpr, det_width)) - 1 (rows with vpb,
 * vpt and line_rows),
 *   top = fmaf(fx, E[j0][i1] - E[j0][i0], E[j0][i0]),
 *   bot = fmaf(fx, E[j1][i1] - E[j1][i0], E[j1][i0]),
 *   sample = fmaf(fy, bot - top, top). */
int xpr_backward(xpr_stream s, void *volume, int vol_dtype, const void *err, int err_dtype,
                 const float *projdata, const xpr_proj_args *a, uint64_t *empty_count);

/* No counterpart in the reference.  xpr_backward that also reports the
 * largest float32 voxel value of the slab after the update, updated or kept
 * (computed before a 16-bit store rounds it): max_bits is a device array of
 * XPR_MAX_SLOTS uint32, each raised to the bits of the largest value of the
 * rows it serves (spread so that the waves' atomics do not queue on one
 * address); the largest slot is the slab's.  Non-negative floats order as
 * their bits; a negative or NaN value reads as a huge one.  Clear the slots
 * before the call.  For callers that bound the next forward's detector sums
 * (xpr_hit_bound). */
#define XPR_MAX_SLOTS 1024
int xpr_backward_max(xpr_stream s, void *volume, int vol_dtype, const void *err, int err_dtype,
                     const float *projdata, const xpr_proj_args *a, uint64_t *empty_count,
                     uint32_t *max_bits);

/* One sub-iteration of the slab (the body of ReconstructionN::exec's loop,
 * reconstruction_n.hpp:66-76, with the values setOrigin / setWeight would
 * set, src/reconstruction.hpp:255-272, in *a and weight):
 *   sums <- 0, then xpr_forward of the volume into them; on the first pass
 *           of a reconstruction (first_pass != 0) no forward and every sum
 *           is 1.0, i.e. 1 << 22 (reconstruction_n.hpp:60,70-72);
 *   err  <- xpr_error(images, sums, weight) over n_views*line_rows*det_width
 *           pixels;
 *   xpr_backward of err into the volume (update in place).
 * images: [n_views][line_rows][det_width] of img_dtype.  sums (acc_bits)
 * and err (err_dtype) are the caller's workspace, n_views*line_rows*
 * det_width elements each; err may alias sums when acc_bits == 32 and
 * err_dtype == XPR_F32.  Their contents on entry do not matter; on return
 * err holds the error images.  a->view_index must be NULL. */
int xpr_subiteration(xpr_stream s, void *volume, int vol_dtype, const void *images,
                     int img_dtype, const float *projdata, const xpr_proj_args *a, float weight,
                     int first_pass, void *sums, int acc_bits, void *err, int err_dtype,
                     uint64_t *empty_count);

/* No counterpart in the reference.  xpr_subiteration for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the sums and error workspace live in a
 * static device array of XPR_STATIC_SCRATCH_BYTES per GPU, held from the
 * library's load until the process ends; the call allocates nothing.
 * n_views*line_rows*det_width*(|acc_bits|/8 + error sample bytes) + 256 must
 * not exceed XPR_STATIC_SCRATCH_BYTES (XPR_E_INVALID otherwise: e.g. 36 views
 * of 384x384 with 64-bit sums and f32 error fit, 36 views of 768x768 do not).
 * The sums are cleared inside the call.  All calls on one GPU share its
 * array: calls that overlap on one GPU (concurrent streams or threads) race
 * and must be serialised by the caller. */
#define XPR_STATIC_SCRATCH_BYTES (64ull << 20)
int xpr_subiteration_static(xpr_stream s, void *volume, int vol_dtype, const void *images,
                            int img_dtype, const float *projdata, const xpr_proj_args *a,
                            float weight, int first_pass, int acc_bits, int err_dtype,
                            uint64_t *empty_count);

/* Q.22 sums to float (FIXED_TO_FLOAT, constants.cl:21; float sums are
 * copied) for n pixels; used to write reprojected views
 * (Reconstruction::forward_exec,
 * src/reconstruction.hpp:139-153). */
int xpr_q22_to_float(xpr_stream s, const void *sums, int acc_bits, float *out, uint64_t n);

/* x <- x * factor for n floats: the de-normalisation of the final layers
 * (DataLoader::saveLayer with finalize, data_loader.hpp:135-139). */
int xpr_scale(xpr_stream s, float *x, uint64_t n, float factor);

/* ---- preprocessing (DataLoader::initializeTempImages,
 *      dataset/data_loader.hpp:48-101) -------------------------------------- */

/* Running global min/max over isnormal() samples (data_loader.hpp:55-74).
 * minmax (device, 2 x int32 order keys) must be initialised with
 * xpr_minmax_init; read it back with xpr_minmax_read. */
int xpr_minmax_init(xpr_stream s, int32_t *minmax);
int xpr_minmax_normal(xpr_stream s, const float *data, uint64_t n, int32_t *minmax);
int xpr_minmax_read(const int32_t *minmax_host, float *mn, float *mx);
/* Normalise (v-min)/(max-min) if normalize, then -log(clamp(v, eps, 1-eps))
 * if mlog (data_loader.hpp:80-86), from float input to dtype out_dtype. */
int xpr_preprocess(xpr_stream s, const float *in, void *out, int out_dtype, uint64_t n,
                   int normalize, int mlog, float mn, float mx);
/* xpr_preprocess of n_views views of n elements in one launch: view j is read
 * from in + j*in_view_stride (elements; >= n) and written to out + j*n.  The
 * views of one sub-iteration lie at a fixed stride in the projection stack
 * (Dataset::getSitImages, dataset.hpp:133-140: views s, s+S, ...). */
int xpr_preprocess_views(xpr_stream s, const float *in, uint64_t in_view_stride, void *out,
                         int out_dtype, uint64_t n, uint32_t n_views, int normalize, int mlog,
                         float mn, float mx);

/* ---- tile-shape tuning (no counterpart in the reference: its kernels have a
 *      fixed work-item shape, gpu/kernel.hpp:73-81) ------------------------------
 *
 * Time the compiled tile shapes of a projector on the caller's buffers and keep
 * the fastest in a per-process table keyed by the launch's shape (dtypes, slab
 * width and depth, views, modules, detector rows); later xpr_forward /
 * xpr_backward calls of that shape use it.  The shapes change only how work is
 * spread over lanes and paths: results are identical whatever is chosen.
 * Forward: LDS tile capacity (entries); the scratch sums receive several
 * forward passes.  Backward: voxels per lane (4, 8, 16: 1, 2 or 4 x positions
 * in each of 4 layers); the scratch volume receives several updates.  *choice receives the shape kept. */
int xpr_tune_forward(xpr_stream s, const void *volume, int vol_dtype, void *sums_scratch,
                     int acc_bits, const float *projdata, const xpr_proj_args *args, int *choice);
int xpr_tune_backward(xpr_stream s, void *volume_scratch, int vol_dtype, const void *err,
                      int err_dtype, const float *projdata, const xpr_proj_args *args, int *choice);
/* Forget every tuned shape (defaults apply again). */
int xpr_tune_reset(void);

/* No counterpart in the reference: a self-check of the two hardware
 * identities the projectors' exact fast paths rely on, over every float32 bit
 * pattern (one pass of a few milliseconds).  mismatches[0] counts w with
 * normal |w| in [1e-6, 1e6] where v_rcp_f32 plus one Newton FMA differs from
 * the IEEE quotient 1/w; mismatches[1] counts u in (-0.5, 2^23) where
 * v_cvt_rpi_i32_f32 differs from (int)roundf(u).  Both are 0 on a correct
 * device; mismatches[2] counts the reciprocal's mismatches outside that range
 * (where the kernels use the IEEE division).  Synchronous. */
int xpr_check_fast_paths(xpr_stream s, uint64_t mismatches[3]);

/* No counterpart in the reference: the device's VALU issue ceiling, the
 * roofline of the projectors (VALU-bound: a dozen vector instructions per
 * voxel-view against well under a byte of HBM traffic).  Kernels of 16
 * independent instruction chains per lane at 8 waves per SIMD on every CU,
 * one kind each: 0 v_fma_f32, 1 v_mul_f32, 2 v_add_f32 (one VGPR read each),
 * 3, 4, 5 with 8, 12, 16 of the chains v_xor_b32 of two VGPRs and the rest
 * v_fma_f32 (gfx950 issues such integer instructions faster than FMAs).
 * rates[k] receives kind k's vector lane-instructions per second (the best of
 * five launches after a warm-up; n_rates up to 6, 0 beyond);
 * xpr_valu_ceiling the best kind's.  Synchronous; allocate a small output
 * buffer. */
int xpr_valu_rates(xpr_stream s, double *rates, int n_rates);
int xpr_valu_ceiling(xpr_stream s, double *lane_instr_per_s);

/* ---- geometry (Geometry, dataset/geometry.hpp:35-76, 182-328) ----------- */

/* JSON "geometry" and "detector" sections (global_parameters.hpp:23-54).
 * rx is in pixels (multiplied by px at load, parameters.hpp:39); sx, sy in mm. */
typedef struct {
    double so, sd, angle;            /* source-object, source-detector (mm), total angle (deg) */
    double heli_offset;              /* mm */
    double px, rx, sx, sy;           /* pixel pitch (mm), axis offset (px), detector shift (mm) */
    double module_angle_offset;      /* deg, used when has_module_angle */
    double module_center_offset_z;   /* mm */
    double roll, pitch, yaw;         /* deg, module rotation about the detector centre
                                        (extension; the reference parses none of them) */
    int64_t module_number;
    int64_t dwidth, dheight, projections;
    int32_t has_module_angle;
    int32_t fp64;                    /* 1: FP64 construction; 0: float32 as the reference */
    const double *angle_list;        /* NULL: angle i = i*angle/projections (deg) */
} xpr_geometry_params;

/* Reconstruction grid (geometry.hpp:50-68): vx = px*so/sd, W = dwidth,
 * H = dheight + |heli_offset/vx|, origin of voxel (0,0,0) = -(W,H,W)*vx/2. */
typedef struct {
    double vx;
    int64_t vwidth, vheight;
    double orig[3];
} xpr_grid;

/* ProjData of every (angle, module), angle-major: out has
 * projections*module_number*20 floats (Geometry::computeMVPs,
 * geometry.hpp:298-328).  Host function, no device work. */
int xpr_geometry(const xpr_geometry_params *p, float *projdata_out, xpr_grid *grid_out);

/* Detector frames of every (angle, module) in FP64, 14 doubles each: source
 * position, upper-left detector corner, column step, row step (mm per pixel),
 * vpl, vpb (geometry.hpp:244-292).  Used by the synthetic-data generator. */
int xpr_geometry_frames(const xpr_geometry_params *p, double *frames_out);

/* No counterpart in the reference.  *hits receives a bound on the voxels of
 * an xpr_forward launch of args *a whose nearest pixel is one pixel of one
 * view (any pixel, any view, summed over a view's modules), valid for every
 * origin in [a->origin, a->origin + a->voxel] per axis: pass the grid origin
 * without jitter.  projdata: host copy of the launch's n_views * modules
 * records.  +inf when some voxel of the box is not in front of the source.
 * Every 64-bit detector sum of the launch is then at most *hits times the
 * largest Q.22 voxel value, and when that is below 2^32 the launch's 32-bit
 * sums (XPR_ACC_U32) equal its 64-bit sums.  a->view_index must be NULL.
 * Host function, no device work. */
int xpr_hit_bound(const float *projdata, const xpr_proj_args *a, double *hits);

/* ---- synthetic data (no counterpart in the reference: test and benchmark
 *      input generation) ------------------------------------------------------ */

/* One ellipsoid of a phantom: centre c, semi-axes a (mm), row-major rotation
 * rot (body = rot * (x - c)), attenuation mu per mm. */
typedef struct {
    double c[3], a[3], rot[9], mu;
} xpr_ellipsoid;

/* FP64 line integrals of a sum of ellipsoids along the ray from the source to
 * the centre of every pixel (pixel (col,row) is centred at detector coordinate
 * (col,row), the reference's convention).  frames: device copy of
 * xpr_geometry_frames for the n_records (view, module) records written; out:
 * device double [n_views][det_height][det_width]. */
int xpr_synth_line_integrals(xpr_stream s, const double *frames, uint32_t n_views,
                             uint32_t modules, uint32_t det_width, uint32_t det_height,
                             const xpr_ellipsoid *ellipsoids, int n_ellipsoids, double *out);
/* The phantom on the reconstruction grid: out[l][k][i] (device float, vwidth^2
 * * vheight) = sum of mu over the ellipsoids (device array) that contain the
 * centre of voxel (i, l, k), orig + (i, l, k)*vx + vx/2 (ground truth for
 * reconstruction-error figures). */
int xpr_synth_phantom(xpr_stream s, double vx, int64_t vwidth, int64_t vheight,
                      const double orig[3], const xpr_ellipsoid *ellipsoids, int n_ellipsoids,
                      float *out);

/* ---- the XPRecon drop-in (src/main.cpp:62-137) ----------------------------- */

/* Runs one reconstruction in-process from a parameter file path or inline JSON
 * (the XPRecon CLI body).  Returns the CLI's exit codes: 0 OK, -1 arguments,
 * -2 parameters, -3 reconstruction.  If record is non-NULL it receives the
 * JSON run record (truncated to record_cap). */
int xpr_reconstruct(const char *params, char *record, size_t record_cap);
/* xpr_reconstruct with the projections and the volume in host memory instead
 * of TIFF files (recon.input / recon.output are not used; reprojections still
 * go to recon.proj_output): projections [count][height][width] raw (as read
 * from the TIFFs, sorted by name), volume_out [vheight][vwidth][vwidth]
 * float32, de-normalised like the written layers.  Any schedule of
 * recon.method. */
int xpr_reconstruct_mem(const char *params, const float *projections, int64_t width,
                        int64_t height, int64_t count, float *volume_out, char *record,
                        size_t record_cap);
/* xpr_reconstruct_mem with both buffers in device memory (HBM) of the
 * current device: the projections are preprocessed straight from
 * `projections` (device float32, [count][height][width], raw) and the
 * de-normalised volume is written to `volume_out` (device float32,
 * [vheight][vwidth][vwidth]); no host copy of either is made. */
int xpr_reconstruct_dev(const char *params, const float *projections, int64_t width,
                        int64_t height, int64_t count, float *volume_out, char *record,
                        size_t record_cap);
/* The method_2 chunk plan (GeometryChunk, dataset/geometry_chunk.hpp:92-197,
 * made by ReconstructionChunk's constructor, reconstruction_chunk.hpp:13-25)
 * for `count` projections of width x height under the parameters' geometry,
 * storage and xpr.device_budget_gb (required).  Host function, no device
 * work and no file access.  chunks_out receives 6 int64 per chunk: vOffset,
 * vSize, iOffset, iSize, vPadTop, vPadBot (the order of the run record's
 * "chunks"); *n_chunks the number of chunks (also when cap is too small,
 * which returns XPR_E_INVALID).  XPR_E_NOMEM when a single layer does not fit
 * the budget. */
int xpr_plan_chunks(const char *params, int64_t width, int64_t height, int64_t count,
                    int64_t *chunks_out, int64_t cap, int64_t *n_chunks);
/* The default parameters as JSON (XPRecon -template).  0 on success, -1 if
 * cap is too small. */
int xpr_template_json(char *out, size_t cap);

/* ---- serving: concurrent reconstructions under an HBM budget ----------------
 * (no counterpart in the reference, which runs one reconstruction per process)
 *
 * Jobs are admitted in arrival order while the device memory they need fits in
 * the budget and a concurrency slot is free; each admitted job runs the naive
 * schedule on its own stream, overlapping the jobs already running.  A job that
 * cannot fit the budget even alone runs with slab streaming (volume held on the
 * host).  Every job owns its buffers and its seeded jitter, so its result does
 * not depend on what else runs. */
typedef struct xpr_server xpr_server;

typedef struct {
    xpr_geometry_params geometry;    /* detector size and projection count included */
    int64_t it, sit;
    float weight, weight_factor;
    int32_t mlog, normalize;
    int32_t storage;                 /* XPR_F32 / XPR_F16 / XPR_BF16 */
    int32_t final_fp32;              /* last sub-iteration in FP32 storage */
    int32_t acc_bits;                /* XPR_ACC_U32 / XPR_ACC_U64 / XPR_ACC_F32 */
    int32_t jitter;                  /* 0: no origin jitter */
    int64_t jitter_seed;
    int64_t slab_height;             /* used by the streaming fallback */
    const float *projections;        /* host, [projections][dheight][dwidth], raw (as read from TIFF) */
    float *volume_out;               /* host, [vheight][vwidth][vwidth] float32 */
} xpr_job;

typedef struct {
    double submit_ms, start_ms, end_ms;   /* steady clock, ms */
    double device_ms;                     /* sum of sub-iteration device times */
    uint64_t device_bytes;                /* memory reserved against the budget */
    uint64_t voxel_views;                 /* projector work (both projectors) */
    int32_t streamed;                     /* 1: ran with the slab-streaming fallback */
    int32_t status;                       /* XPR_OK or the job's error code */
} xpr_job_stats;

int xpr_server_create(xpr_server **out, uint64_t budget_bytes, int max_concurrent, int device);
/* The job description is copied; its host buffers must stay valid until
 * xpr_server_wait returns for that job. */
int xpr_server_submit(xpr_server *sv, const xpr_job *job, uint64_t *job_id);
int xpr_server_wait(xpr_server *sv, uint64_t job_id, xpr_job_stats *stats);
/* Device memory a job needs when resident (the admission figure). */
int xpr_job_device_bytes(const xpr_job *job, uint64_t *bytes);
int xpr_server_destroy(xpr_server *sv);   /* waits for every job */

#ifdef __cplusplus
}
#endif
#endif
