/*
 * libdd — MI355X (gfx950) distance-driven CT projectors, C ABI.
 *
 * Plain pointers and sizes only: device pointers for data, host pointers for
 * geometry structs, a hipStream_t passed as void*.  Every entry point is
 * stream-ordered and returns a dd_status; none synchronises the host and none
 * allocates memory the caller can see.  Scratch space is passed in
 * (dd*_workspace_size + the `ws` argument), except for dd2d_run_noalloc,
 * which uses a fixed per-device buffer inside the library.
 *
 * Reference interface replaced: the raw-pointer C level of deeprecon,
 *   deeprecon/include/dis2d/{FanArc,FanFlat,Para}DisCUDA2d.h and
 *   deeprecon/include/dis3d/{ConeFlat,ConeArc,Spiral}DisCUDA3d.h,
 * called from deeprecon/interface/RadonTorch.cpp.  The geometry structs carry
 * exactly the fields of the reference's Options tensors
 * (deeprecon/deeprecon/torch/projector/projector.py:113, :286, :460).
 */
#ifndef DD_H_
#define DD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  DD_OK = 0,
  DD_BAD_ARG = 1,      /* null pointer, non-positive size, unknown op */
  DD_UNSUPPORTED = 2,  /* op/geometry pair the reference does not define */
  DD_LAUNCH_FAIL = 3,  /* a kernel launch or HIP call failed */
  DD_NO_WORKSPACE = 4, /* workspace too small or absent */
  DD_NO_DEVICE = 5,    /* no HIP device visible */
  DD_TOO_LARGE = 6     /* the problem needs more scratch than dd2d_run_noalloc's
                          fixed buffer (DD_NOALLOC_SCRATCH_BYTES) holds */
} dd_status;

/* Scan geometries (the reference's scan_type, projector.py:104,273,452). */
typedef enum {
  DD_PARA = 0,     /* 2D parallel beam */
  DD_FAN_FLAT = 1, /* 2D fan beam, flat detector */
  DD_FAN_ARC = 2,  /* 2D fan beam, equiangular (arc) detector */
  DD_CONE_FLAT = 3,
  DD_CONE_ARC = 4,
  DD_SPIRAL = 5
} dd_scan;

/* The six operators of every geometry (RadonTorch.cpp:1376-1418). */
typedef enum {
  DD_PROJ = 0,        /* A: image -> sinogram (Proj*) */
  DD_PROJ_T = 1,      /* A^T, the matched adjoint (ProjTrans*) */
  DD_BPROJ = 2,       /* FBP-style backprojection (BackProj*) */
  DD_BPROJ_T = 3,     /* its transpose (BackProjTrans*) */
  DD_BPROJ_W = 4,     /* distance-weighted backprojection (BackProjWeighted*) */
  DD_BPROJ_TW = 5     /* its transpose (BackProjTransWeighted*) */
} dd_op;

/* 2D geometry: the 11-entry Options tensor of projector2d
 * [W, H, D, dx, dy, det, sid, sdd, shift_x, shift_y, bin_shift]
 * (projector.py:113).  For DD_FAN_ARC `det` and `bin_shift` are angles in
 * radians (projector.py:110-112). */
typedef struct {
  int width;
  int height;
  int num_det;
  float pix_x;
  float pix_y;
  float det;
  float iso_source;
  float source_detector;
  float shift_x;
  float shift_y;
  float bin_shift;
} dd_geom2d;

/* Bytes of scratch device memory dd2d_run needs for this call. */
size_t dd2d_workspace_size(int scan, int op, const dd_geom2d* g, int batch,
                           int views);

/* Runs one 2D op on `stream`.
 *   input  : image [batch, H, W] for DD_PROJ / DD_BPROJ_T / DD_BPROJ_TW,
 *            otherwise sinogram [batch, views, D]; float32, device.
 *   output : the other shape; fully overwritten (no pre-zeroing needed).
 *   angles : [views] float32, device.
 *   ws     : at least dd2d_workspace_size() bytes of device memory.
 * Replaces ProjFanArcDisCUDA2d & co. (deeprecon/include/dis2d/FanArcDisCUDA2d.h etc.). */
dd_status dd2d_run(int scan, int op, const dd_geom2d* g, const float* input,
                   float* output, const float* angles, int batch, int views,
                   void* ws, size_t ws_bytes, void* stream);

/* dd2d_run without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation).  No counterpart in the reference.  Scratch comes
 * from a fixed buffer inside the library, one per device, resident until the
 * process ends; the call allocates and frees nothing.  A problem whose
 * dd2d_workspace_size() exceeds DD_NOALLOC_SCRATCH_BYTES is rejected with
 * DD_TOO_LARGE (a 512x512 fan-arc slice with 720 views and 736 bins needs
 * 27 MB at batch 1).  Results equal dd2d_run's bit for bit.  Every call on
 * one device shares that device's buffer: calls that overlap on one device
 * (two streams, two host threads) race and must be ordered by the caller. */
#define DD_NOALLOC_SCRATCH_BYTES (64u << 20)
dd_status dd2d_run_noalloc(int scan, int op, const dd_geom2d* g, const float* input,
                           float* output, const float* angles, int batch, int views,
                           void* stream);

/* View-resolved sinogram -> image op (DD_PROJ_T, DD_BPROJ or DD_BPROJ_W):
 * output [batch][views][H][W] holds each view's contribution alone, equal
 * bit for bit to dd2d_run of that op on the one view.  Replaces the per-view
 * loop of single-view backprojections in deeprecon's bprj_sv_fun
 * (deeprecon/deeprecon/torch/model/vvbp_tensor_net.py:11-24), one launch for
 * all views.  views <= 65535. */
size_t dd2d_views_workspace_size(int scan, int op, const dd_geom2d* g, int batch,
                                 int views);
dd_status dd2d_run_views(int scan, int op, const dd_geom2d* g, const float* input,
                         float* output, const float* angles, int batch, int views,
                         void* ws, size_t ws_bytes, void* stream);

/* System matrix with the FP weights as COO entries (row = v*D + bin,
 * column = row*W + col), built on the device in two passes: count the
 * entries of every (view, bin) into `counts` [views*D], let the caller turn
 * them into exclusive offsets, then fill.  Replaces GetSysMatrix*2dKernel
 * (deeprecon/include/dis2d/FanArcDisCPU2d.h:21-23, FanFlatDisCPU2d.h:21-23,
 * ParaDisCPU2d.h:16-17), which the reference runs on the CPU only. */
size_t dd2d_sysmatrix_workspace_size(int scan, const dd_geom2d* g, int views);
dd_status dd2d_sysmatrix_count(int scan, const dd_geom2d* g,
                               const float* angles, int views, int64_t* counts,
                               void* ws, size_t ws_bytes, void* stream);
dd_status dd2d_sysmatrix_fill(int scan, const dd_geom2d* g,
                              const float* angles, int views,
                              const int64_t* offsets, int64_t* rows,
                              int64_t* cols, float* vals, void* ws,
                              size_t ws_bytes, void* stream);

/* Planned use: the per-view geometry tables (bin edges on the common axis,
 * FP weights, per-line pixel-edge maps) live in caller memory and are built
 * once by dd2d_geometry(); dd2d_run_planned() then only moves data.  Used by
 * the batched path and by graph-captured steps, where the same geometry is
 * replayed many times.  dd2d_run() is dd2d_geometry() + dd2d_run_planned()
 * on one workspace. */
size_t dd2d_geometry_size(int scan, const dd_geom2d* g, int views);
dd_status dd2d_geometry(int scan, const dd_geom2d* g, const float* angles,
                        int views, void* geo, size_t geo_bytes, void* stream);
/* Kernel variants of the sinogram -> image gathers (all exact and equal):
 * each pixel finds its bins on its own; a 16x16 / 8x8 pixel tile stages its
 * lines' bin-edge positions per view in LDS; or one thread walks the bins of
 * a 2x2 pixel block once for its two lines of two pixels. */
#define DD_ADJ_PER_PIXEL 0
#define DD_ADJ_TILE16 1
#define DD_ADJ_TILE8 2
#define DD_ADJ_QUAD 3
/* Uneven batches only: the 4 x 4 pixel blocks, each line's bins read as one
 * window of loads instead of one bin at a time. */
#define DD_ADJ_BLOCK_WINDOW 4
/* The 4 x 4 pixel blocks through per-view running integrals of the sinogram
 * along the detector (FP64 knots at the bin edges): each pixel's value from
 * a view is the difference of the running integral at its two edges, found
 * by one merge walk per 4-pixel run.  PROJ_T and BPROJ only (BPROJ_W, whose
 * weight varies per pixel, keeps a walk).  In uneven batches; and in
 * dd2d_run / planned runs at instance tile 1, where the views are split over
 * workgroups (partial images added in order): a 512^2 x 720-view slice's
 * ProjTrans 0.3 ms instead of the walk's 1.06 ms.  Same operator, different
 * rounding: not bit-equal to the other variants (5e-8 from FP64 against the
 * walks' 2-3e-7), so the automatic choice keeps the walks and this variant
 * is selected explicitly or by the autotuner. */
#define DD_ADJ_BLOCK_PREFIX 5
/* Flag OR-ed into a variant: the image -> sinogram ops PROJ and BPROJ_T
 * integrate each bin over each line as a difference of the line's running
 * pixel sums (accumulated in FP64, stored as FP32 hi/lo pairs) instead of
 * walking the pixels the bin covers.  In uneven batches, and in dd2d_run /
 * planned runs at instance tile 1 (dd2d_set_adjoint_variant or a plan's
 * variant; ignored at larger tiles).  Same operator, different rounding (not
 * bit-equal to the walk; a non-finite pixel then reaches every later bin of
 * its line).  BPROJ_TW keeps the walk. */
#define DD_FWD_PREFIX 0x100

/* `tile` = instances sharing one geometry evaluation (1, 2, 4, 8 or 16;
 * 0 = automatic); it must be the same in both calls.  `variant` selects the
 * sinogram -> image kernel (DD_ADJ_*; -1 = process default). */
size_t dd2d_planned_workspace_size(int scan, int op, const dd_geom2d* g,
                                   int batch, int views, int tile);
dd_status dd2d_run_planned(int scan, int op, const dd_geom2d* g,
                           const void* geo, const float* input, float* output,
                           int batch, int views, int tile, int variant,
                           void* ws, size_t ws_bytes, void* stream);

/* Output epilogue of a planned op, for fused update steps such as the
 * data-consistency step x - w * FBP(A x - y) (PrideDiff's fidelity module,
 * model/pridediff.py:56-68; no counterpart at the reference's C level):
 *   output = alpha * s * op(input) + beta * c,
 * s = *scale (a device scalar, e.g. the learned step weight w) or 1 when
 * scale is NULL, c a device array shaped like the output (NULL: no term;
 * it must not alias the output).  Both are read on the device when the op
 * runs, so a captured graph follows later changes of *scale and of c. */
typedef struct {
  const float* c;
  const float* scale;
  float alpha;
  float beta;
} dd_epilogue;

/* Workgroup shape of the planned gathers, the (views x detector bins x image
 * lines) tile a workgroup covers; 0 in a field keeps the built-in choice.
 *   fwd_waves : image -> sinogram: 63-bin wavefronts per workgroup (1, 2, 4);
 *   fwd_views : image -> sinogram: views per workgroup, one after the other
 *               (1 - 8); the lines are one of the 8 per-XCD line chunks;
 *   pix_width : per-pixel sinogram -> image variant: pixels per workgroup
 *               row (16, 32 or 64; 256 threads, so 16, 8 or 4 rows).
 * The shape never changes a result bit; it is what the autotuner sweeps. */
typedef struct {
  int fwd_waves;
  int fwd_views;
  int pix_width;
} dd_shape;

/* dd2d_run_planned with a workgroup shape (NULL: built-in) and an epilogue
 * applied where the op writes its output (in the forward's chunk
 * reduction, or in the backprojection gather), so the residual A x - y and
 * the update x - w BP(r) cost no extra pass (NULL: none).  With both NULL it
 * is dd2d_run_planned. */
dd_status dd2d_run_planned_ep(int scan, int op, const dd_geom2d* g,
                              const void* geo, const float* input, float* output,
                              int batch, int views, int tile, int variant,
                              const dd_shape* shape, const dd_epilogue* ep, void* ws,
                              size_t ws_bytes, void* stream);

/* Batched projection of many instances of uneven geometry.
 *
 * Each item describes one instance (or `batch` instances sharing a
 * geometry) and where its image and sinogram live inside one packed image
 * buffer and one packed sinogram buffer (offsets in floats; image
 * [batch, H, W], sinogram [batch, views, D]).  dd2d_batch_create is host
 * only; dd2d_batch_setup builds every instance's geometry in the device
 * memory it is given (dd2d_batch_device_bytes) and waits for it once;
 * dd2d_batch_run then runs one op for all instances as two launches
 * (layout, gather; plus the chunk reduction for image -> sinogram ops),
 * capturable into a HIP graph.  The result for
 * each instance equals dd2d_run on that instance bit for bit.  No
 * counterpart in the reference, which launches once per instance and op
 * (deeprecon/cuda/dis2d/FanArcDisCUDA2d.cu:758-1016). */
typedef struct {
  int scan;               /* DD_PARA, DD_FAN_FLAT or DD_FAN_ARC */
  dd_geom2d geom;
  int views;
  int batch;              /* instances sharing this geometry */
  const float* angles;    /* [views], device */
  long long image_offset; /* floats into the packed image buffer */
  long long sino_offset;  /* floats into the packed sinogram buffer */
} dd2d_batch_item;

typedef struct dd2d_batch dd2d_batch;

dd2d_batch* dd2d_batch_create(const dd2d_batch_item* items, int n);
size_t dd2d_batch_device_bytes(const dd2d_batch* b);
/* Grid size in workgroups, each striding over the work items (0 = one
 * workgroup per work item). */
void dd2d_batch_set_grid(dd2d_batch* b, int workgroups);
/* Sinogram -> image kernel of the batch: DD_ADJ_PER_PIXEL, DD_ADJ_TILE16,
 * DD_ADJ_QUAD (in a batch: one thread per 4x4 pixel block, four lines of
 * four pixels; the default), DD_ADJ_BLOCK_WINDOW or DD_ADJ_BLOCK_PREFIX,
 * optionally OR-ed with DD_FWD_PREFIX for the image -> sinogram ops. */
void dd2d_batch_set_variant(dd2d_batch* b, int variant);
dd_status dd2d_batch_setup(dd2d_batch* b, void* device_mem, size_t bytes,
                           void* stream);
dd_status dd2d_batch_run(const dd2d_batch* b, int op, const float* input,
                         float* output, void* stream);
void dd2d_batch_destroy(dd2d_batch* b);

/* Kernel classes timed by the optional profiler. */
#define DD_KERNEL_IMAGE_TO_SINO 0 /* gathers writing sinograms (FP side) */
#define DD_KERNEL_SINO_TO_IMAGE 1 /* gathers writing images (adjoint side) */
#define DD_KERNEL_CLASSES 2

/* Brackets every launch of the gather kernels with HIP events on the launch
 * stream (never inside stream capture).  dd_profile_read waits for the
 * recorded events, returns the summed milliseconds and launch count of one
 * class since the last read, and resets that class. */
void dd_profile_enable(int on);
int dd_profile_read(int kernel_class, double* total_ms, long long* launches);

/* Default sinogram -> image kernel variant of dd2d_run and of planned runs
 * given variant -1 (DD_ADJ_PER_PIXEL .. DD_ADJ_QUAD or DD_ADJ_BLOCK_PREFIX,
 * optionally | DD_FWD_PREFIX; -1 = automatic: DD_ADJ_QUAD for up to 4
 * instances per launch when that still gives 4 workgroups per compute unit,
 * DD_ADJ_PER_PIXEL otherwise, walks in both directions).  The running sums
 * apply at instance tile 1 only; above it DD_ADJ_BLOCK_PREFIX falls back to
 * the automatic choice and DD_FWD_PREFIX is ignored.  dd2d_workspace_size
 * follows the current setting. */
void dd2d_set_adjoint_variant(int variant);

/* Instance tile used per launch (1, 2, 4, 8 or 16 instances share one
 * geometry evaluation); 0 restores the built-in choice.  Process-wide. */
void dd2d_set_instance_tile(int tile);

/* Library version string. */
const char* dd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DD_H_ */
