/*
 * libdd — 3D distance-driven projectors (cone-flat, cone-arc, spiral), C ABI.
 * Same conventions as dd.h.
 *
 * Reference interface replaced: deeprecon/include/dis3d/ConeFlatDisCUDA3d.h,
 * ConeArcDisCUDA3d.h and SpiralDisCUDA3d.h (raw-pointer launchers called from
 * deeprecon/interface/RadonTorch.cpp:572-1267).
 */
#ifndef DD3D_H_
#define DD3D_H_

#include "dd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 3D geometry: the 17-entry Options tensor of projector3d
 * [W, H, Dz, cols, rows, dx, dy, dz, col, row, sid, sdd, sx, sy, sz,
 *  col_shift, row_shift] (projector.py:286).  The spiral projector's
 * 16-entry tensor (projector.py:460) has no sz; pass shift_z = 0.
 * For DD_CONE_ARC and DD_SPIRAL `col_size` and `col_shift` are radians. */
typedef struct {
  int width;
  int height;
  int depth;
  int num_col;
  int num_row;
  float pix_x;
  float pix_y;
  float pix_z;
  float col_size;
  float row_size;
  float iso_source;
  float source_detector;
  float shift_x;
  float shift_y;
  float shift_z;
  float col_shift;
  float row_shift;
} dd_geom3d;

/* Workspace bytes for one call (0 for an invalid geometry).  Geometries
 * whose per-view detector block (rows + 2) * cols * padded batch reaches
 * 2^29 floats are rejected (0 here, DD_BAD_ARG from dd3d_run): the
 * sinogram -> volume gather addresses a view's block with 32-bit offsets.
 * The size depends on dd3d_set_compensated() at the time of the query. */
size_t dd3d_workspace_size(int scan, int op, const dd_geom3d* g, int batch,
                           int views);

/* Runs one 3D op on `stream`.
 *   input  : volume [batch, Dz, H, W] for DD_PROJ / DD_BPROJ_T / DD_BPROJ_TW,
 *            otherwise sinogram [batch, views, rows, cols]; float32, device.
 *   output : the other shape, fully overwritten.
 *   angles : [views] float32, device.
 *   source_pos : DD_SPIRAL only, [4, views] float32, device: rows are the
 *            axial source position, radius shift, angle shift and axial
 *            focal-spot shift (RadonTorch.cpp:1040-1043); NULL otherwise.
 * Replaces ProjConeFlatDisCUDA3d & co. (deeprecon/include/dis3d/). */
dd_status dd3d_run(int scan, int op, const dd_geom3d* g, const float* input,
                   float* output, const float* angles, const float* source_pos,
                   int batch, int views, void* ws, size_t ws_bytes,
                   void* stream);

/* Running sums over detector rows of the sinogram -> volume ops:
 *   1: compensated FP32 pairs (hi + lo, plus each row's own value), so every
 *      slice's sum is rounded relative to the slice;
 *   0: rounded to FP32 (a third of the workspace; error relative to the
 *      whole running sum, i.e. growing with the number of detector rows);
 *   2 (default, and any other value): compensated when the instance tile is
 *      1 or 2 (no extra load per evaluation), rounded at tile 4 (batches of
 *      3 or more), where the compensated gather needs a third 16-byte load.
 * Process-wide; affects calls issued afterwards and the workspace they need.
 * No counterpart in the reference, whose adjoint walks rows and slices with
 * atomics (ConeFlatDisCUDA3d.cu:284-536). */
void dd3d_set_compensated(int mode);
/* Whether the sinogram -> volume gather stages each view's row running sums
 * for its 16 x 16 pixel tile in LDS: 0 never, 1 always, 2 (default)
 * automatic — with compensated sums or batches of 1-2 (cone-flat 256^3 x 2:
 * adjoint 20.0 -> 15.6 ms), not for rounded sums at batches of 3+ (helical
 * 512x512x64 x 4: 94 ms unstaged, 105 ms staged).  The same bits either
 * way; a view whose block exceeds 32 KB of LDS is read from global memory.
 * No counterpart in the reference. */
void dd3d_set_adjoint_staging(int mode);
/* Dispatch order of the volume -> sinogram gather: 0 views first (two
 * detector-row chunks per workgroup), 1 detector columns first (one chunk),
 * 2 (default) automatic: columns first when one instance tile's volume copy
 * exceeds 192 MB (helical 512x512x64 x 4, 268 MB: forward 75.4 -> 65.7 ms),
 * views first otherwise (cone-flat 256^3 x 2, 134 MB: 17.7 ms views first,
 * 19.5 ms columns first).  The same bits either way.  No
 * counterpart in the reference. */
void dd3d_set_forward_order(int mode);

#ifdef __cplusplus
}
#endif

#endif /* DD3D_H_ */
