// 3D distance-driven projectors (cone-flat, cone-arc, spiral) for gfx950.
//
// Behaviour followed: deeprecon/cuda/dis3d/ConeFlatDisCUDA3d.cu,
// ConeArcDisCUDA3d.cu and SpiralDisCUDA3d.cu (op table RadonTorch.cpp:572-1267).
// Separable footprint as in the reference: per view and image line, the
// column overlap on the 2D common axis times, per detector column, the z
// overlap on the axis plane, with detector rows magnified at the column
// centre (ConeFlatDisCUDA3d.cu:92-93) and voxel slices at the line centre
// (:120-121).  The spiral adds per-view source z, radius, angle and axial
// focal-spot shifts (SpiralDisCUDA3d.cu:72-83).
//
// Launches per call (stream-ordered, no host sync):
//   k_setup3d            per view: FP64 column edges, magnifications, row-z
//                        law, per-line x and z maps;
//   k_interleave_volume  volume -> [tile][line][slice][pixel][b] copies (rows
//                        and bottom-up columns) for image -> sinogram ops;
//   k_bins3d             image -> sinogram: one thread per detector column and
//                        chunk of rows, gathering over the lines its ray band
//                        meets in z;
//   k_prefix_sino3d      sinogram -> per-column running sums over detector
//                        rows, [view][row][col][b] (x FP weight);
//   k_voxels3d           sinogram -> image: one thread per pixel and chunk of
//                        slices, gathering over views; the row/slice overlap
//                        sums come from differences of the running sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>

#include "dd3d.h"
#include "dd_common.h"
#include "dd_profile.h"

namespace dd {
namespace {

constexpr int kThreads3 = 256;
// k_bins3d workgroups: 64 detector columns x one or two row chunks (one wave
// per chunk), in one of two dispatch orders (dd3d_set_forward_order):
//  * views first: grid (views, column blocks, row-chunk pairs x instance
//    tiles), two chunks per workgroup.  Consecutive workgroups are
//    consecutive views of one block, so the blocks in flight sweep the whole
//    volume copy; best while that copy stays in the 256 MB Infinity Cache
//    (cone-flat 256^3 x 2: 17.5 ms; with one / four chunks per workgroup
//    17.3 / 17.7 ms; helical 512x512x64 x 4: 77.4 / 77.6 against 75.4 at
//    two).
//  * columns first: grid (column blocks, views, row chunks x instance tiles),
//    one chunk per workgroup.  The workgroups in flight are the column blocks
//    of a few hundred consecutive views for one row chunk: a band of slices,
//    and on each XCD a few column wedges of it.  Best once the copy outgrows
//    the cache (the helical copy, 268 MB: 75.4 -> 65.7 ms; cone-flat
//    19.5 ms).
constexpr int kColsPerBlock = 64;
constexpr int kFwdThreads = 2 * kColsPerBlock;
constexpr int kChunksViewsFirst = 2;
constexpr int kChunksColsFirst = 1;

// Dispatch order of k_bins3d: 0 views first, 1 columns first, 2 automatic
// (columns first when one instance tile's volume copy exceeds 192 MB, three
// quarters of the Infinity Cache, and the views fit a grid dimension).
std::atomic<int> g_fwd_order{2};
constexpr int kTile3 = 16;         // k_voxels3d: 16 x 16 pixels x one slice chunk

// Running sums of the sinogram -> volume ops (dd3d_set_compensated):
// 0 rounded to FP32, 1 compensated (hi/lo pairs), 2 automatic: compensated
// when it adds no vector-memory instruction to the gather (instance tiles of
// 1 and 2: its hi and (lo, row) tiles take the two loads the rounded form's
// rows r and r + 1 take), rounded at tile 4, where it costs a third load.
std::atomic<int> g_compensated{2};

bool use_compensated(int bt) {
  const int m = g_compensated.load();
  return m == 1 || (m == 2 && bt <= 2);
}

// Whether k_voxels3d stages each view's row running sums for its pixel tile
// in LDS (dd3d_set_adjoint_staging; the same bits either way): 0 never, 1
// always, 2 automatic: with compensated sums or instance tiles of 1-2 (16
// slices per thread), where it measured faster; at tile 4 with rounded sums
// (8 slices per thread, less reuse per staged view) global reads win.
std::atomic<int> g_stage3{2};

bool use_stage3(int bt, bool comp) {
  const int m = g_stage3.load();
  return m == 1 || (m == 2 && (comp || bt <= 2));
}

enum WMode3 { W3_FP = 0, W3_BP = 1, W3_BPW = 2 };

struct Geo3 {
  int W, H, Dz, C, R, V, B, Bp, BT, Lmax, scan;
  float dx, dy, dz, col, row, sid, sdd, shx, shy, shz, csh, rsh;
};

struct ViewGeo3 {
  double sx, sy, sz;   // source position
  double s, c;         // sin / cos of the source angle
  double theta_ref;    // arc: detector angle base + column shift
  double colp, cshp;   // column pitch and shift (radians, or virtual mm)
  double zdet0, rowp;  // detector row edge r at zdet0 + r * rowp
  double sid_eff;      // source-isocentre distance of this view
  float sxf, syf, kbp; // float copies; BackProj constant
  float zlo, zhi;      // world z range the view's rays can reach in the volume
  float cref, sref;    // cos, sin of theta_ref in float (first-column guesses)
  int xaxis, flip;
};

struct LineGeo3 {
  double a0, inv_p;    // pixel edges on the axis: a0 + c / inv_p
  double z0, inv_pz;   // slice edges on the axis plane: z0 + s / inv_pz
  double pz;           // slice interval on the axis plane
  float p, pzf;        // pixel / slice intervals in float
};

// Row edge r of column k on the axis plane.  Explicit fma so every kernel
// evaluates it to the same bits.
__device__ __forceinline__ double row_z(const ViewGeo3& vg, double mag, int r) {
  const double zdet = fma(static_cast<double>(r), vg.rowp, vg.zdet0);
  return fma(mag, zdet - vg.sz, vg.sz);
}

// FP weight of a detector cell: perpendicular voxel size / (column width on
// the axis * row height on the axis plane * cosine of the cell-centre ray),
// the cosine taken at the absolute row-centre z as the reference does
// (ConeFlatDisCUDA3d.cu:99-100, SpiralDisCUDA3d.cu:111).
__device__ __forceinline__ float fp_weight3(const Geo3& g, const ViewGeo3& vg,
                                            double e0, double e1, double z0,
                                            double z1) {
  const double s_ax = vg.xaxis ? vg.sx : vg.sy;
  const double s_pp = vg.xaxis ? vg.sy : vg.sx;
  const double a = 0.5 * (e0 + e1) - s_ax, b = 0.5 * (z0 + z1);
  const double cs = fabs(s_pp) / sqrt(a * a + b * b + s_pp * s_pp);
  const double perp = vg.xaxis ? g.dy : g.dx;
  return static_cast<float>(perp / ((e1 - e0) * (z1 - z0) * cs));
}

__device__ __forceinline__ void pixel_centre3(const Geo3& g, int xaxis, int l,
                                              int i, float* xc, float* yc) {
  if (xaxis) {
    *xc = (i - g.W * 0.5f + 0.5f) * g.dx + g.shx;
    *yc = (g.H * 0.5f - l - 0.5f) * g.dy + g.shy;
  } else {
    *xc = (l - g.W * 0.5f + 0.5f) * g.dx + g.shx;
    *yc = (i - g.H * 0.5f + 0.5f) * g.dy + g.shy;
  }
}

// FBP distance weight: flat (r/(r-d))^2 (ConeFlatDisCUDA3d.cu:16-19), arc and
// spiral 1/|S-P|^2 (ConeArcDisCUDA3d.cu:24-26).
__device__ __forceinline__ float fbp_weight3(const Geo3& g, const ViewGeo3& vg,
                                             float xc, float yc) {
  if (g.scan == DD_CONE_FLAT) {
    const float r = g.sid;
    const float d = (vg.sxf * xc + vg.syf * yc) / r;
    return (r * r) / ((r - d) * (r - d));
  }
  const float ddx = vg.sxf - xc, ddy = vg.syf - yc;
  return 1.0f / (ddx * ddx + ddy * ddy);
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads3) void k_setup3d(
    Geo3 g, const float* __restrict__ angles, const float* __restrict__ spos,
    ViewGeo3* __restrict__ views, double* __restrict__ cedge,
    double* __restrict__ mag, LineGeo3* __restrict__ lines) {
  const int v = blockIdx.x;
  const int C = g.C;
  double sid = g.sid, sdd = g.sdd, colp = g.col, csh = g.csh, sz = 0.0, zoff = 0.0;
  float th_src_f = angles[v];
  const double th_det = static_cast<double>(angles[v]);
  if (g.scan == DD_SPIRAL) {
    const double pa = spos[v], dr = spos[g.V + v], da = spos[2 * g.V + v],
                 dzs = spos[3 * g.V + v];
    sid = static_cast<double>(g.sid) + dr;
    sdd = static_cast<double>(g.sdd) + dr;
    colp = g.col * static_cast<double>(g.sdd) / sdd;
    csh = g.csh * static_cast<double>(g.sdd) / sdd - da * (sdd - sid) / sdd;
    th_src_f = angles[v] + spos[2 * g.V + v];
    sz = pa + dzs;
    zoff = pa;
  }
  int xaxis, flip;
  choose_axis(th_src_f, &xaxis, &flip);
  const double th_src = static_cast<double>(th_src_f);
  const double s = sin(th_src), c = cos(th_src);
  const double sx = -s * sid, sy = c * sid;
  const double vcol = static_cast<double>(g.sid) / g.sdd * g.col;
  const double vcsh = static_cast<double>(g.sid) / g.sdd * g.csh;
  const bool flat = (g.scan == DD_CONE_FLAT);
  ViewGeo3 vg;
  {
    vg.sx = sx;
    vg.sy = sy;
    vg.sz = sz;
    vg.s = s;
    vg.c = c;
    vg.theta_ref = th_det + csh;
    vg.cref = static_cast<float>(cos(th_det + csh));
    vg.sref = static_cast<float>(sin(th_det + csh));
    vg.colp = flat ? vcol : colp;
    vg.cshp = flat ? vcsh : csh;
    const double rs = flat ? static_cast<double>(g.sid) / g.sdd : 1.0;
    vg.zdet0 = flat ? (-g.R / 2.0 * g.row + g.rsh) * rs : (-g.R / 2.0 * g.row + g.rsh + zoff);
    vg.rowp = g.row * rs;
    vg.sid_eff = sid;
    vg.sxf = static_cast<float>(sx);
    vg.syf = static_cast<float>(sy);
    vg.kbp = static_cast<float>(flat ? vcol : sid * colp);
    vg.xaxis = xaxis;
    vg.flip = flip;
  }
  double mag_lo = 1e300, mag_hi = -1e300, m_lo = 1e300, m_hi = -1e300;
  // column edges on the common axis, walk order (ConeFlatDisCUDA3d.cu:78-91,
  // ConeArcDisCUDA3d.cu:82-93)
  for (int k = threadIdx.x; k <= C; k += blockDim.x) {
    const double off = flip ? (C / 2.0 - k) : (k - C / 2.0);
    double e;
    if (flat) {
      const double u = off * vcol + vcsh, px = u * c, py = u * s;
      e = xaxis ? (sx * py - sy * px) / (py - sy) : (sy * px - sx * py) / (px - sx);
    } else {
      const double t = tan(off * colp + csh + th_det);
      e = xaxis ? sx + sy * t : sy + sx / t;
    }
    cedge[static_cast<size_t>(v) * (C + 1) + k] = e;
  }
  // magnification of the detector rows at each column centre (:92-93)
  for (int k = threadIdx.x; k < C; k += blockDim.x) {
    const double offc = flip ? (C / 2.0 - k - 0.5) : (k + 0.5 - C / 2.0);
    double dcx, dcy;
    if (flat) {
      const double u = offc * vcol + vcsh;
      dcx = u * c;
      dcy = u * s;
    } else {
      const double gc = offc * colp + csh + th_det;
      dcx = sin(gc) * sdd + sx;
      dcy = -cos(gc) * sdd + sy;
    }
    const double mk = xaxis ? sy / (sy - dcy) : sx / (sx - dcx);
    mag[static_cast<size_t>(v) * C + k] = mk;
    mag_lo = fmin(mag_lo, mk);
    mag_hi = fmax(mag_hi, mk);
  }
  // per line: pixel edges on the axis and slice edges on the axis plane
  const int L = xaxis ? g.H : g.W;
  const double z0v = -g.Dz / 2.0 * g.dz + (g.scan == DD_SPIRAL ? 0.0 : g.shz);
  for (int l = threadIdx.x; l < L; l += blockDim.x) {
    double a0, p, m;
    if (xaxis) {
      const double yl = (g.H / 2.0 - l - 0.5) * g.dy + g.shy;
      const double x0 = -g.W / 2.0 * g.dx + g.shx, x1 = g.W / 2.0 * g.dx + g.shx;
      a0 = (sx * yl - sy * x0) / (yl - sy);
      p = ((sx * yl - sy * x1) / (yl - sy) - a0) / g.W;
      m = sy / (sy - yl);
    } else {
      const double xl = (l - g.W / 2.0 + 0.5) * g.dx + g.shx;
      const double y0 = -g.H / 2.0 * g.dy + g.shy, y1 = g.H / 2.0 * g.dy + g.shy;
      a0 = (sy * xl - sx * y0) / (xl - sx);
      p = ((sy * xl - sx * y1) / (xl - sx) - a0) / g.H;
      m = sx / (sx - xl);
    }
    LineGeo3 lg;
    lg.a0 = a0;
    lg.inv_p = 1.0 / p;
    lg.z0 = m * (z0v - sz) + sz;
    lg.pz = m * g.dz;
    lg.inv_pz = 1.0 / lg.pz;
    lg.p = static_cast<float>(p);
    lg.pzf = static_cast<float>(lg.pz);
    lines[static_cast<size_t>(v) * g.Lmax + l] = lg;
    m_lo = fmin(m_lo, m);
    m_hi = fmax(m_hi, m);
  }
  // World z range of the view: a slice at z meets row band zdet on line l
  // and column k only if m_l (z - sz) lies within mag_k (zdet - sz), so z - sz
  // is bounded by (mag_k / m_l) times the detector's z extent, over all k, l.
  __shared__ double red[4][kThreads3];
  red[0][threadIdx.x] = mag_lo;
  red[1][threadIdx.x] = -mag_hi;
  red[2][threadIdx.x] = m_lo;
  red[3][threadIdx.x] = -m_hi;
  __syncthreads();
  for (int w = kThreads3 / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int q = 0; q < 4; ++q)
        red[q][threadIdx.x] = fmin(red[q][threadIdx.x], red[q][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rho_lo = red[0][0] / -red[3][0], rho_hi = -red[1][0] / red[2][0];
    const double d0 = vg.zdet0 - sz, d1 = fma(static_cast<double>(g.R), vg.rowp, vg.zdet0) - sz;
    const double c0 = rho_lo * d0, c1 = rho_lo * d1, c2 = rho_hi * d0, c3 = rho_hi * d1;
    const double lo = fmin(fmin(c0, c1), fmin(c2, c3)), hi = fmax(fmax(c0, c1), fmax(c2, c3));
    // one slice of margin on each side, plus a relative guard
    const double pad = g.dz + 1e-6 * (fabs(lo) + fabs(hi) + fabs(sz));
    vg.zlo = static_cast<float>(sz + lo - pad);
    vg.zhi = static_cast<float>(sz + hi + pad);
    views[v] = vg;
  }
}

// ---------------------------------------------------------------------------
// layout prologues
// ---------------------------------------------------------------------------
// x [B][Dz][H][W] -> xrow [tile][H][Dz][W][BT], xcol [tile][W][Dz][H][BT]
// (bottom-up), tile = BT consecutive instances: a pixel's BT values and its
// neighbour's are contiguous, whatever the padded batch.  One workgroup per
// 16 x 16 pixels of one slice: the column copy runs down the image's
// columns, so the tile passes through LDS and is written transposed, and a
// wave's lanes write contiguous pixels of both copies.
constexpr int kItile3 = 16;

template <int BT>
__global__ __launch_bounds__(kItile3 * kItile3) void k_interleave_volume(
    Geo3 g, const float* __restrict__ x, float* __restrict__ xrow,
    float* __restrict__ xcol) {
  // pixel (r, c) of the tile, value b, at t[r * kRowPitch + c * (BT + 1) + b]
  // (odd strides: neither access order below piles onto one LDS bank)
  constexpr int kRowPitch = kItile3 * (BT + 1) + 1;
  __shared__ float t[kItile3 * kRowPitch];
  const int ntw = (g.W + kItile3 - 1) / kItile3;
  const int tx = static_cast<int>(threadIdx.x) % kItile3;
  const int ty = static_cast<int>(threadIdx.x) / kItile3;
  const int w0 = (blockIdx.x % ntw) * kItile3, h0 = (blockIdx.x / ntw) * kItile3;
  const int s = blockIdx.y;
  const int b0 = blockIdx.z * BT;
  const size_t tile = static_cast<size_t>(blockIdx.z) * g.W * g.H * g.Dz * BT;
  if (h0 + ty < g.H && w0 + tx < g.W) {
    const int h = h0 + ty, w = w0 + tx;
    float v[BT];
#pragma unroll
    for (int b = 0; b < BT; ++b)
      v[b] = (b0 + b < g.B)
                 ? x[((static_cast<size_t>(b0 + b) * g.Dz + s) * g.H + h) * g.W + w]
                 : 0.0f;
    store_tile<BT>(xrow + tile + ((static_cast<size_t>(h) * g.Dz + s) * g.W + w) * BT, v);
#pragma unroll
    for (int b = 0; b < BT; ++b) t[ty * kRowPitch + tx * (BT + 1) + b] = v[b];
  }
  __syncthreads();
  // transposed: consecutive lanes take consecutive rows of one image column
  if (h0 + tx < g.H && w0 + ty < g.W) {
    const int h = h0 + tx, w = w0 + ty;
    float v[BT];
#pragma unroll
    for (int b = 0; b < BT; ++b) v[b] = t[tx * kRowPitch + ty * (BT + 1) + b];
    store_tile<BT>(xcol + tile + ((static_cast<size_t>(w) * g.Dz + s) * g.H + (g.H - 1 - h)) * BT,
                   v);
  }
}

// p [B][V][R][C] -> Q [V][R+2][C][Bp]: per view and column, the running sum
// over detector rows (columns in walk order, optionally times the FP weight
// of each cell), Q[0] = 0, and Q[R+1] = Q[R] so an interpolation at the top
// edge (row R, fraction 0) needs no special case.  Accumulated in FP64.
//
// COMP = false: Q stored rounded to FP32 (one rounding per entry, no drift).
// COMP = true (compensated): Q kept as an unevaluated FP32 pair hi + lo
// (hi = Q rounded, lo = Q - hi rounded, the exact two-sum split of the FP64
// running sum), and beside lo the row's own value x_r, the term the sum
// added at row r (zero at rows R and R + 1).  Planes: hi [V][R+2][C][Bp] in
// `q`, and (lo, x) pairs [V][R+2][C][Bp][2] in `q2`.  The gather then forms
// each slice's sum as differences of hi and of (lo + fraction x) that are no
// larger than the slice's own rows, so its rounding error is relative to
// the slice, not to the running sum (which, R rows deep, is up to R times
// larger).
template <int BT, bool FOLD_W, bool COMP>
__global__ __launch_bounds__(kThreads3) void k_prefix_sino3d(
    Geo3 g, const ViewGeo3* __restrict__ views, const double* __restrict__ cedge,
    const double* __restrict__ mag, const float* __restrict__ p,
    float* __restrict__ q, float* __restrict__ q2) {
  const int nkc = (g.C + kThreads3 - 1) / kThreads3;
  const int k = (blockIdx.x % nkc) * kThreads3 + threadIdx.x;
  const int v = blockIdx.x / nkc;
  const int b0 = blockIdx.z * BT;
  if (k >= g.C) return;
  const ViewGeo3 vg = views[v];
  const int kk = vg.flip ? g.C - 1 - k : k;
  double e1 = 0.0, e2 = 0.0, m = 0.0;
  if (FOLD_W) {
    const double* ce = cedge + static_cast<size_t>(v) * (g.C + 1);
    e1 = ce[k];
    e2 = ce[k + 1];
    m = mag[static_cast<size_t>(v) * g.C + k];
  }
  const size_t rstride = static_cast<size_t>(g.C) * g.Bp;
  float* __restrict__ qv = q + static_cast<size_t>(v) * (g.R + 2) * rstride +
                           static_cast<size_t>(k) * g.Bp + b0;
  // (lo, x) pairs of this tile: the same element offset, doubled
  float* __restrict__ q2v = COMP ? q2 + 2 * (static_cast<size_t>(v) * (g.R + 2) * rstride +
                                             static_cast<size_t>(k) * g.Bp + b0)
                                 : nullptr;
  double acc[BT];
  float val[BT], pair[2 * BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) {
    acc[b] = 0.0;
    val[b] = 0.0f;
  }
  store_tile<BT>(qv, val);
  double zlo = FOLD_W ? row_z(vg, m, 0) : 0.0;
  for (int r = 0; r < g.R; ++r) {
    float w = 1.0f;
    if (FOLD_W) {
      const double zhi = row_z(vg, m, r + 1);
      w = fp_weight3(g, vg, e1, e2, zlo, zhi);
      zlo = zhi;
    }
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const float x = (b0 + b < g.B)
                          ? p[((static_cast<size_t>(b0 + b) * g.V + v) * g.R + r) * g.C + kk]
                          : 0.0f;
      const float xw = x * w;
      if constexpr (COMP) {
        // row r's pair: lo of Q[r] (still in acc) and the row's term
        pair[2 * b] = static_cast<float>(acc[b] - static_cast<double>(val[b]));
        pair[2 * b + 1] = xw;
      }
      acc[b] += static_cast<double>(xw);
      val[b] = static_cast<float>(acc[b]);
    }
    if constexpr (COMP) store_tile<2 * BT>(q2v + 2 * static_cast<size_t>(r) * rstride, pair);
    store_tile<BT>(qv + (r + 1) * rstride, val);
  }
  store_tile<BT>(qv + (g.R + 1) * rstride, val);
  if constexpr (COMP) {
    // rows R and R + 1: lo of Q[R], no row term
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      pair[2 * b] = static_cast<float>(acc[b] - static_cast<double>(val[b]));
      pair[2 * b + 1] = 0.0f;
    }
    store_tile<2 * BT>(q2v + 2 * static_cast<size_t>(g.R) * rstride, pair);
    store_tile<2 * BT>(q2v + 2 * static_cast<size_t>(g.R + 1) * rstride, pair);
  }
}

// ---------------------------------------------------------------------------
// image -> sinogram: one thread per (view, row, column)
// ---------------------------------------------------------------------------
// First and last line (inclusive) whose slab [z0_l, z0_l + Dz pz_l] meets the
// row band (zlo, zhi).  Both conditions are monotone in the line index, so
// the set is an interval; lines outside it have zero z overlap.
__device__ __forceinline__ void line_range(const LineGeo3* __restrict__ lv,
                                           int L, int Dz, double zlo,
                                           double zhi, int* first, int* last) {
  auto below = [&](int l) { return lv[l].z0 < zhi; };
  auto above = [&](int l) { return fma(static_cast<double>(Dz), lv[l].pz, lv[l].z0) > zlo; };
  int lo = 0, hi = L - 1;
  // restrict with each monotone predicate
  for (int pass = 0; pass < 2; ++pass) {
    auto pred = [&](int l) { return pass == 0 ? below(l) : above(l); };
    const bool a = pred(lo), b = pred(hi);
    if (!a && !b) {
      *first = 1;
      *last = 0;
      return;
    }
    if (a && b) continue;
    int good = a ? lo : hi, bad = a ? hi : lo;
    while ((good > bad ? good - bad : bad - good) > 1) {
      const int mid = (good + bad) / 2;
      if (pred(mid)) good = mid; else bad = mid;
    }
    if (a) hi = good; else lo = good;
  }
  *first = lo;
  *last = hi;
}

// Detector row edges of column k on the slice grid of line l.  Row edge r
// sits at z = m (zdet0 + r rowp - sz) + sz on the axis plane, so in slice
// units t_r = A + r B, affine in r.  The forward kernel walks these edges;
// the adjoint inverts the same law (u = (t - A) / B).
struct ZAffine {
  double A, B;
};

__device__ __forceinline__ ZAffine z_affine(const ViewGeo3& vg, const LineGeo3& lg,
                                            double m) {
  const double u = fma(m, vg.zdet0 - vg.sz, vg.sz);
  return {(u - lg.z0) * lg.inv_pz, (m * vg.rowp) * lg.inv_pz};
}

// Row edges r0 + j (j = 0 .. a few) of one column on one line's slice grid:
// t = A + (r0 + j) B, kept as floor(t_r0) plus an FP32 offset (|offset| is
// a few slices, so FP32 holds it to ~1e-6 of a slice), clamped to
// [-2, Dz + 2] like edge_pos_t.
struct RowEdges {
  int base;
  float dt, b, lo, hi;
  __device__ RowEdges(const ZAffine& z, int r0, int Dz) {
    const double t0 = fma(static_cast<double>(r0), z.B, z.A);
    const double tb = floor(fmin(fmax(t0, -1.0e6), 1.0e6));
    base = static_cast<int>(tb);
    dt = static_cast<float>(fmin(fmax(t0, -1.0e6), 1.0e6) - tb);
    b = static_cast<float>(z.B);
    lo = static_cast<float>(-2 - base);
    hi = static_cast<float>(Dz + 2 - base);
  }
  // edge j relative to base, clamped
  __device__ float rel(int j) const {
    return fminf(fmaxf(fmaf(static_cast<float>(j), b, dt), lo), hi);
  }
};

// Rows (forward) and slices (adjoint) one thread carries in registers.
// Forward: 8 rows per thread at instance tiles 1-2, 4 at tile 4 (92 VGPRs,
// 5 waves per SIMD); against 4 / 2 the per-line geometry work and the slice
// a chunk boundary splits are shared by twice the rows, and 16 / 8 (120-124
// VGPRs, 4 waves) loses more in occupancy than it saves.
#ifndef DD3D_FWD_ROWS
#define DD3D_FWD_ROWS 8
#endif
#ifndef DD3D_ADJ_SLICES
#define DD3D_ADJ_SLICES 16
#endif
template <int BT>
constexpr int rows3() {
  return BT >= 4 ? (DD3D_FWD_ROWS > 1 ? DD3D_FWD_ROWS / 2 : 1) : DD3D_FWD_ROWS;
}
template <int BT>
constexpr int slices3() {
  return BT >= 4 ? DD3D_ADJ_SLICES / 2 : DD3D_ADJ_SLICES;
}
#ifndef DD3D_ADJ_WAVES
#define DD3D_ADJ_WAVES 4
#endif
template <int BT>
constexpr int adj_waves3() {
  return DD3D_ADJ_WAVES;
}

// image -> sinogram: one thread per (view, column, chunk of S rows).  Per
// line the column's x footprint and the z law are evaluated once and shared
// by the S rows; the slices the chunk reaches are then visited once each,
// every slice's voxels loaded once and spread over the S rows by their z
// overlaps (slice-major: a row-by-row walk reloads each slice for every
// row that overlaps it).
template <int BT, int WM, bool COLS_FIRST>
__global__ __launch_bounds__(kFwdThreads) void k_bins3d(
    Geo3 g, const ViewGeo3* __restrict__ views, const double* __restrict__ cedge,
    const double* __restrict__ mag, const LineGeo3* __restrict__ lines,
    const float* __restrict__ xrow, const float* __restrict__ xcol,
    float* __restrict__ out) {
  constexpr int S = rows3<BT>();
  constexpr int CH = COLS_FIRST ? kChunksColsFirst : kChunksViewsFirst;
  const int v = COLS_FIRST ? blockIdx.y : blockIdx.x;
  const int k = (COLS_FIRST ? blockIdx.x : blockIdx.y) * kColsPerBlock +
                (threadIdx.x % kColsPerBlock);
  const int nch = (g.R + S - 1) / S;
  const int ncb = (nch + CH - 1) / CH;
  const int ch = (blockIdx.z % ncb) * CH + threadIdx.x / kColsPerBlock;
  const int b0 = (blockIdx.z / ncb) * BT;
  const int r0 = ch * S;
  if (k >= g.C || r0 >= g.R) return;
  const int nr = min(S, g.R - r0);
  const ViewGeo3 vg = views[v];
  const int xaxis = vg.xaxis;
  const int L = xaxis ? g.H : g.W;
  const int N = xaxis ? g.W : g.H;
  const float* __restrict__ src =
      (xaxis ? xrow : xcol) + static_cast<size_t>(b0) * g.W * g.H * g.Dz;  // tile b0 / BT
  const double* ce = cedge + static_cast<size_t>(v) * (g.C + 1);
  const double e1 = ce[k], e2 = ce[k + 1];
  const double m = mag[static_cast<size_t>(v) * g.C + k];
  const LineGeo3* __restrict__ lv = lines + static_cast<size_t>(v) * g.Lmax;
  float acc[S][BT];
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int b = 0; b < BT; ++b) acc[j][b] = 0.0f;
  int lfirst, llast;
  line_range(lv, L, g.Dz, row_z(vg, m, r0), row_z(vg, m, r0 + nr), &lfirst, &llast);
  // one line of margin: the range test and the slice walk round differently
  lfirst = max(lfirst - 1, 0);
  llast = min(llast + 1, L - 1);
  // the next line's geometry is loaded while this line is worked on
  LineGeo3 lg_next = lv[min(lfirst, L - 1)];
  for (int l = lfirst; l <= llast; ++l) {
    const LineGeo3 lg = lg_next;
    lg_next = lv[min(l + 1, llast)];
    const EdgePos p1 = edge_pos(e1, lg.a0, lg.inv_p, N);
    const EdgePos p2 = edge_pos(e2, lg.a0, lg.inv_p, N);
    const int cs = max(p1.c, 0), cl = min(p2.c, N - 1);
    if (cs > cl) continue;
    const ZAffine z = z_affine(vg, lg, m);
    // the chunk's row edges relative to floor(t_r0); rows past nr are empty
    const RowEdges re(z, r0, g.Dz);
    float t[S + 1];
#pragma unroll
    for (int j = 0; j <= S; ++j) t[j] = re.rel(min(j, nr));
    const int s_lo = max(re.base + static_cast<int>(floorf(t[0])), 0);
    const int s_hi = min(re.base + static_cast<int>(floorf(t[S])), g.Dz - 1);
    if (s_lo > s_hi) continue;  // the chunk's rows miss the slab
    const float wl = (WM == W3_FP) ? lg.p * lg.pzf : vg.kbp;
    const float* __restrict__ line = src + static_cast<size_t>(l) * g.Dz * N * BT;
    // Columns in pairs: each slice's pair is loaded once and its x-weighted
    // sum spread over the chunk's rows by their z overlaps.  The second
    // column of a pair past cl is masked (its load stays inside the
    // buffer, which carries one column of padding).
    for (int c0 = cs; c0 <= cl; c0 += 2) {
      const bool two = c0 + 1 <= cl;
      float w0 = overlap_in_pixel(p1, p2, c0) * wl;
      float w1 = two ? overlap_in_pixel(p1, p2, c0 + 1) * wl : 0.0f;
      if constexpr (WM == W3_BPW) {
        float xc, yc;
        pixel_centre3(g, xaxis, l, c0, &xc, &yc);
        w0 *= fbp_weight3(g, vg, xc, yc);
        pixel_centre3(g, xaxis, l, c0 + 1, &xc, &yc);
        w1 *= fbp_weight3(g, vg, xc, yc);
      }
      const float* __restrict__ px = line + (static_cast<size_t>(s_lo) * N + c0) * BT;
      const size_t sstride = static_cast<size_t>(N) * BT;
      for (int sl = s_lo; sl <= s_hi; ++sl, px += sstride) {
        float x0[BT], x1[BT], sx[BT];
        load_pair<BT>(px, BT, x0, x1);
#pragma unroll
        for (int b = 0; b < BT; ++b) {
          sx[b] = w0 * x0[b];
          sx[b] = two ? fmaf(w1, x1[b], sx[b]) : sx[b];
        }
        // row j covers [t_j, t_j+1]; its overlap with slice [rs, rs + 1] is
        // the difference of the two edges clamped to the slice
        const float rs = static_cast<float>(sl - re.base), rs1 = rs + 1.0f;
        float clo = __builtin_amdgcn_fmed3f(t[0], rs, rs1);
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const float chi = __builtin_amdgcn_fmed3f(t[j + 1], rs, rs1);
          const float wz = chi - clo;
          clo = chi;
#pragma unroll
          for (int b = 0; b < BT; ++b) acc[j][b] = fmaf(wz, sx[b], acc[j][b]);
        }
      }
    }
  }
  const int kk = vg.flip ? g.C - 1 - k : k;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    if (j < nr) {
      const int r = r0 + j;
      const float wk = (WM == W3_FP)
                           ? fp_weight3(g, vg, e1, e2, row_z(vg, m, r), row_z(vg, m, r + 1))
                           : 1.0f;
#pragma unroll
      for (int b = 0; b < BT; ++b)
        if (b0 + b < g.B)
          out[((static_cast<size_t>(b0 + b) * g.V + v) * g.R + r) * g.C + kk] = acc[j][b] * wk;
    }
  }
}

// ---------------------------------------------------------------------------
// sinogram -> image: one thread per (pixel, chunk of S slices), all views
// ---------------------------------------------------------------------------
// First-column guess in float; callers fix it exactly with edge_pos.
__device__ __forceinline__ int guess_col_fast(const Geo3& g, const ViewGeo3& vg, double a) {
  float kf;
  if (g.scan == DD_CONE_FLAT) {
    const float R = g.sid, af = static_cast<float>(a);
    const float sf = static_cast<float>(vg.s), cf = static_cast<float>(vg.c);
    const float u = vg.xaxis ? R * af * cf * __builtin_amdgcn_rcpf(R + sf * af)
                             : R * af * sf * __builtin_amdgcn_rcpf(R - cf * af);
    const float d = u - static_cast<float>(vg.cshp);
    kf = g.C * 0.5f + (vg.flip ? -d : d) * __builtin_amdgcn_rcpf(static_cast<float>(vg.colp));
  } else {
    // d = gamma - theta_ref (mod pi) from tan d = (Y cos - X sin) / (X cos + Y sin),
    // atan by Abramowitz & Stegun 4.4.49 (|error| <= 1.2e-5 on [-1, 1])
    const float Y = vg.xaxis ? static_cast<float>(a - vg.sx) : vg.sxf;
    const float X = vg.xaxis ? vg.syf : static_cast<float>(a - vg.sy);
    const float num = Y * vg.cref - X * vg.sref;
    const float den = X * vg.cref + Y * vg.sref;
    const float t = fminf(fmaxf(num * __builtin_amdgcn_rcpf(den), -1.0f), 1.0f);
    const float t2 = t * t;
    const float d =
        t * (0.9998660f +
             t2 * (-0.3302995f + t2 * (0.1801410f + t2 * (-0.0851330f + t2 * 0.0208351f))));
    kf = g.C * 0.5f + (vg.flip ? -d : d) * __builtin_amdgcn_rcpf(static_cast<float>(vg.colp));
  }
  kf = fminf(fmaxf(kf, 0.0f), g.C - 1.0f);
  return static_cast<int>(kf);
}

// Columns [kb, ke) that hold every pixel edge of one 16 x 16 pixel tile in
// one view, and the range of their magnifications (k_voxels3d's staging).
struct VoxRange {
  double mmin, mmax;
  int kb, ke;
};

// One thread per (view, pixel tile).  The tile's axis extent is spanned by
// its corners (a pixel edge's axis point is monotone along and across the
// lines), widened by one column edge each side against rounding.
__global__ __launch_bounds__(kThreads3) void k_voxel_ranges(
    Geo3 g, const ViewGeo3* __restrict__ views, const double* __restrict__ cedge,
    const double* __restrict__ mag, const LineGeo3* __restrict__ lines,
    VoxRange* __restrict__ out) {
  const int ntx = (g.W + kTile3 - 1) / kTile3;
  const int tiles = ntx * ((g.H + kTile3 - 1) / kTile3);
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(tiles) * g.V) return;
  const int v = static_cast<int>(idx / tiles), t = static_cast<int>(idx % tiles);
  const int c0 = (t % ntx) * kTile3, r0 = (t / ntx) * kTile3;
  const int c1 = min(c0 + kTile3, g.W), r1 = min(r0 + kTile3, g.H);
  const ViewGeo3 vg = views[v];
  const LineGeo3* __restrict__ lv = lines + static_cast<size_t>(v) * g.Lmax;
  double amin = 1e300, amax = -1e300;
  auto corner = [&](int l, int i) {
    const double a = lv[l].a0 + i / lv[l].inv_p;
    amin = fmin(amin, a);
    amax = fmax(amax, a);
  };
  if (vg.xaxis) {  // lines = rows, pixel edges = columns
    corner(r0, c0), corner(r0, c1), corner(r1 - 1, c0), corner(r1 - 1, c1);
  } else {  // lines = columns, pixel edges from the bottom
    corner(c0, g.H - r1), corner(c0, g.H - r0), corner(c1 - 1, g.H - r1),
        corner(c1 - 1, g.H - r0);
  }
  const double* __restrict__ ce = cedge + static_cast<size_t>(v) * (g.C + 1);
  int lo = 0, hi = g.C;  // last edge <= amin
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (ce[mid] <= amin) lo = mid; else hi = mid - 1;
  }
  const int kb = max(lo - 1, 0);
  lo = 0, hi = g.C;  // first edge >= amax
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (ce[mid] >= amax) hi = mid; else lo = mid + 1;
  }
  const int ke = max(min(lo + 1, g.C), kb + 1);
  double mmin = 1e300, mmax = -1e300;
  for (int k = kb; k < ke; ++k) {
    const double m = mag[static_cast<size_t>(v) * g.C + k];
    mmin = fmin(mmin, m);
    mmax = fmax(mmax, m);
  }
  out[idx] = VoxRange{mmin, mmax, kb, ke};
}

// Floats of LDS the staged k_voxels3d holds one view's rows x columns of
// running sums in (32 KB).
constexpr int kStage3Floats = 8192;
// Staged rows are laid out with an odd number of columns (one unused column
// after an even count), so that lanes reading the same column on
// neighbouring rows fall in different LDS banks.
constexpr bool kStage3OddPitch = true;

// STG: per view, the block of running sums the workgroup's 16 x 16 pixel
// tile and slice chunk can read is staged in LDS first (rows x columns from
// the view's column range and the chunk's z extent; two barriers per view),
// and every evaluation reads it there; a view whose block exceeds the
// buffer reads global memory.  The values are copies, so the bits are those
// of the unstaged gather.
template <int BT, int WM, bool COMP, bool STG>
__global__ __launch_bounds__(kThreads3)
__attribute__((amdgpu_waves_per_eu(adj_waves3<BT>()))) void k_voxels3d(
    Geo3 g, const ViewGeo3* __restrict__ views, const double* __restrict__ cedge,
    const double* __restrict__ mag, const LineGeo3* __restrict__ lines,
    const float* __restrict__ sino, const float* __restrict__ sino2,
    float* __restrict__ out, const VoxRange* __restrict__ ranges) {
  constexpr int S = slices3<BT>();
  // staged view: rows [srb, srb + sn / sncols) x columns [skb, skb + sncols)
  // of hi (or the rounded sums) in lq[0 ..), spitch entries per row, and the
  // (lo, row value) pairs from lq[spair] (16-byte aligned) on
  __shared__ __align__(16) float lq[STG ? kStage3Floats : 1];
  const int ntx = (g.W + kTile3 - 1) / kTile3;
  const int tile = blockIdx.x;
  const int tiles = ntx * ((g.H + kTile3 - 1) / kTile3);
  const int col = (blockIdx.x % ntx) * kTile3 + (threadIdx.x % kTile3);
  const int row = (blockIdx.x / ntx) * kTile3 + (threadIdx.x / kTile3);
  const int s0 = blockIdx.y * S;
  const int b0 = blockIdx.z * BT;
  if (!STG && (col >= g.W || row >= g.H)) return;
  const bool active = col < g.W && row < g.H;
  float xcw = 0.0f, ycw = 0.0f;
  if constexpr (WM == W3_BPW) pixel_centre3(g, 1, row, col, &xcw, &ycw);
  float acc[S][BT];
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int b = 0; b < BT; ++b) acc[j][b] = 0.0f;
  // world z extent of this thread's slices (block-uniform)
  const float z0w = -g.Dz * 0.5f * g.dz + (g.scan == DD_SPIRAL ? 0.0f : g.shz);
  const float zc_lo = z0w + s0 * g.dz, zc_hi = z0w + min(s0 + S, g.Dz) * g.dz;
  for (int v = 0; v < g.V; ++v) {
    const ViewGeo3 vg = views[v];
    if (zc_hi < vg.zlo || zc_lo > vg.zhi) continue;  // the view's cone misses the chunk
    // staging (block-uniform): the rows the chunk's slices can reach on the
    // tile's lines through its columns.  A row coordinate r = ((z - sz) / m
    // + sz - zdet0) / rowp is monotone in the slice edge's z (extreme on the
    // first and last line at the chunk's first and last slice edge) and in
    // the column magnification m, so the box corners bound it; one row of
    // margin each side (two above: rows r and r + 1 are read).
    bool staged = false;
    int srb = 0, skb = 0, sncols = 1, sn = 0, spitch = 1, spair = 0;
    if constexpr (STG) {
      const VoxRange vr = ranges[static_cast<size_t>(v) * tiles + tile];
      const int l0 = vg.xaxis ? (blockIdx.x / ntx) * kTile3 : (blockIdx.x % ntx) * kTile3;
      const int l1 = min(l0 + kTile3, vg.xaxis ? g.H : g.W) - 1;
      const LineGeo3* __restrict__ lvv = lines + static_cast<size_t>(v) * g.Lmax;
      double zmin = 1e300, zmax = -1e300;
      for (int e = 0; e < 2; ++e) {
        const LineGeo3 lg = lvv[e ? l1 : l0];
        zmin = fmin(zmin, fma(static_cast<double>(s0), lg.pz, lg.z0));
        zmax = fmax(zmax, fma(static_cast<double>(s0 + S), lg.pz, lg.z0));
      }
      double rmin = 1e300, rmax = -1e300;
      for (int e = 0; e < 4; ++e) {
        const double z = (e & 1) ? zmax : zmin, m = (e & 2) ? vr.mmax : vr.mmin;
        const double r = ((z - vg.sz) / m + vg.sz - vg.zdet0) / vg.rowp;
        rmin = fmin(rmin, r);
        rmax = fmax(rmax, r);
      }
      srb = static_cast<int>(fmin(fmax(floor(rmin) - 1.0, 0.0), static_cast<double>(g.R + 1)));
      const int sre =
          static_cast<int>(fmin(fmax(floor(rmax) + 2.0, 0.0), static_cast<double>(g.R + 1)));
      skb = vr.kb;
      sncols = vr.ke - vr.kb;
      spitch = kStage3OddPitch ? (sncols | 1) : sncols;
      const int nrow = sre - srb + 1;
      sn = nrow * sncols;
      spair = (nrow * spitch * BT + 3) & ~3;
      staged = (COMP ? spair + 2 * nrow * spitch * BT : nrow * spitch * BT) <= kStage3Floats;
      __syncthreads();  // the previous view's walks are done with lq
      if (staged) {
        const size_t vblock = static_cast<size_t>(v) * (g.R + 2) * g.C * g.Bp;
        for (int e = threadIdx.x; e < sn; e += kThreads3) {
          const size_t src = vblock + (static_cast<size_t>(srb + e / sncols) * g.C + skb +
                                       e % sncols) * g.Bp + b0;
          const int dst = (e / sncols) * spitch + e % sncols;
          float h[BT];
          load_tile<BT>(sino + src, h);
#pragma unroll
          for (int b = 0; b < BT; ++b) lq[dst * BT + b] = h[b];
          if constexpr (COMP) {
            float lx[2 * BT];
            load_tile<2 * BT>(sino2 + 2 * src, lx);
#pragma unroll
            for (int b = 0; b < 2 * BT; ++b) lq[spair + dst * 2 * BT + b] = lx[b];
          }
        }
      }
      __syncthreads();
      if (!active) continue;
    }
    int l, i, N;
    if (vg.xaxis) {
      l = row;
      i = col;
      N = g.W;
    } else {
      l = col;
      i = g.H - 1 - row;
      N = g.H;
    }
    const LineGeo3 lg = lines[static_cast<size_t>(v) * g.Lmax + l];
    const double* __restrict__ ce = cedge + static_cast<size_t>(v) * (g.C + 1);
    const double* __restrict__ mv = mag + static_cast<size_t>(v) * g.C;
    // first column: the last whose lower edge lies before pixel i (or 0);
    // the edge positions found on the way start the walk
    int k = guess_col_fast(g, vg, lg.a0 + (i + 0.5) * static_cast<double>(lg.p));
    EdgePos p1 = edge_pos(ce[k], lg.a0, lg.inv_p, N);
    while (k > 0 && p1.c >= i) p1 = edge_pos(ce[--k], lg.a0, lg.inv_p, N);
    EdgePos p2 = edge_pos(ce[k + 1], lg.a0, lg.inv_p, N);
    bool past = false;
    while (p2.c < i) {
      if (++k >= g.C) {
        past = true;
        break;
      }
      p1 = p2;
      p2 = edge_pos(ce[k + 1], lg.a0, lg.inv_p, N);
    }
    if (past) continue;
    float gw = (WM == W3_FP) ? lg.p * lg.pzf : vg.kbp;
    if constexpr (WM == W3_BPW) gw *= fbp_weight3(g, vg, xcw, ycw);
    while (p1.c <= i) {
      const float wx = overlap_in_pixel(p1, p2, i) * gw;
      const ZAffine z = z_affine(vg, lg, mv[k]);
      // Slice sl receives sum_r Q-weighted row overlaps = B (L(u_sl+1) -
      // L(u_sl)), where u = (t - A) / B is the detector-row coordinate of
      // slice edge t and L(u) interpolates the row prefix Q linearly.  u is
      // kept relative to the chunk's first row so float holds it exactly
      // enough.
      const double inv_b = 1.0 / z.B;
      const double u0 = (static_cast<double>(s0) - z.A) * inv_b;
      const double rb = floor(u0);
      const int rbase = static_cast<int>(fmin(fmax(rb, -1.0e6), 1.0e6));
      const float du = static_cast<float>(u0 - rb);
      const float invb = static_cast<float>(inv_b);
      const float ulo = static_cast<float>(-rbase), uhi = static_cast<float>(g.R - rbase);
      if (STG && staged && uhi > 0.0f && fmaf(static_cast<float>(S), invb, du) > ulo) {
        // the same evaluations as below, reading the staged block: element
        // (row rbase + fu, column k) of the hi block at lb + fu * rsl
        const float wxb = wx * static_cast<float>(z.B);
        const int lb = ((rbase - srb) * spitch + (k - skb)) * BT;
        const int rsl = spitch * BT;
        const __attribute__((address_space(3))) float* lh =
            (const __attribute__((address_space(3))) float*)(lq);
        if constexpr (COMP) {
          auto eval = [&](int j, float (&h)[BT], float (&m)[BT]) {
            const float u = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(j), invb, du), ulo, uhi);
            const float fu = floorf(u);
            const float fr = u - fu;
            const int e = lb + static_cast<int>(fu) * rsl;
#pragma unroll
            for (int b = 0; b < BT; ++b) {
              h[b] = lh[e + b];
              m[b] = fmaf(fr, lh[spair + 2 * (e + b) + 1], lh[spair + 2 * (e + b)]);
            }
          };
          float hprev[BT], mprev[BT];
          eval(0, hprev, mprev);
#pragma unroll
          for (int j = 0; j < S; ++j) {
            float hnext[BT], mnext[BT];
            eval(j + 1, hnext, mnext);
#pragma unroll
            for (int b = 0; b < BT; ++b) {
              acc[j][b] = fmaf(wxb, (hnext[b] - hprev[b]) + (mnext[b] - mprev[b]), acc[j][b]);
              hprev[b] = hnext[b];
              mprev[b] = mnext[b];
            }
          }
        } else {
          auto eval = [&](int j, float (&l)[BT]) {
            const float u = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(j), invb, du), ulo, uhi);
            const float fu = floorf(u);
            const float fr = u - fu;
            const int e = lb + static_cast<int>(fu) * rsl;
#pragma unroll
            for (int b = 0; b < BT; ++b) l[b] = fmaf(fr, lh[e + rsl + b] - lh[e + b], lh[e + b]);
          };
          float lprev[BT];
          eval(0, lprev);
#pragma unroll
          for (int j = 0; j < S; ++j) {
            float lnext[BT];
            eval(j + 1, lnext);
#pragma unroll
            for (int b = 0; b < BT; ++b) {
              acc[j][b] = fmaf(wxb, lnext[b] - lprev[b], acc[j][b]);
              lprev[b] = lnext[b];
            }
          }
        }
      } else if (uhi > 0.0f && fmaf(static_cast<float>(S), invb, du) > ulo) {
        const float wxb = wx * static_cast<float>(z.B);
        // the view's block of Q (wave-uniform base) plus 32-bit byte
        // offsets: a block holds (R + 2) C Bp < 2^30 floats (checked on the
        // host), so base + offset maps onto scalar-base vector loads
        const size_t vblock = static_cast<size_t>(v) * (g.R + 2) * g.C * g.Bp;
        const char* __restrict__ qv = reinterpret_cast<const char*>(sino + vblock);
        const unsigned rstride = 4u * static_cast<unsigned>(g.C * g.Bp);
        const unsigned okb =
            4u * static_cast<unsigned>(k * g.Bp + b0) + static_cast<unsigned>(rbase) * rstride;
        if constexpr (COMP) {
          // Compensated form.  At slice edge u (row r = floor u, fraction f)
          // the running sum is hi[r] + (lo[r] + f x_r); slice j receives
          //   (hi[r'] - hi[r]) + ((lo + f x)[r'] - (lo + f x)[r])
          // between its edges r + f and r' + f'.  Both differences are of the
          // size of the slice's rows, so FP32 rounds them relative to the
          // slice rather than to the R-row running sum.
          const char* __restrict__ q2v = reinterpret_cast<const char*>(sino2 + 2 * vblock);
          auto eval = [&](int j, float (&h)[BT], float (&m)[BT]) {
            const float u = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(j), invb, du), ulo, uhi);
            const float fu = floorf(u);
            const float fr = u - fu;
            const unsigned off = okb + static_cast<unsigned>(static_cast<int>(fu)) * rstride;
            float lx[2 * BT];
            load_tile<BT>(reinterpret_cast<const float*>(qv + off), h);
            load_tile<2 * BT>(reinterpret_cast<const float*>(q2v + 2u * off), lx);
#pragma unroll
            for (int b = 0; b < BT; ++b) m[b] = fmaf(fr, lx[2 * b + 1], lx[2 * b]);
          };
          float hprev[BT], mprev[BT];
          eval(0, hprev, mprev);
#pragma unroll
          for (int j = 0; j < S; ++j) {
            float hnext[BT], mnext[BT];
            eval(j + 1, hnext, mnext);
#pragma unroll
            for (int b = 0; b < BT; ++b) {
              acc[j][b] = fmaf(wxb, (hnext[b] - hprev[b]) + (mnext[b] - mprev[b]), acc[j][b]);
              hprev[b] = hnext[b];
              mprev[b] = mnext[b];
            }
          }
        } else {
          auto eval = [&](int j, float (&l)[BT]) {
            // u in [-rbase, R - rbase]: rows r, r + 1 with r <= R (Q[R+1] = Q[R])
            const float u = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(j), invb, du), ulo, uhi);
            const float fu = floorf(u);
            const float fr = u - fu;
            const unsigned off = okb + static_cast<unsigned>(static_cast<int>(fu)) * rstride;
            float q0[BT], q1[BT];
            load_tile<BT>(reinterpret_cast<const float*>(qv + off), q0);
            load_tile<BT>(reinterpret_cast<const float*>(qv + (off + rstride)), q1);
#pragma unroll
            for (int b = 0; b < BT; ++b) l[b] = fmaf(fr, q1[b] - q0[b], q0[b]);
          };
          float lprev[BT];
          eval(0, lprev);
#pragma unroll
          for (int j = 0; j < S; ++j) {
            float lnext[BT];
            eval(j + 1, lnext);
#pragma unroll
            for (int b = 0; b < BT; ++b) {
              acc[j][b] = fmaf(wxb, lnext[b] - lprev[b], acc[j][b]);
              lprev[b] = lnext[b];
            }
          }
        }
      }
      if (++k >= g.C) break;
      p1 = p2;
      if (p1.c > i) break;
      p2 = edge_pos(ce[k + 1], lg.a0, lg.inv_p, N);
    }
  }
  if (!active) return;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int sl = s0 + j;
    if (sl < g.Dz) {
#pragma unroll
      for (int b = 0; b < BT; ++b)
        if (b0 + b < g.B)
          out[((static_cast<size_t>(b0 + b) * g.Dz + sl) * g.H + row) * g.W + col] = acc[j][b];
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
constexpr size_t kAlign3 = 256;

struct Layout3 {
  size_t views, cedge, mag, lines, geo_total;
  size_t buf0, buf1, ranges, data_total;
};

bool to_sino3(int op) {
  return op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW;
}

int pick_tile3(int batch) {
  int t = 1;
  while (t < batch && t < 4) t *= 2;
  return t;
}

Layout3 make_layout3(const Geo3& g, int op) {
  Layout3 L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off = align_up(off + bytes, kAlign3);
    return at;
  };
  L.views = take(sizeof(ViewGeo3) * g.V);
  L.cedge = take(sizeof(double) * g.V * (g.C + 1));
  L.mag = take(sizeof(double) * g.V * g.C);
  L.lines = take(sizeof(LineGeo3) * g.V * g.Lmax);
  L.geo_total = off;
  off = 0;
  const size_t vox = static_cast<size_t>(g.W) * g.H * g.Dz * g.Bp;
  if (to_sino3(op)) {
    // one pixel of padding: k_bins3d loads pixels in pairs
    L.buf0 = take(sizeof(float) * (vox + g.BT));
    L.buf1 = take(sizeof(float) * (vox + g.BT));
  } else {
    const size_t q = sizeof(float) * static_cast<size_t>(g.V) * (g.R + 2) * g.C * g.Bp;
    L.buf0 = take(q);
    // compensated running sums: a second plane of (lo, row value) pairs
    L.buf1 = use_compensated(g.BT) ? take(2 * q) : L.buf0;
    // column ranges of the adjoint's LDS staging, per (view, 16 x 16 tile)
    L.ranges = take(sizeof(VoxRange) * g.V * ((g.W + kTile3 - 1) / kTile3) *
                    ((g.H + kTile3 - 1) / kTile3));
  }
  L.data_total = off;
  return L;
}

bool fill_geo3(Geo3* out, int scan, const dd_geom3d* g, int batch, int views) {
  if (!g || batch <= 0 || views <= 0) return false;
  if (g->width <= 0 || g->height <= 0 || g->depth <= 0 || g->num_col <= 0 ||
      g->num_row <= 0)
    return false;
  Geo3 q;
  q.W = g->width;
  q.H = g->height;
  q.Dz = g->depth;
  q.C = g->num_col;
  q.R = g->num_row;
  q.V = views;
  q.B = batch;
  q.BT = pick_tile3(batch);
  q.Bp = align_up(batch, q.BT);
  q.Lmax = std::max(q.W, q.H);
  q.scan = scan;
  q.dx = g->pix_x;
  q.dy = g->pix_y;
  q.dz = g->pix_z;
  q.col = g->col_size;
  q.row = g->row_size;
  q.sid = g->iso_source;
  q.sdd = g->source_detector;
  q.shx = g->shift_x;
  q.shy = g->shift_y;
  q.shz = g->shift_z;
  q.csh = g->col_shift;
  q.rsh = g->row_shift;
  // the adjoint addresses one view's row running sums with 32-bit byte
  // offsets (twice the offset for the (lo, row value) plane)
  if (static_cast<long long>(q.R + 2) * q.C * q.Bp >= (1LL << 29)) return false;
  *out = q;
  return true;
}

// sinogram -> volume: running sums over rows, then the voxel gather
template <int BT, bool COMP, bool STG>
void launch_voxels3(const Geo3& g, int op, dim3 grid, const ViewGeo3* views,
                    const double* cedge, const double* mag, const LineGeo3* lines, float* q,
                    float* q2, float* output, VoxRange* ranges, hipStream_t st) {
  if (op == DD_PROJ_T)
    hipLaunchKernelGGL((k_voxels3d<BT, W3_FP, COMP, STG>), grid, dim3(kThreads3), 0, st, g,
                       views, cedge, mag, lines, q, q2, output, ranges);
  else if (op == DD_BPROJ)
    hipLaunchKernelGGL((k_voxels3d<BT, W3_BP, COMP, STG>), grid, dim3(kThreads3), 0, st, g,
                       views, cedge, mag, lines, q, q2, output, ranges);
  else
    hipLaunchKernelGGL((k_voxels3d<BT, W3_BPW, COMP, STG>), grid, dim3(kThreads3), 0, st, g,
                       views, cedge, mag, lines, q, q2, output, ranges);
}

template <int BT, bool COMP>
void launch_adjoint3(const Geo3& g, int op, const float* input, float* output,
                     const ViewGeo3* views, const double* cedge, const double* mag,
                     const LineGeo3* lines, float* q, float* q2, VoxRange* ranges,
                     hipStream_t st) {
  dim3 sgrid((g.C + kThreads3 - 1) / kThreads3 * g.V, 1, g.Bp / BT);
  if (op == DD_PROJ_T)
    hipLaunchKernelGGL((k_prefix_sino3d<BT, true, COMP>), sgrid, dim3(kThreads3), 0, st, g,
                       views, cedge, mag, input, q, q2);
  else
    hipLaunchKernelGGL((k_prefix_sino3d<BT, false, COMP>), sgrid, dim3(kThreads3), 0, st, g,
                       views, cedge, mag, input, q, q2);
  const int ntx = (g.W + kTile3 - 1) / kTile3, nty = (g.H + kTile3 - 1) / kTile3;
  constexpr int S = slices3<BT>();
  dim3 grid(ntx * nty, (g.Dz + S - 1) / S, g.Bp / BT);
  ProfileScope prof(DD_KERNEL_SINO_TO_IMAGE, st);
  if (use_stage3(BT, COMP)) {
    const long long n = 1LL * ntx * nty * g.V;
    hipLaunchKernelGGL(k_voxel_ranges, dim3(static_cast<unsigned>((n + kThreads3 - 1) / kThreads3)),
                       dim3(kThreads3), 0, st, g, views, cedge, mag, lines, ranges);
    launch_voxels3<BT, COMP, true>(g, op, grid, views, cedge, mag, lines, q, q2, output, ranges,
                                   st);
  } else {
    launch_voxels3<BT, COMP, false>(g, op, grid, views, cedge, mag, lines, q, q2, output,
                                    ranges, st);
  }
}

bool use_cols_first(const Geo3& g, int bt) {
  const int m = g_fwd_order.load();
  if (g.V > 65535) return false;  // the views are grid dimension y
  if (m != 2) return m == 1;
  const double tile_bytes = 4.0 * g.W * g.H * g.Dz * bt;
  return tile_bytes > 192.0 * (1 << 20);
}

template <int BT, bool COLS_FIRST>
void launch_bins3(const Geo3& g, int op, dim3 grid, int chunks, const ViewGeo3* views,
                  const double* cedge, const double* mag, const LineGeo3* lines,
                  const float* buf0, const float* buf1, float* output, hipStream_t st) {
  const dim3 block(kColsPerBlock * chunks);
  if (op == DD_PROJ)
    hipLaunchKernelGGL((k_bins3d<BT, W3_FP, COLS_FIRST>), grid, block, 0, st, g, views, cedge,
                       mag, lines, buf0, buf1, output);
  else if (op == DD_BPROJ_T)
    hipLaunchKernelGGL((k_bins3d<BT, W3_BP, COLS_FIRST>), grid, block, 0, st, g, views, cedge,
                       mag, lines, buf0, buf1, output);
  else
    hipLaunchKernelGGL((k_bins3d<BT, W3_BPW, COLS_FIRST>), grid, block, 0, st, g, views, cedge,
                       mag, lines, buf0, buf1, output);
}

template <int BT>
dd_status launch3(const Geo3& g, int op, const float* input, float* output,
                  char* geo, char* ws, const Layout3& L, hipStream_t st) {
  auto* views = reinterpret_cast<ViewGeo3*>(geo + L.views);
  auto* cedge = reinterpret_cast<double*>(geo + L.cedge);
  auto* mag = reinterpret_cast<double*>(geo + L.mag);
  auto* lines = reinterpret_cast<LineGeo3*>(geo + L.lines);
  auto* buf0 = reinterpret_cast<float*>(ws + L.buf0);
  auto* buf1 = reinterpret_cast<float*>(ws + L.buf1);
  const int ntile = g.Bp / BT;
  if (to_sino3(op)) {
    hipLaunchKernelGGL(k_interleave_volume<BT>,
                       dim3(((g.W + kItile3 - 1) / kItile3) * ((g.H + kItile3 - 1) / kItile3),
                            g.Dz, ntile),
                       dim3(kItile3 * kItile3), 0, st, g, input, buf0, buf1);
    constexpr int S = rows3<BT>();
    const int nch = (g.R + S - 1) / S;
    const unsigned ncolb = (g.C + kColsPerBlock - 1) / kColsPerBlock;
    ProfileScope prof(DD_KERNEL_IMAGE_TO_SINO, st);
    if (use_cols_first(g, BT)) {
      const int ncb = (nch + kChunksColsFirst - 1) / kChunksColsFirst;
      launch_bins3<BT, true>(g, op, dim3(ncolb, g.V, ncb * ntile), kChunksColsFirst, views,
                             cedge, mag, lines, buf0, buf1, output, st);
    } else {
      const int ncb = (nch + kChunksViewsFirst - 1) / kChunksViewsFirst;
      launch_bins3<BT, false>(g, op, dim3(g.V, ncolb, ncb * ntile), kChunksViewsFirst, views,
                              cedge, mag, lines, buf0, buf1, output, st);
    }
  } else {
    auto* ranges = reinterpret_cast<VoxRange*>(ws + L.ranges);
    if (buf1 != buf0)
      launch_adjoint3<BT, true>(g, op, input, output, views, cedge, mag, lines, buf0, buf1,
                                ranges, st);
    else
      launch_adjoint3<BT, false>(g, op, input, output, views, cedge, mag, lines, buf0, buf1,
                                 ranges, st);
  }
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

}  // namespace
}  // namespace dd

using namespace dd;

extern "C" {

void dd3d_set_compensated(int mode) {
  g_compensated = (mode == 0 || mode == 1) ? mode : 2;
}

void dd3d_set_adjoint_staging(int mode) { g_stage3 = (mode == 0 || mode == 1) ? mode : 2; }

void dd3d_set_forward_order(int mode) { g_fwd_order = (mode == 0 || mode == 1) ? mode : 2; }

size_t dd3d_workspace_size(int scan, int op, const dd_geom3d* g, int batch,
                           int views) {
  Geo3 q;
  if (scan < DD_CONE_FLAT || scan > DD_SPIRAL) return 0;
  if (!fill_geo3(&q, scan, g, batch, views)) return 0;
  const Layout3 L = make_layout3(q, op);
  return L.geo_total + L.data_total;
}

dd_status dd3d_run(int scan, int op, const dd_geom3d* g, const float* input,
                   float* output, const float* angles, const float* source_pos,
                   int batch, int views, void* ws, size_t ws_bytes,
                   void* stream) {
  if (scan < DD_CONE_FLAT || scan > DD_SPIRAL) return DD_BAD_ARG;
  if (op < DD_PROJ || op > DD_BPROJ_TW) return DD_BAD_ARG;
  if (!input || !output || !angles) return DD_BAD_ARG;
  if (scan == DD_SPIRAL && !source_pos) return DD_BAD_ARG;
  Geo3 q;
  if (!fill_geo3(&q, scan, g, batch, views)) return DD_BAD_ARG;
  const Layout3 L = make_layout3(q, op);
  if (!ws || ws_bytes < L.geo_total + L.data_total) return DD_NO_WORKSPACE;
  auto st = static_cast<hipStream_t>(stream);
  char* w = static_cast<char*>(ws);
  hipLaunchKernelGGL(k_setup3d, dim3(q.V), dim3(kThreads3), 0, st, q, angles,
                     source_pos, reinterpret_cast<ViewGeo3*>(w + L.views),
                     reinterpret_cast<double*>(w + L.cedge),
                     reinterpret_cast<double*>(w + L.mag),
                     reinterpret_cast<LineGeo3*>(w + L.lines));
  char* data = w + L.geo_total;
  switch (q.BT) {
    case 1: return launch3<1>(q, op, input, output, w, data, L, st);
    case 2: return launch3<2>(q, op, input, output, w, data, L, st);
    default: return launch3<4>(q, op, input, output, w, data, L, st);
  }
}

}  // extern "C"
