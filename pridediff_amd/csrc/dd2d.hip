// 2D distance-driven projectors (parallel, fan-flat, fan-arc) for gfx950.
//
// Behaviour followed: deeprecon/cpu/dis2d/FanArcDisCPU2d.cpp (fan arc),
// FanFlatDisCPU2d.cpp (fan flat), ParaDisCPU2d.cpp (parallel), and the op
// table of deeprecon/interface/RadonTorch.cpp:17-568.
//
// One call = four stream-ordered launches, no host synchronisation:
//   1. k_setup2d        per view: axis/flip choice, bin edges on the common
//                       axis (FP64), FP weights, per-line pixel-edge maps.
//   2. k_interleave_*   copies the input into an instance-interleaved layout
//                       ([line][pixel][b] or [view][bin][b]) so one geometry
//                       evaluation feeds BT instances through vector loads.
//   3. k_bins           image -> sinogram ops: one thread per detector bin
//                       gathers along every image line (no atomics).
//      k_pixels         sinogram -> image ops: one thread per pixel gathers
//                       over every view (no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "dd.h"
#include "dd_common.h"
#include "dd_profile.h"

namespace dd {
namespace {

constexpr int kThreads = 256;
constexpr int kTile = 16;  // pixel tile edge of k_pixels (16 x 16 = 256)

enum WMode { W_FP = 0, W_BP = 1, W_BPW = 2 };

int g_forced_tile = 0;

struct Geo2 {
  int W, H, D, V, B, Bp, BT, Lmax, scan;
  float dx, dy, det, sid, sdd, shx, shy, bin;
};

// ---------------------------------------------------------------------------
// 1. per-view setup (FP64 boundary mapping)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_setup2d(
    Geo2 g, const float* __restrict__ angles, ViewGeo* __restrict__ views,
    double* __restrict__ edges, float* __restrict__ wfp,
    LineGeo* __restrict__ lines, float* __restrict__ pint) {
  const int v = blockIdx.x;
  const float thf = angles[v];
  const double th = static_cast<double>(thf);
  int xaxis, flip;
  choose_axis(thf, &xaxis, &flip);
  const double s = sin(th), c = cos(th);
  const double sid = g.sid, sdd = g.sdd, det = g.det, bin = g.bin;
  const double dx = g.dx, dy = g.dy, shx = g.shx, shy = g.shy;
  const double sx = -s * sid, sy = c * sid;
  const int D = g.D;
  if (threadIdx.x == 0) {
    ViewGeo vg;
    vg.sx = sx;
    vg.sy = sy;
    vg.s = s;
    vg.c = c;
    vg.theta_ref = th + bin;
    vg.xaxis = xaxis;
    vg.flip = flip;
    vg.sxf = static_cast<float>(sx);
    vg.syf = static_cast<float>(sy);
    views[v] = vg;
  }
  // Bin edges, walk-ordered (FanArcDisCPU2d.cpp:77-86; FanFlatDisCPU2d.cpp:
  // 72-84; ParaDisCPU2d.cpp:51-55).
  double* ev = edges + static_cast<size_t>(v) * (D + 1);
  for (int k = threadIdx.x; k <= D; k += blockDim.x) {
    const double off = flip ? (D / 2.0 - k) : (k - D / 2.0);
    double e;
    if (g.scan == DD_FAN_ARC) {
      const double t = tan(off * det + bin + th);
      e = xaxis ? sx + sy * t : sy + sx / t;
    } else if (g.scan == DD_FAN_FLAT) {
      const double u = off * (sid / sdd * det) + sid / sdd * bin;
      const double px = u * c, py = u * s;
      e = xaxis ? (sx * py - sy * px) / (py - sy)
                : (sy * px - sx * py) / (px - sx);
    } else {
      e = (off * det + bin) / (xaxis ? c : s);
    }
    ev[k] = e;
  }
  __syncthreads();
  // FP weight: perpendicular pixel size / (bin width on the axis * cosine of
  // the bin-centre ray to the axis normal) (FanArcDisCPU2d.cpp:87-93);
  // parallel beam dy/det or dx/det (ParaDisCPU2d.cpp:60,113).
  const double perp = xaxis ? dy : dx;
  for (int k = threadIdx.x; k < D; k += blockDim.x) {
    double w;
    if (g.scan == DD_PARA) {
      w = perp / det;
    } else {
      const double e0 = ev[k], e1 = ev[k + 1];
      const double mid = 0.5 * (e0 + e1);
      const double cs = xaxis ? fabs(sy) / sqrt((mid - sx) * (mid - sx) + sy * sy)
                              : fabs(sx) / sqrt((mid - sy) * (mid - sy) + sx * sx);
      w = perp / ((e1 - e0) * cs);
    }
    wfp[static_cast<size_t>(v) * D + k] = static_cast<float>(w);
  }
  // Pixel edges of each line on the axis (FanArcDisCPU2d.cpp:95-101,165-171;
  // ParaDisCPU2d.cpp:57-64,110-117).
  const int L = xaxis ? g.H : g.W;
  for (int l = threadIdx.x; l < L; l += blockDim.x) {
    double a0, p;
    if (xaxis) {
      const double yl = (g.H / 2.0 - l - 0.5) * dy + shy;
      const double x0 = -g.W / 2.0 * dx + shx, x1 = g.W / 2.0 * dx + shx;
      if (g.scan == DD_PARA) {
        a0 = x0 + yl * s / c;
        p = dx;
      } else {
        a0 = (sx * yl - sy * x0) / (yl - sy);
        const double a1 = (sx * yl - sy * x1) / (yl - sy);
        p = (a1 - a0) / g.W;
      }
    } else {
      const double xl = (l - g.W / 2.0 + 0.5) * dx + shx;
      const double y0 = -g.H / 2.0 * dy + shy, y1 = g.H / 2.0 * dy + shy;
      if (g.scan == DD_PARA) {
        a0 = y0 + xl * c / s;
        p = dy;
      } else {
        a0 = (sy * xl - sx * y0) / (xl - sx);
        const double a1 = (sy * xl - sx * y1) / (xl - sx);
        p = (a1 - a0) / g.H;
      }
    }
    LineGeo lg;
    lg.a0 = a0;
    lg.inv_p = 1.0 / p;
    lines[static_cast<size_t>(v) * g.Lmax + l] = lg;
    pint[static_cast<size_t>(v) * g.Lmax + l] = static_cast<float>(p);
  }
}

// ---------------------------------------------------------------------------
// 2. layout prologues
// ---------------------------------------------------------------------------
// x [B][H][W] -> xrow [H][W][Bp] (lines = rows) and
//               xcol [W][H][Bp] with xcol[l][i] = x[H-1-i][l] (lines = columns,
//               pixel index from the bottom, FanArcDisCPU2d.cpp:194).
template <int BT>
__global__ __launch_bounds__(kThreads) void k_interleave_image(
    Geo2 g, const float* __restrict__ x, float* __restrict__ xrow,
    float* __restrict__ xcol) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = blockIdx.y;
  const int b0 = blockIdx.z * BT;
  if (w >= g.W) return;
  float v[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b)
    v[b] = (b0 + b < g.B)
               ? x[(static_cast<size_t>(b0 + b) * g.H + h) * g.W + w]
               : 0.0f;
  store_tile<BT>(xrow + (static_cast<size_t>(h) * g.W + w) * g.Bp + b0, v);
  store_tile<BT>(xcol + (static_cast<size_t>(w) * g.H + (g.H - 1 - h)) * g.Bp + b0, v);
}

// p [B][V][D] -> s [V][D][Bp] in walk order, optionally scaled by the FP
// weight of the bin (the adjoint's left factor).
template <int BT, bool FOLD_W>
__global__ __launch_bounds__(kThreads) void k_interleave_sino(
    Geo2 g, const ViewGeo* __restrict__ views, const float* __restrict__ wfp,
    const float* __restrict__ p, float* __restrict__ s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  const int b0 = blockIdx.z * BT;
  if (k >= g.D) return;
  const int kk = views[v].flip ? g.D - 1 - k : k;
  const float wk = FOLD_W ? wfp[static_cast<size_t>(v) * g.D + k] : 1.0f;
  float val[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b)
    val[b] = (b0 + b < g.B)
                 ? p[(static_cast<size_t>(b0 + b) * g.V + v) * g.D + kk] * wk
                 : 0.0f;
  store_tile<BT>(s + (static_cast<size_t>(v) * g.D + k) * g.Bp + b0, val);
}

// ---------------------------------------------------------------------------
// weights shared by both gather kernels
// ---------------------------------------------------------------------------
// BackProj constant: arc SID*dgamma (FanArcDisCPU2d.cpp:451), flat virtual
// pitch (FanFlatDisCPU2d.cpp:478), parallel det (ParaDisCPU2d.cpp:202); the
// reference divides it by the pixel interval and multiplies by the overlap on
// the axis, which is this constant times the overlap in pixel units.
__host__ __device__ inline float bp_const(const Geo2& g) {
  if (g.scan == DD_FAN_ARC) return g.sid * g.det;
  if (g.scan == DD_FAN_FLAT) return g.sid / g.sdd * g.det;
  return g.det;
}

// FBP distance weight of a pixel centre: arc 1/|S-P|^2
// (FanArcDisCPU2d.cpp:27-29,507), flat (r/(r-d))^2 (FanFlatDisCPU2d.cpp:19-22).
__device__ __forceinline__ float fbp_weight(const Geo2& g, const ViewGeo& vg,
                                            float xc, float yc) {
  if (g.scan == DD_FAN_ARC) {
    const float ddx = vg.sxf - xc, ddy = vg.syf - yc;
    return 1.0f / (ddx * ddx + ddy * ddy);
  }
  const float r = g.sid;
  const float d = (vg.sxf * xc + vg.syf * yc) / r;
  return (r * r) / ((r - d) * (r - d));
}

// Pixel centre of (line, pixel) in world coordinates.
__device__ __forceinline__ void pixel_centre(const Geo2& g, int xaxis, int l,
                                             int i, float* xc, float* yc) {
  if (xaxis) {
    *xc = (i - g.W * 0.5f + 0.5f) * g.dx + g.shx;
    *yc = (g.H * 0.5f - l - 0.5f) * g.dy + g.shy;
  } else {
    *xc = (l - g.W * 0.5f + 0.5f) * g.dx + g.shx;
    *yc = (i - g.H * 0.5f + 0.5f) * g.dy + g.shy;
  }
}

// ---------------------------------------------------------------------------
// 3a. image -> sinogram: one thread per (view, bin), BT instances
// ---------------------------------------------------------------------------
template <int BT, int WM>
__global__ __launch_bounds__(kThreads) void k_bins(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const float* __restrict__ wfp, const LineGeo* __restrict__ lines,
    const float* __restrict__ pint, const float* __restrict__ xrow,
    const float* __restrict__ xcol, float* __restrict__ out) {
  const int v = blockIdx.x;
  const int k = blockIdx.y * blockDim.x + threadIdx.x;
  const int b0 = blockIdx.z * BT;
  if (k >= g.D) return;
  const ViewGeo vg = views[v];
  const int xaxis = vg.xaxis;
  const int L = xaxis ? g.H : g.W;
  const int N = xaxis ? g.W : g.H;
  const float* __restrict__ src = (xaxis ? xrow : xcol) + b0;
  const double e1 = edges[static_cast<size_t>(v) * (g.D + 1) + k];
  const double e2 = edges[static_cast<size_t>(v) * (g.D + 1) + k + 1];
  const LineGeo* __restrict__ lv = lines + static_cast<size_t>(v) * g.Lmax;
  const float* __restrict__ pv = pint + static_cast<size_t>(v) * g.Lmax;
  const float kbp = bp_const(g);
  float acc[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) acc[b] = 0.0f;

  for (int l = 0; l < L; ++l) {
    const LineGeo lg = lv[l];
    const EdgePos p1 = edge_pos(e1, lg.a0, lg.inv_p, N);
    const EdgePos p2 = edge_pos(e2, lg.a0, lg.inv_p, N);
    const int cs = max(p1.c, 0);
    const int ce = min(p2.c, N - 1);
    if (cs > ce) continue;
    const float lw = (WM == W_FP) ? pv[l] : kbp;
    const float* __restrict__ line = src + static_cast<size_t>(l) * N * g.Bp;
    for (int c = cs; c <= ce; ++c) {
      float w = overlap_in_pixel(p1, p2, c) * lw;
      if constexpr (WM == W_BPW) {
        float xc, yc;
        pixel_centre(g, xaxis, l, c, &xc, &yc);
        w *= fbp_weight(g, vg, xc, yc);
      }
      float xv[BT];
      load_tile<BT>(line + static_cast<size_t>(c) * g.Bp, xv);
#pragma unroll
      for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, xv[b], acc[b]);
    }
  }
  const int kk = vg.flip ? g.D - 1 - k : k;
  const float wk = (WM == W_FP) ? wfp[static_cast<size_t>(v) * g.D + k] : 1.0f;
#pragma unroll
  for (int b = 0; b < BT; ++b)
    if (b0 + b < g.B)
      out[(static_cast<size_t>(b0 + b) * g.V + v) * g.D + kk] = acc[b] * wk;
}

// ---------------------------------------------------------------------------
// 3b. sinogram -> image: one thread per pixel, BT instances, all views
// ---------------------------------------------------------------------------
// First walk-ordered bin worth testing for an axis point: the inverse of the
// edge map, closed form per geometry.  Only a starting guess; the exact range
// is then fixed against the edge table with the shared edge_pos().
__device__ __forceinline__ int guess_bin(const Geo2& g, const ViewGeo& vg,
                                         double a) {
  double kf;
  if (g.scan == DD_FAN_ARC) {
    double gam = vg.xaxis ? atan2(a - vg.sx, vg.sy) : atan2(vg.sx, a - vg.sy);
    double d = gam - vg.theta_ref;
    d -= M_PI * rint(d / M_PI);
    kf = g.D * 0.5 + (vg.flip ? -d : d) / static_cast<double>(g.det);
  } else if (g.scan == DD_FAN_FLAT) {
    const double R = g.sid;
    const double u = vg.xaxis ? R * a * vg.c / (R + vg.s * a)
                              : R * a * vg.s / (R - vg.c * a);
    const double vdet = static_cast<double>(g.sid) / g.sdd * g.det;
    const double vbin = static_cast<double>(g.sid) / g.sdd * g.bin;
    kf = g.D * 0.5 + (vg.flip ? -(u - vbin) : (u - vbin)) / vdet;
  } else {
    const double u = vg.xaxis ? a * vg.c : a * vg.s;
    kf = g.D * 0.5 + (vg.flip ? -(u - g.bin) : (u - g.bin)) / static_cast<double>(g.det);
  }
  kf = fmin(fmax(kf, 0.0), g.D - 1.0);
  return static_cast<int>(kf);
}

template <int BT, int WM>
__global__ __launch_bounds__(kThreads) void k_pixels(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out) {
  const int c = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int r = blockIdx.y * kTile + (threadIdx.x / kTile);
  const int b0 = blockIdx.z * BT;
  if (c >= g.W || r >= g.H) return;
  const float kbp = bp_const(g);
  float xcw = 0.0f, ycw = 0.0f;
  if constexpr (WM == W_BPW) pixel_centre(g, 1, r, c, &xcw, &ycw);
  float acc[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) acc[b] = 0.0f;

  for (int v = 0; v < g.V; ++v) {
    const ViewGeo vg = views[v];
    int l, i, N;
    if (vg.xaxis) {
      l = r;
      i = c;
      N = g.W;
    } else {
      l = c;
      i = g.H - 1 - r;
      N = g.H;
    }
    const LineGeo lg = lines[static_cast<size_t>(v) * g.Lmax + l];
    const float pl = pint[static_cast<size_t>(v) * g.Lmax + l];
    const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
    int k = guess_bin(g, vg, lg.a0 + (i + 0.5) * static_cast<double>(pl));
    while (k > 0 && edge_pos(ev[k], lg.a0, lg.inv_p, N).c >= i) --k;
    while (k < g.D && edge_pos(ev[k + 1], lg.a0, lg.inv_p, N).c < i) ++k;
    if (k >= g.D) continue;
    float gw = (WM == W_FP) ? pl : kbp;
    if constexpr (WM == W_BPW) gw *= fbp_weight(g, vg, xcw, ycw);
    EdgePos p1 = edge_pos(ev[k], lg.a0, lg.inv_p, N);
    const float* __restrict__ sv = sino + static_cast<size_t>(v) * g.D * g.Bp + b0;
    while (p1.c <= i) {
      const EdgePos p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
      const float w = overlap_in_pixel(p1, p2, i) * gw;
      float q[BT];
      load_tile<BT>(sv + static_cast<size_t>(k) * g.Bp, q);
#pragma unroll
      for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, q[b], acc[b]);
      p1 = p2;
      if (++k >= g.D) break;
    }
  }
#pragma unroll
  for (int b = 0; b < BT; ++b)
    if (b0 + b < g.B)
      out[(static_cast<size_t>(b0 + b) * g.H + r) * g.W + c] = acc[b];
}

// ---------------------------------------------------------------------------
// system matrix (COO, FP weights): count pass and fill pass of one walk
// ---------------------------------------------------------------------------
// Entry (v*D + detector, row*W + col) = overlap on the axis * FP weight,
// the values GetSysMatrixFanArc2dKernel emits (FanArcDisCPU2d.cpp:1130-1149).
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_sysmatrix(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const float* __restrict__ wfp, const LineGeo* __restrict__ lines,
    const float* __restrict__ pint, int64_t* __restrict__ counts,
    const int64_t* __restrict__ offsets, int64_t* __restrict__ rows,
    int64_t* __restrict__ cols, float* __restrict__ vals) {
  const int v = blockIdx.x;
  const int k = blockIdx.y * blockDim.x + threadIdx.x;
  if (k >= g.D) return;
  const ViewGeo vg = views[v];
  const int xaxis = vg.xaxis;
  const int L = xaxis ? g.H : g.W;
  const int N = xaxis ? g.W : g.H;
  const double e1 = edges[static_cast<size_t>(v) * (g.D + 1) + k];
  const double e2 = edges[static_cast<size_t>(v) * (g.D + 1) + k + 1];
  const int kk = vg.flip ? g.D - 1 - k : k;
  const size_t slot = static_cast<size_t>(v) * g.D + kk;
  const float wk = wfp[static_cast<size_t>(v) * g.D + k];
  int64_t n = 0;
  int64_t at = FILL ? offsets[slot] : 0;
  for (int l = 0; l < L; ++l) {
    const LineGeo lg = lines[static_cast<size_t>(v) * g.Lmax + l];
    const EdgePos p1 = edge_pos(e1, lg.a0, lg.inv_p, N);
    const EdgePos p2 = edge_pos(e2, lg.a0, lg.inv_p, N);
    const int cs = max(p1.c, 0);
    const int ce = min(p2.c, N - 1);
    if (cs > ce) continue;
    if constexpr (!FILL) {
      n += ce - cs + 1;
    } else {
      const float pl = pint[static_cast<size_t>(v) * g.Lmax + l];
      for (int c = cs; c <= ce; ++c) {
        const int row = xaxis ? l : g.H - 1 - c;
        const int col = xaxis ? c : l;
        rows[at] = static_cast<int64_t>(slot);
        cols[at] = static_cast<int64_t>(row) * g.W + col;
        vals[at] = overlap_in_pixel(p1, p2, c) * pl * wk;
        ++at;
      }
    }
  }
  if constexpr (!FILL) counts[slot] = n;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Layout {
  size_t views, edges, wfp, lines, pint, geo_total;  // geometry tables
  size_t buf0, buf1, data_total;                      // data buffers
};

constexpr size_t kAlign = 256;

int pick_tile(int batch) {
  if (g_forced_tile > 0) return g_forced_tile;
  int t = 1;
  while (t < batch && t < 16) t *= 2;
  return t;
}

bool to_sino(int op) {
  return op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW;
}

Layout make_layout(const Geo2& g, int op) {
  Layout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off = align_up(off + bytes, kAlign);
    return at;
  };
  L.views = take(sizeof(ViewGeo) * g.V);
  L.edges = take(sizeof(double) * g.V * (g.D + 1));
  L.wfp = take(sizeof(float) * g.V * g.D);
  L.lines = take(sizeof(LineGeo) * g.V * g.Lmax);
  L.pint = take(sizeof(float) * g.V * g.Lmax);
  L.geo_total = off;
  off = 0;
  if (to_sino(op)) {
    L.buf0 = take(sizeof(float) * static_cast<size_t>(g.H) * g.W * g.Bp);
    L.buf1 = take(sizeof(float) * static_cast<size_t>(g.H) * g.W * g.Bp);
  } else {
    L.buf0 = take(sizeof(float) * static_cast<size_t>(g.V) * g.D * g.Bp);
    L.buf1 = L.buf0;
  }
  L.data_total = off;
  return L;
}

bool fill_geo(Geo2* out, int scan, const dd_geom2d* g, int batch, int views,
              int tile) {
  if (!g || batch <= 0 || views <= 0) return false;
  if (g->width <= 0 || g->height <= 0 || g->num_det <= 0) return false;
  Geo2 q;
  q.W = g->width;
  q.H = g->height;
  q.D = g->num_det;
  q.V = views;
  q.B = batch;
  q.BT = tile;
  q.Bp = align_up(batch, tile);
  q.Lmax = std::max(q.W, q.H);
  q.scan = scan;
  q.dx = g->pix_x;
  q.dy = g->pix_y;
  q.det = g->det;
  q.sid = g->iso_source;
  q.sdd = g->source_detector;
  q.shx = g->shift_x;
  q.shy = g->shift_y;
  q.bin = g->bin_shift;
  *out = q;
  return true;
}

struct GeoPtrs {
  const ViewGeo* views;
  const double* edges;
  const float* wfp;
  const LineGeo* lines;
  const float* pint;
};

GeoPtrs geo_ptrs(const void* geo, const Layout& L) {
  const char* b = static_cast<const char*>(geo);
  return {reinterpret_cast<const ViewGeo*>(b + L.views),
          reinterpret_cast<const double*>(b + L.edges),
          reinterpret_cast<const float*>(b + L.wfp),
          reinterpret_cast<const LineGeo*>(b + L.lines),
          reinterpret_cast<const float*>(b + L.pint)};
}

void launch_setup(const Geo2& g, const float* angles, void* geo,
                  const Layout& L, hipStream_t st) {
  char* b = static_cast<char*>(geo);
  hipLaunchKernelGGL(k_setup2d, dim3(g.V), dim3(kThreads), 0, st, g, angles,
                     reinterpret_cast<ViewGeo*>(b + L.views),
                     reinterpret_cast<double*>(b + L.edges),
                     reinterpret_cast<float*>(b + L.wfp),
                     reinterpret_cast<LineGeo*>(b + L.lines),
                     reinterpret_cast<float*>(b + L.pint));
}

template <int BT>
dd_status launch_tile(const Geo2& g, int op, const float* input, float* output,
                      const GeoPtrs& G, char* ws, const Layout& L,
                      hipStream_t st) {
  auto* buf0 = reinterpret_cast<float*>(ws + L.buf0);
  auto* buf1 = reinterpret_cast<float*>(ws + L.buf1);
  const int ntile = g.Bp / BT;
  if (to_sino(op)) {
    hipLaunchKernelGGL(k_interleave_image<BT>,
                       dim3((g.W + kThreads - 1) / kThreads, g.H, ntile),
                       dim3(kThreads), 0, st, g, input, buf0, buf1);
    dim3 grid(g.V, (g.D + kThreads - 1) / kThreads, ntile);
    ProfileScope prof(DD_KERNEL_IMAGE_TO_SINO, st);
    if (op == DD_PROJ)
      hipLaunchKernelGGL((k_bins<BT, W_FP>), grid, dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.wfp, G.lines, G.pint, buf0, buf1, output);
    else if (op == DD_BPROJ_T)
      hipLaunchKernelGGL((k_bins<BT, W_BP>), grid, dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.wfp, G.lines, G.pint, buf0, buf1, output);
    else
      hipLaunchKernelGGL((k_bins<BT, W_BPW>), grid, dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.wfp, G.lines, G.pint, buf0, buf1, output);
  } else {
    dim3 sgrid((g.D + kThreads - 1) / kThreads, g.V, ntile);
    if (op == DD_PROJ_T)
      hipLaunchKernelGGL((k_interleave_sino<BT, true>), sgrid, dim3(kThreads),
                         0, st, g, G.views, G.wfp, input, buf0);
    else
      hipLaunchKernelGGL((k_interleave_sino<BT, false>), sgrid, dim3(kThreads),
                         0, st, g, G.views, G.wfp, input, buf0);
    dim3 grid((g.W + kTile - 1) / kTile, (g.H + kTile - 1) / kTile, ntile);
    ProfileScope prof(DD_KERNEL_SINO_TO_IMAGE, st);
    if (op == DD_PROJ_T)
      hipLaunchKernelGGL((k_pixels<BT, W_FP>), grid, dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.lines, G.pint, buf0, output);
    else if (op == DD_BPROJ)
      hipLaunchKernelGGL((k_pixels<BT, W_BP>), grid, dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.lines, G.pint, buf0, output);
    else
      hipLaunchKernelGGL((k_pixels<BT, W_BPW>), grid, dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.lines, G.pint, buf0, output);
  }
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

dd_status run_tiles(const Geo2& q, int op, const float* input, float* output,
                    const GeoPtrs& G, char* w, const Layout& L, hipStream_t st) {
  switch (q.BT) {
    case 1: return launch_tile<1>(q, op, input, output, G, w, L, st);
    case 2: return launch_tile<2>(q, op, input, output, G, w, L, st);
    case 4: return launch_tile<4>(q, op, input, output, G, w, L, st);
    case 8: return launch_tile<8>(q, op, input, output, G, w, L, st);
    default: return launch_tile<16>(q, op, input, output, G, w, L, st);
  }
}

dd_status check_op(int scan, int op) {
  if (scan < DD_PARA || scan > DD_FAN_ARC) return DD_BAD_ARG;
  if (op < DD_PROJ || op > DD_BPROJ_TW) return DD_BAD_ARG;
  if (scan == DD_PARA && op >= DD_BPROJ_W) return DD_UNSUPPORTED;
  return DD_OK;
}

}  // namespace
}  // namespace dd

using namespace dd;

extern "C" {

const char* dd_version(void) { return "libdd 0.1 (gfx950)"; }

void dd2d_set_instance_tile(int tile) {
  g_forced_tile = (tile == 1 || tile == 2 || tile == 4 || tile == 8 || tile == 16) ? tile : 0;
}

size_t dd2d_workspace_size(int scan, int op, const dd_geom2d* g, int batch,
                           int views) {
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, pick_tile(batch))) return 0;
  const Layout L = make_layout(q, op);
  return L.geo_total + L.data_total;
}

dd_status dd2d_run(int scan, int op, const dd_geom2d* g, const float* input,
                   float* output, const float* angles, int batch, int views,
                   void* ws, size_t ws_bytes, void* stream) {
  dd_status rc = check_op(scan, op);
  if (rc != DD_OK) return rc;
  if (!input || !output || !angles) return DD_BAD_ARG;
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, pick_tile(batch))) return DD_BAD_ARG;
  const Layout L = make_layout(q, op);
  if (!ws || ws_bytes < L.geo_total + L.data_total) return DD_NO_WORKSPACE;
  auto st = static_cast<hipStream_t>(stream);
  char* w = static_cast<char*>(ws);
  launch_setup(q, angles, w, L, st);
  return run_tiles(q, op, input, output, geo_ptrs(w, L), w + L.geo_total, L, st);
}

size_t dd2d_geometry_size(int scan, const dd_geom2d* g, int views) {
  Geo2 q;
  if (!fill_geo(&q, scan, g, 1, views, 1)) return 0;
  return make_layout(q, DD_PROJ).geo_total;
}

dd_status dd2d_geometry(int scan, const dd_geom2d* g, const float* angles,
                        int views, void* geo, size_t geo_bytes, void* stream) {
  dd_status rc = check_op(scan, DD_PROJ);
  if (rc != DD_OK) return rc;
  Geo2 q;
  if (!angles || !fill_geo(&q, scan, g, 1, views, 1)) return DD_BAD_ARG;
  const Layout L = make_layout(q, DD_PROJ);
  if (!geo || geo_bytes < L.geo_total) return DD_NO_WORKSPACE;
  launch_setup(q, angles, geo, L, static_cast<hipStream_t>(stream));
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

size_t dd2d_planned_workspace_size(int scan, int op, const dd_geom2d* g,
                                   int batch, int views) {
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, pick_tile(batch))) return 0;
  return make_layout(q, op).data_total;
}

dd_status dd2d_run_planned(int scan, int op, const dd_geom2d* g,
                           const void* geo, const float* input, float* output,
                           int batch, int views, void* ws, size_t ws_bytes,
                           void* stream) {
  dd_status rc = check_op(scan, op);
  if (rc != DD_OK) return rc;
  if (!geo || !input || !output) return DD_BAD_ARG;
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, pick_tile(batch))) return DD_BAD_ARG;
  const Layout L = make_layout(q, op);
  if (!ws || ws_bytes < L.data_total) return DD_NO_WORKSPACE;
  return run_tiles(q, op, input, output, geo_ptrs(geo, L),
                   static_cast<char*>(ws), L, static_cast<hipStream_t>(stream));
}

}  // extern "C"

namespace {
// Geometry tables only (no data buffers) for the system matrix.
bool sys_setup(int scan, const dd_geom2d* g, const float* angles, int views,
               void* ws, size_t ws_bytes, hipStream_t st, Geo2* q, Layout* L) {
  if (scan < DD_PARA || scan > DD_FAN_ARC || !angles) return false;
  if (!fill_geo(q, scan, g, 1, views, 1)) return false;
  *L = make_layout(*q, DD_PROJ_T);
  if (!ws || ws_bytes < L->geo_total) return false;
  char* w = static_cast<char*>(ws);
  hipLaunchKernelGGL(k_setup2d, dim3(q->V), dim3(kThreads), 0, st, *q, angles,
                     reinterpret_cast<ViewGeo*>(w + L->views),
                     reinterpret_cast<double*>(w + L->edges),
                     reinterpret_cast<float*>(w + L->wfp),
                     reinterpret_cast<LineGeo*>(w + L->lines),
                     reinterpret_cast<float*>(w + L->pint));
  return true;
}
}  // namespace

extern "C" {

size_t dd2d_sysmatrix_workspace_size(int scan, const dd_geom2d* g, int views) {
  Geo2 q;
  if (!fill_geo(&q, scan, g, 1, views, 1)) return 0;
  return make_layout(q, DD_PROJ_T).geo_total;
}

dd_status dd2d_sysmatrix_count(int scan, const dd_geom2d* g,
                               const float* angles, int views, int64_t* counts,
                               void* ws, size_t ws_bytes, void* stream) {
  auto st = static_cast<hipStream_t>(stream);
  Geo2 q;
  Layout L;
  if (!counts || !sys_setup(scan, g, angles, views, ws, ws_bytes, st, &q, &L))
    return DD_BAD_ARG;
  char* w = static_cast<char*>(ws);
  hipLaunchKernelGGL((k_sysmatrix<false>), dim3(q.V, (q.D + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, st, q,
                     reinterpret_cast<const ViewGeo*>(w + L.views),
                     reinterpret_cast<const double*>(w + L.edges),
                     reinterpret_cast<const float*>(w + L.wfp),
                     reinterpret_cast<const LineGeo*>(w + L.lines),
                     reinterpret_cast<const float*>(w + L.pint), counts,
                     nullptr, nullptr, nullptr, nullptr);
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

dd_status dd2d_sysmatrix_fill(int scan, const dd_geom2d* g,
                              const float* angles, int views,
                              const int64_t* offsets, int64_t* rows,
                              int64_t* cols, float* vals, void* ws,
                              size_t ws_bytes, void* stream) {
  auto st = static_cast<hipStream_t>(stream);
  Geo2 q;
  Layout L;
  if (!offsets || !rows || !cols || !vals ||
      !sys_setup(scan, g, angles, views, ws, ws_bytes, st, &q, &L))
    return DD_BAD_ARG;
  char* w = static_cast<char*>(ws);
  hipLaunchKernelGGL((k_sysmatrix<true>), dim3(q.V, (q.D + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, st, q,
                     reinterpret_cast<const ViewGeo*>(w + L.views),
                     reinterpret_cast<const double*>(w + L.edges),
                     reinterpret_cast<const float*>(w + L.wfp),
                     reinterpret_cast<const LineGeo*>(w + L.lines),
                     reinterpret_cast<const float*>(w + L.pint), nullptr,
                     offsets, rows, cols, vals);
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

}  // extern "C"
