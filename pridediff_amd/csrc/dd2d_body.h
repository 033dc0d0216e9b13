// Device-side bodies of the 2D distance-driven kernels, shared by the
// one-call kernels (dd2d.hip) and the persistent batched kernels
// (ddbatch.hip), so both paths execute the same arithmetic.
//
// Behaviour followed: deeprecon/cpu/dis2d/FanArcDisCPU2d.cpp (fan arc),
// FanFlatDisCPU2d.cpp (fan flat), ParaDisCPU2d.cpp (parallel).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "dd.h"
#include "dd_common.h"

namespace dd {
namespace body2d {


constexpr int kThreads = 256;
constexpr int kTile = 16;  // pixel tile edge of k_pixels (16 x 16 = 256)

enum WMode { W_FP = 0, W_BP = 1, W_BPW = 2 };

struct Geo2 {
  int W, H, D, V, B, Bp, BT, Lmax, scan;
  float dx, dy, det, sid, sdd, shx, shy, bin;
};

// Output epilogue of an op (fused update steps): out = alpha * s * result
// + beta * c, s = *scale (a device scalar) or 1, c a device array shaped
// like the output or null.  The default leaves the result untouched, bit for
// bit.
struct Epi {
  const float* c = nullptr;
  const float* scale = nullptr;
  float alpha = 1.0f, beta = 0.0f;
};

__device__ __forceinline__ float epi_value(const Epi& e, float v, size_t at) {
  if (e.scale) v *= e.alpha * *e.scale;
  else if (e.alpha != 1.0f) v *= e.alpha;
  if (e.c) v = fmaf(e.beta, e.c[at], v);
  return v;
}

// ---------------------------------------------------------------------------
// 1. per-view setup (FP64 boundary mapping)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void setup2d_view(
    const Geo2& g, const float* __restrict__ angles, ViewGeo* __restrict__ views,
    double* __restrict__ edges, float* __restrict__ wfp,
    LineGeo* __restrict__ lines, float* __restrict__ pint, int v) {
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
    vg.cref = static_cast<float>(cos(th + bin));
    vg.sref = static_cast<float>(sin(th + bin));
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
// Instance-interleaved image copies for tiles of 8+ instances are quad-major,
// [Bp / 4][line][pixel][4]: the 4 float4 loads a lane makes for one pixel of
// a 16-instance tile then each read consecutive pixels across the wavefront
// (one 16-byte slice per pixel, the wave's pixels contiguous), instead of
// every load touching every pixel record's cache line.  Tiles of 1-4 keep
// [line][pixel][Bp].
template <int BT>
constexpr bool image_quad_major() {
  return BT >= 8;
}

// The sinogram copy of the sinogram -> image gathers likewise:
// [Bp / 4][view][bin][4] for tiles of 8+ instances, [view][bin][Bp] below.
template <int BT>
constexpr bool sino_quad_major() {
  return BT >= 8;
}

// View v of the sinogram copy for the tile starting at instance b0.
template <int BT>
__device__ __forceinline__ const float* sino_view(const Geo2& g, const float* sino, int v,
                                                  int b0) {
  if constexpr (sino_quad_major<BT>())
    return sino + (static_cast<size_t>(b0 / 4) * g.V * g.D + static_cast<size_t>(v) * g.D) * 4;
  else
    return sino + static_cast<size_t>(v) * g.D * g.Bp + b0;
}

// The BT values of bin k of a view returned by sino_view.
template <int BT>
__device__ __forceinline__ void load_sino(const Geo2& g, const float* __restrict__ sv, int k,
                                          float (&q)[BT]) {
  if constexpr (sino_quad_major<BT>()) {
    const size_t plane4 = static_cast<size_t>(g.V) * g.D * 4;
#pragma unroll
    for (int j = 0; j < BT / 4; ++j) {
      const float4 t =
          *reinterpret_cast<const float4*>(sv + j * plane4 + static_cast<size_t>(k) * 4);
      q[4 * j] = t.x;
      q[4 * j + 1] = t.y;
      q[4 * j + 2] = t.z;
      q[4 * j + 3] = t.w;
    }
  } else {
    load_tile<BT>(sv + static_cast<size_t>(k) * g.Bp, q);
  }
}

template <int BT>
__device__ __forceinline__ void load_image_px(const Geo2& g, const float* __restrict__ x, int h,
                                              int w, int b0, float (&v)[BT]) {
#pragma unroll
  for (int b = 0; b < BT; ++b)
    v[b] = (b0 + b < g.B)
               ? x[(static_cast<size_t>(b0 + b) * g.H + h) * g.W + w]
               : 0.0f;
}

// Pixel (h, w)'s BT values into the row copy (COL false) or the column copy
// (COL true, lines read bottom-up).
template <int BT, bool COL>
__device__ __forceinline__ void store_image_px(const Geo2& g, float* __restrict__ xc, int h,
                                               int w, int b0, const float (&v)[BT]) {
  const size_t px = COL ? static_cast<size_t>(w) * g.H + (g.H - 1 - h)
                        : static_cast<size_t>(h) * g.W + w;
  if constexpr (image_quad_major<BT>()) {
    const size_t plane = static_cast<size_t>(g.W) * g.H;  // pixels per quad plane
#pragma unroll
    for (int q = 0; q < BT / 4; ++q)
      *reinterpret_cast<float4*>(xc + (static_cast<size_t>(b0 / 4 + q) * plane + px) * 4) =
          make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  } else {
    store_tile<BT>(xc + px * g.Bp + b0, v);
  }
}

template <int BT>
__device__ __forceinline__ void interleave_image_px(
    const Geo2& g, const float* __restrict__ x, float* __restrict__ xrow,
    float* __restrict__ xcol, int h, int w, int b0) {
  float v[BT];
  load_image_px<BT>(g, x, h, w, b0, v);
  store_image_px<BT, false>(g, xrow, h, w, b0, v);
  store_image_px<BT, true>(g, xcol, h, w, b0, v);
}

// p [B][V][D] -> s [V][D][Bp] in walk order, optionally scaled by the FP
// weight of the bin (the adjoint's left factor).
template <int BT, bool FOLD_W>
__device__ __forceinline__ void interleave_sino_bin(
    const Geo2& g, const ViewGeo* __restrict__ views, const float* __restrict__ wfp,
    const float* __restrict__ p, float* __restrict__ s, int v, int k, int b0) {
  const int kk = views[v].flip ? g.D - 1 - k : k;
  const float wk = FOLD_W ? wfp[static_cast<size_t>(v) * g.D + k] : 1.0f;
  float val[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b)
    val[b] = (b0 + b < g.B)
                 ? p[(static_cast<size_t>(b0 + b) * g.V + v) * g.D + kk] * wk
                 : 0.0f;
  if constexpr (sino_quad_major<BT>()) {
#pragma unroll
    for (int j = 0; j < BT / 4; ++j)
      *reinterpret_cast<float4*>(
          s + ((static_cast<size_t>(b0 / 4 + j) * g.V + v) * g.D + k) * 4) =
          make_float4(val[4 * j], val[4 * j + 1], val[4 * j + 2], val[4 * j + 3]);
  } else {
    store_tile<BT>(s + (static_cast<size_t>(v) * g.D + k) * g.Bp + b0, val);
  }
}

// ---------------------------------------------------------------------------
// weights shared by both gather kernels
// ---------------------------------------------------------------------------
// BackProj constant: arc SID*dgamma (FanArcDisCPU2d.cpp:451), flat virtual
// pitch (FanFlatDisCPU2d.cpp:478), parallel det (ParaDisCPU2d.cpp:202); the
// reference divides it by the pixel interval and multiplies by the overlap on
// the axis, which is this constant times the overlap in pixel units.
__host__ __device__ __forceinline__ float bp_const(const Geo2& g) {
  if (g.scan == DD_FAN_ARC) return g.sid * g.det;
  if (g.scan == DD_FAN_FLAT) return g.sid / g.sdd * g.det;
  return g.det;
}

// FBP distance weight of a pixel centre: arc 1/|S-P|^2
// (FanArcDisCPU2d.cpp:27-29,507), flat (r/(r-d))^2 (FanFlatDisCPU2d.cpp:19-22).
// Every product that feeds a sum here and in pixel_centre is an explicit
// fmaf: left to contraction, which product the compiler fuses depends on
// the code around the inlined call, and every kernel must get the same bits.
__device__ __forceinline__ float fbp_weight(const Geo2& g, const ViewGeo& vg,
                                            float xc, float yc) {
  if (g.scan == DD_FAN_ARC) {
    const float ddx = vg.sxf - xc, ddy = vg.syf - yc;
    const float dy2 = ddy * ddy;
    return 1.0f / fmaf(ddx, ddx, dy2);
  }
  const float r = g.sid;
  const float syy = vg.syf * yc;
  const float d = fmaf(vg.sxf, xc, syy) / r;
  return (r * r) / ((r - d) * (r - d));
}

// Pixel centre of (line, pixel) in world coordinates.  The offsets
// (i - W/2 + 1/2 etc.) are exact in FP32, only the scaling rounds.
__device__ __forceinline__ void pixel_centre(const Geo2& g, int xaxis, int l,
                                             int i, float* xc, float* yc) {
  const int cx = xaxis ? i : l, ry = xaxis ? l : i;
  *xc = fmaf(static_cast<float>(cx) - g.W * 0.5f + 0.5f, g.dx, g.shx);
  *yc = xaxis ? fmaf(g.H * 0.5f - static_cast<float>(ry) - 0.5f, g.dy, g.shy)
              : fmaf(static_cast<float>(ry) - g.H * 0.5f + 0.5f, g.dy, g.shy);
}

// ---------------------------------------------------------------------------
// 3a. image -> sinogram: one thread per (view, bin), BT instances
// ---------------------------------------------------------------------------
// Image -> sinogram gathers split each bin's line loop into kSplit chunks
// (one per XCD) and add the chunk sums in chunk order in a second pass.
constexpr int kSplit = 8;

// Forward gathers run 63 bins per wavefront: consecutive lanes hold
// consecutive bins, every lane evaluates the position of its lower edge on
// each line and takes its upper edge's from the next lane (a DPP wave shift,
// the same bits), and lane 63 only serves lane 62's upper edge.  Each edge is
// evaluated once per line instead of twice.
constexpr int kBinsPerWave = 63;
// Tiles from this size up share pixels between forward lanes (below).
constexpr int kSharedPixelMinBT = 2;

__device__ __forceinline__ EdgePos edge_from_next_lane(EdgePos p) {
  constexpr int kWaveShl1 = 0x130;  // DPP wave_shl:1: lane i reads lane i + 1
  EdgePos q;
  q.c = __builtin_amdgcn_update_dpp(0, p.c, kWaveShl1, 0xf, 0xf, false);
  q.f = __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, p.f), kWaveShl1, 0xf, 0xf,
                                         false));
  return q;
}

// Sum over lines [l0, l1) of one bin's overlaps with the image (no bin
// weight); e1 is the bin's lower edge.  The whole wavefront calls this
// together (every lane active: the DPP reads the next lane); `edge_lane`
// marks lane 63.
template <int BT, int WM>
__device__ __forceinline__ void bins_accumulate(
    const Geo2& g, const ViewGeo& vg, double e1, const LineGeo* __restrict__ lv,
    const float* __restrict__ pv, const float* __restrict__ src, float kbp, int l0, int l1,
    int N, float (&acc)[BT], bool edge_lane) {
  const int xaxis = vg.xaxis;
  if (l0 >= l1) return;
  // the next line's geometry is loaded while this line is worked on
  LineGeo lg_next = lv[l0];
  float lw_next = (WM == W_FP) ? const_load(pv + l0) : kbp;
  for (int l = l0; l < l1; ++l) {
    const LineGeo lg = lg_next;
    const float lw = lw_next;
    lg_next = lv[min(l + 1, l1 - 1)];
    if constexpr (WM == W_FP) lw_next = const_load(pv + min(l + 1, l1 - 1));
    const EdgePos p1 = edge_pos(e1, lg.a0, lg.inv_p, N);
    const EdgePos p2 = edge_from_next_lane(p1);
    const float* __restrict__ line = src + static_cast<size_t>(l) * N *
                                               (image_quad_major<BT>() ? 4 : g.Bp);
    auto load_px = [&](int c, float (&xv)[BT]) {
      if constexpr (image_quad_major<BT>()) {
        const size_t plane4 = static_cast<size_t>(g.W) * g.H * 4;
#pragma unroll
        for (int q = 0; q < BT / 4; ++q) {
          const float4 t = *reinterpret_cast<const float4*>(line + q * plane4 +
                                                            static_cast<size_t>(c) * 4);
          xv[4 * q] = t.x;
          xv[4 * q + 1] = t.y;
          xv[4 * q + 2] = t.z;
          xv[4 * q + 3] = t.w;
        }
      } else {
        load_tile<BT>(line + static_cast<size_t>(c) * g.Bp, xv);
      }
    };
    auto weight = [&](int c) {
      float w = overlap_in_pixel(p1, p2, c) * lw;
      if constexpr (WM == W_BPW) {
        float xc, yc;
        pixel_centre(g, xaxis, l, c, &xc, &yc);
        w *= fbp_weight(g, vg, xc, yc);
      }
      return w;
    };
    const int cs = max(p1.c, 0);
    if constexpr (BT >= kSharedPixelMinBT) {
      // A bin's last pixel is the pixel of its upper edge, which is the next
      // bin's first pixel: every lane (lane 63 too) loads its first pixel and
      // takes its last one from the next lane, so each pixel of the line is
      // loaded once per wavefront instead of once per bin that reaches it.
      float x0[BT];
      if (p1.c <= N - 1) {
        load_px(cs, x0);
      } else {
#pragma unroll
        for (int b = 0; b < BT; ++b) x0[b] = 0.0f;
      }
      float xn[BT];
#pragma unroll
      for (int b = 0; b < BT; ++b)
        xn[b] = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                                              0, __builtin_bit_cast(int, x0[b]), 0x130, 0xf,
                                              0xf, false));
      if (edge_lane) continue;
      const int ce = min(p2.c, N - 1);
      if (cs > ce) continue;
      {
        const float w = weight(cs);
#pragma unroll
        for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, x0[b], acc[b]);
      }
      if (ce > cs) {
        for (int c = cs + 1; c < ce; ++c) {  // footprints wider than two pixels
          float xv[BT];
          load_px(c, xv);
          const float w = weight(c);
#pragma unroll
          for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, xv[b], acc[b]);
        }
        // the next lane's first pixel is this bin's last whenever the upper
        // edge lies inside the image (p2.c == ce); past the image, load it
        if (p2.c != ce) load_px(ce, xn);
        const float w = weight(ce);
#pragma unroll
        for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, xn[b], acc[b]);
      }
    } else {
      if (edge_lane) continue;
      const int ce = min(p2.c, N - 1);
      if (cs > ce) continue;
      for (int c = cs; c <= ce; ++c) {
        const float w = weight(c);
        float xv[BT];
        load_px(c, xv);
#pragma unroll
        for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, xv[b], acc[b]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3a'. image -> sinogram through running sums along the lines (one instance
// per geometry, no per-pixel weight)
// ---------------------------------------------------------------------------
// A bin's integral over a line is F(t2) - F(t1), F the running sum of the
// line's pixels, linear inside each pixel: F(t) = C_c + x_c (t - c),
// C_c = x_0 + ... + x_{c-1}, t in pixel units clamped to [0, n].  Each
// (bin, line) is then one lookup instead of a walk over the 1-3 pixels the
// bin covers.  C is accumulated in FP64 and stored as two floats, hi = fl(C)
// and lo = fl(C - hi), beside the pixel itself, {hi_c, lo_c, x_c, 0} (one
// aligned 16-byte load), so differences of nearby sums lose nothing to
// cancellation: F(t2) - F(t1) = (hi2 - hi1) + (m2 - m1) with m = lo + x f,
// where hi2 - hi1 rounds once, relative to the result.  A line of n pixels
// holds n + 1 entries (entry n: the total, x = 0).
__host__ __device__ constexpr int prefix_entries(int n) { return n + 1; }

// Running sums of one line of n pixels x[0], x[stride], ... into
// out[0 .. n], by one workgroup of blockDim.x threads (a power of two
// <= 1024), each owning a contiguous segment.
__device__ __forceinline__ void prefix_line(const float* __restrict__ x, long long stride,
                                            int n, float4* __restrict__ out,
                                            double* __restrict__ lds) {
  const int t = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
  const int seg = (n + nt - 1) / nt;
  const int p0 = min(t * seg, n), p1 = min(p0 + seg, n);
  double s = 0.0;
  for (int p = p0; p < p1; ++p) s += static_cast<double>(x[p * stride]);
  lds[t] = s;
  __syncthreads();
  // inclusive scan of the segment sums (Hillis-Steele)
  for (int d = 1; d < nt; d *= 2) {
    const double add = (t >= d) ? lds[t - d] : 0.0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  double c = lds[t] - s;  // sum of the segments before this one
  auto entry = [](double v, float xv) {
    const float hi = static_cast<float>(v);
    return make_float4(hi, static_cast<float>(v - static_cast<double>(hi)), xv, 0.0f);
  };
  for (int p = p0; p < p1; ++p) {
    const float xv = x[p * stride];
    out[p] = entry(c, xv);
    c += static_cast<double>(xv);
  }
  if (t == nt - 1) out[n] = entry(lds[nt - 1], 0.0f);
  __syncthreads();  // lds is reused by the next line
}

// Sum over lines [l0, l1) of one bin's integrals over the lines (no bin
// weight), from the lines' running-sum tables tab[line][n + 1]; e1 is the
// bin's lower edge, the upper edge's sum comes from the next lane.  The whole
// wavefront calls this together (the DPP reads the next lane); lane 63's sum
// is meaningless and not written.
template <int WM>
__device__ __forceinline__ float bins_accumulate_prefix(
    double e1, const LineGeo* __restrict__ lv, const float* __restrict__ pv,
    const float4* __restrict__ tab, float kbp, int l0, int l1, int n) {
  float acc = 0.0f;
  if (l0 >= l1) return acc;
  LineGeo lg_next = lv[l0];
  float lw_next = (WM == W_FP) ? const_load(pv + l0) : kbp;
  const float4* __restrict__ row = tab + static_cast<size_t>(l0) * prefix_entries(n);
  for (int l = l0; l < l1; ++l, row += prefix_entries(n)) {
    const LineGeo lg = lg_next;
    const float lw = lw_next;
    lg_next = lv[min(l + 1, l1 - 1)];
    if constexpr (WM == W_FP) lw_next = const_load(pv + min(l + 1, l1 - 1));
    const double t = fmin(fmax((e1 - lg.a0) * lg.inv_p, 0.0), static_cast<double>(n));
    const double fl = floor(t);
    const float f = static_cast<float>(t - fl);
    const float4 e = row[static_cast<int>(fl)];
    const float m = fmaf(f, e.z, e.y);
    constexpr int kWaveShl1 = 0x130;  // DPP wave_shl:1: lane i reads lane i + 1
    const float hn = __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, e.x), kWaveShl1, 0xf, 0xf,
                                           false));
    const float mn = __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, m), kWaveShl1, 0xf, 0xf,
                                           false));
    acc = fmaf((hn - e.x) + (mn - m), lw, acc);
  }
  return acc;
}

// ---------------------------------------------------------------------------
// 3b. sinogram -> image: one thread per pixel, BT instances, all views
// ---------------------------------------------------------------------------
// First walk-ordered bin worth testing for an axis point a: the inverse of
// the edge map, closed form per geometry, in FP32 with hardware reciprocals
// (v_rcp_f32) instead of IEEE divisions.  Only a starting guess; the exact
// bin is then fixed against the edge table with edge_pos().
__device__ __forceinline__ int guess_bin_fast(const Geo2& g, const ViewGeo& vg,
                                              double a) {
  float kf;
  if (g.scan == DD_FAN_ARC) {
    // Fan angle of the ray through a relative to the detector centre,
    // d = gamma - theta_ref (mod pi): tan d = (Y cos - X sin) / (X cos + Y sin)
    // with (X, Y) the ray direction; atan by the odd polynomial of
    // Abramowitz & Stegun 4.4.49 (|error| <= 1.2e-5 on [-1, 1]).
    const float Y = vg.xaxis ? static_cast<float>(a - vg.sx) : vg.sxf;
    const float X = vg.xaxis ? vg.syf : static_cast<float>(a - vg.sy);
    const float num = Y * vg.cref - X * vg.sref;
    const float den = X * vg.cref + Y * vg.sref;
    const float t = fminf(fmaxf(num * __builtin_amdgcn_rcpf(den), -1.0f), 1.0f);
    const float t2 = t * t;
    const float d =
        t * (0.9998660f +
             t2 * (-0.3302995f + t2 * (0.1801410f + t2 * (-0.0851330f + t2 * 0.0208351f))));
    kf = g.D * 0.5f + (vg.flip ? -d : d) * __builtin_amdgcn_rcpf(g.det);
  } else if (g.scan == DD_FAN_FLAT) {
    const float R = g.sid, af = static_cast<float>(a);
    const float sf = static_cast<float>(vg.s), cf = static_cast<float>(vg.c);
    const float u = vg.xaxis ? R * af * cf * __builtin_amdgcn_rcpf(R + sf * af)
                             : R * af * sf * __builtin_amdgcn_rcpf(R - cf * af);
    const float vdet = g.sid / g.sdd * g.det, vbin = g.sid / g.sdd * g.bin;
    kf = g.D * 0.5f + (vg.flip ? -(u - vbin) : (u - vbin)) * __builtin_amdgcn_rcpf(vdet);
  } else {
    const float u = static_cast<float>(vg.xaxis ? a * vg.c : a * vg.s);
    kf = g.D * 0.5f + (vg.flip ? -(u - g.bin) : (u - g.bin)) * __builtin_amdgcn_rcpf(g.det);
  }
  kf = fminf(fmaxf(kf, -4.0f), g.D + 4.0f);
  return static_cast<int>(floorf(kf));
}

// Contribution of view v to pixel (r, c): find the bins whose edges bracket
// the pixel in its line (closed-form guess, fixed exactly with edge_pos),
// then accumulate overlap * weight * sinogram.
template <int BT, int WM>
__device__ __forceinline__ void pixel_one_view(
    const Geo2& g, const ViewGeo& vg, int v, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, int r, int c, int b0, float xcw, float ycw,
    float (&acc)[BT]) {
  // line geometry loaded in the branch, so a caller whose row r is
  // wave-uniform gets scalar loads on row-axis views
  int i, N;
  LineGeo lg;
  float pl;
  if (vg.xaxis) {
    i = c;
    N = g.W;
    lg = lines[static_cast<size_t>(v) * g.Lmax + r];
    pl = pint[static_cast<size_t>(v) * g.Lmax + r];
  } else {
    i = g.H - 1 - r;
    N = g.H;
    lg = lines[static_cast<size_t>(v) * g.Lmax + c];
    pl = pint[static_cast<size_t>(v) * g.Lmax + c];
  }
  const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
  // k = the last bin whose lower edge lies before pixel i (or 0): unique,
  // whatever the guess.  The positions of edges k and k + 1 found on the way
  // start the walk.
  int k = min(max(guess_bin_fast(g, vg, lg.a0 + (i + 0.5) * static_cast<double>(pl)), 0),
              g.D - 1);
  EdgePos p1 = edge_pos(ev[k], lg.a0, lg.inv_p, N);
  while (k > 0 && p1.c >= i) p1 = edge_pos(ev[--k], lg.a0, lg.inv_p, N);
  EdgePos p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  while (p2.c < i) {
    if (++k >= g.D) return;
    p1 = p2;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
  float gw = (WM == W_FP) ? pl : bp_const(g);
  if constexpr (WM == W_BPW) gw *= fbp_weight(g, vg, xcw, ycw);
  const float* __restrict__ sv = sino_view<BT>(g, sino, v, b0);
  while (p1.c <= i) {
    const float w = overlap_in_pixel(p1, p2, i) * gw;
    float q[BT];
    load_sino<BT>(g, sv, k, q);
#pragma unroll
    for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, q[b], acc[b]);
    if (++k >= g.D) break;
    p1 = p2;
    if (p1.c > i) break;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
}

// First-bin guess for the pixel centred at (xc, yc) from the view alone: the
// axis coordinate of the ray from the source through the centre (the line's
// pixel-edge map evaluated at the centre), in FP32.  It needs no per-line
// table, so a view's search can be started before the previous view's walk.
// Only a guess: the search below it is exact whatever it returns.
__device__ __forceinline__ int guess_bin_centre(const Geo2& g, const ViewGeo& vg, float xc,
                                                float yc) {
  float a;
  if (g.scan == DD_PARA) {
    const float s = static_cast<float>(vg.s), c = static_cast<float>(vg.c);
    a = vg.xaxis ? xc + yc * s * __builtin_amdgcn_rcpf(c) : yc + xc * c * __builtin_amdgcn_rcpf(s);
  } else {
    a = vg.xaxis ? (vg.sxf * yc - vg.syf * xc) * __builtin_amdgcn_rcpf(yc - vg.syf)
                 : (vg.syf * xc - vg.sxf * yc) * __builtin_amdgcn_rcpf(xc - vg.sxf);
  }
  return min(max(guess_bin_fast(g, vg, static_cast<double>(a)), 0), g.D - 1);
}

// What a pixel's search in one view starts from: the line's pixel-edge map,
// the guessed bin and that bin's two edges, all loaded ahead.
struct PixViewStart {
  LineGeo lg;
  float pl;
  int k0;
  double ea, eb;  // edges k0 and k0 + 1
};

__device__ __forceinline__ PixViewStart pixel_view_start(
    const Geo2& g, const ViewGeo& vg, int v, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint, int r, int c, float xc,
    float yc) {
  PixViewStart s;
  // a branch per axis, so a caller whose row r is wave-uniform gets scalar
  // loads on row-axis views
  if (vg.xaxis) {
    s.lg = lines[static_cast<size_t>(v) * g.Lmax + r];
    s.pl = pint[static_cast<size_t>(v) * g.Lmax + r];
  } else {
    s.lg = lines[static_cast<size_t>(v) * g.Lmax + c];
    s.pl = pint[static_cast<size_t>(v) * g.Lmax + c];
  }
  s.k0 = guess_bin_centre(g, vg, xc, yc);
  const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
  s.ea = ev[s.k0];
  s.eb = ev[s.k0 + 1];
  return s;
}

// pixel_one_view from a start loaded ahead: the same search (edge k is
// P(k) = edge_pos(ev[k]); going down, the old lower edge becomes the upper
// one), the same walk.  Same bins, same order, same values: equal bit for
// bit.
template <int BT, int WM>
__device__ __forceinline__ void pixel_view_from(
    const Geo2& g, const ViewGeo& vg, int v, const double* __restrict__ edges,
    const PixViewStart& s, const float* __restrict__ sino, int r, int c, int b0, float xcw,
    float ycw, float (&acc)[BT]) {
  const int i = vg.xaxis ? c : g.H - 1 - r;
  const int N = vg.xaxis ? g.W : g.H;
  const LineGeo lg = s.lg;
  const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
  int k = s.k0;
  EdgePos p1 = edge_pos(s.ea, lg.a0, lg.inv_p, N);
  EdgePos p2 = edge_pos(s.eb, lg.a0, lg.inv_p, N);
  while (k > 0 && p1.c >= i) {
    p2 = p1;
    p1 = edge_pos(ev[--k], lg.a0, lg.inv_p, N);
  }
  while (p2.c < i) {
    if (++k >= g.D) return;
    p1 = p2;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
  float gw = (WM == W_FP) ? s.pl : bp_const(g);
  if constexpr (WM == W_BPW) gw *= fbp_weight(g, vg, xcw, ycw);
  const float* __restrict__ sv = sino_view<BT>(g, sino, v, b0);
  while (p1.c <= i) {
    const float w = overlap_in_pixel(p1, p2, i) * gw;
    float q[BT];
    load_sino<BT>(g, sv, k, q);
#pragma unroll
    for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, q[b], acc[b]);
    if (++k >= g.D) break;
    p1 = p2;
    if (p1.c > i) break;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
}

// Contributions of view v to the two consecutive pixels i0, i0 + 1 of one
// line (i0 + 1 only when has1): one first-bin search for i0, then one walk
// over the bins that reach either pixel, each sinogram tile loaded once.  A
// bin contributes to pixel i exactly when its edge cells bracket i
// (p1.c <= i <= p2.c), which is the set and order of bins pixel_one_view
// walks for i, so each pixel's sum is the same bit for bit.
template <int BT, int WM>
__device__ __forceinline__ void pixel_pair_view(
    const Geo2& g, const ViewGeo& vg, int v, const double* __restrict__ edges,
    const LineGeo& lg, float pl, const float* __restrict__ sino, int i0, bool has1, int N,
    int b0, float gw0, float gw1, float (&acc0)[BT], float (&acc1)[BT]) {
  const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
  const int i1 = i0 + 1, ilast = has1 ? i1 : i0;
  int k = min(max(guess_bin_fast(g, vg, lg.a0 + (i0 + 0.5) * static_cast<double>(pl)), 0),
              g.D - 1);
  EdgePos p1 = edge_pos(ev[k], lg.a0, lg.inv_p, N);
  while (k > 0 && p1.c >= i0) p1 = edge_pos(ev[--k], lg.a0, lg.inv_p, N);
  EdgePos p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  while (p2.c < i0) {
    if (++k >= g.D) return;
    p1 = p2;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
  const float* __restrict__ sv = sino_view<BT>(g, sino, v, b0);
  while (p1.c <= ilast) {
    float q[BT];
    load_sino<BT>(g, sv, k, q);
    if (p1.c <= i0 && i0 <= p2.c) {
      const float w = overlap_in_pixel(p1, p2, i0) * gw0;
#pragma unroll
      for (int b = 0; b < BT; ++b) acc0[b] = fmaf(w, q[b], acc0[b]);
    }
    if (has1 && p1.c <= i1 && i1 <= p2.c) {
      const float w = overlap_in_pixel(p1, p2, i1) * gw1;
#pragma unroll
      for (int b = 0; b < BT; ++b) acc1[b] = fmaf(w, q[b], acc1[b]);
    }
    if (++k >= g.D) break;
    p1 = p2;
    if (p1.c > ilast) break;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
}

// sinogram -> image for the 2 x 2 pixel block at (r0, c0), all views: per
// view the block is two lines of two consecutive pixels whichever the axis
// (image rows c0, c0 + 1 on rows r0, r0 + 1; or columns, read bottom-up).
// Equal bit for bit to pixels_body on each of the four pixels.
template <int BT, int WM>
__device__ __forceinline__ void quads_body(
    const Geo2& g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out, int r0, int c0, int b0,
    const Epi& ep = Epi{}) {
  const bool hr = r0 + 1 < g.H, hc = c0 + 1 < g.W;
  float xc[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}}, yc[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  if constexpr (WM == W_BPW) {
#pragma unroll
    for (int dr = 0; dr < 2; ++dr)
#pragma unroll
      for (int dc = 0; dc < 2; ++dc) pixel_centre(g, 1, r0 + dr, c0 + dc, &xc[dr][dc], &yc[dr][dc]);
  }
  float acc[2][2][BT];  // [row offset][column offset][instance]
#pragma unroll
  for (int dr = 0; dr < 2; ++dr)
#pragma unroll
    for (int dc = 0; dc < 2; ++dc)
#pragma unroll
      for (int b = 0; b < BT; ++b) acc[dr][dc][b] = 0.0f;
  const float kbp = bp_const(g);
  for (int v = 0; v < g.V; ++v) {
    const ViewGeo vg = views[v];
    const LineGeo* __restrict__ lv = lines + static_cast<size_t>(v) * g.Lmax;
    const float* __restrict__ pv = pint + static_cast<size_t>(v) * g.Lmax;
    if (vg.xaxis) {
      // lines = image rows r0, r0 + 1; pixels c0, c0 + 1
#pragma unroll
      for (int dr = 0; dr < 2; ++dr) {
        if (dr == 1 && !hr) break;
        const int l = r0 + dr;
        const float pl = pv[l];
        float gw0 = (WM == W_FP) ? pl : kbp, gw1 = gw0;
        if constexpr (WM == W_BPW) {
          gw0 *= fbp_weight(g, vg, xc[dr][0], yc[dr][0]);
          gw1 *= fbp_weight(g, vg, xc[dr][1], yc[dr][1]);
        }
        pixel_pair_view<BT, WM>(g, vg, v, edges, lv[l], pl, sino, c0, hc, g.W, b0, gw0, gw1,
                                acc[dr][0], acc[dr][1]);
      }
    } else {
      // lines = image columns c0, c0 + 1; pixels i = H - 1 - r: the pair
      // (i0, i0 + 1) is rows (r0 + 1, r0), or row r0 alone on the last row
#pragma unroll
      for (int dc = 0; dc < 2; ++dc) {
        if (dc == 1 && !hc) break;
        const int l = c0 + dc;
        const float pl = pv[l];
        float gw0 = (WM == W_FP) ? pl : kbp, gw1 = gw0;
        if (hr) {
          if constexpr (WM == W_BPW) {
            gw0 *= fbp_weight(g, vg, xc[1][dc], yc[1][dc]);
            gw1 *= fbp_weight(g, vg, xc[0][dc], yc[0][dc]);
          }
          pixel_pair_view<BT, WM>(g, vg, v, edges, lv[l], pl, sino, g.H - 2 - r0, true, g.H,
                                  b0, gw0, gw1, acc[1][dc], acc[0][dc]);
        } else {
          if constexpr (WM == W_BPW) gw0 *= fbp_weight(g, vg, xc[0][dc], yc[0][dc]);
          pixel_pair_view<BT, WM>(g, vg, v, edges, lv[l], pl, sino, g.H - 1 - r0, false, g.H,
                                  b0, gw0, gw1, acc[0][dc], acc[1][dc]);
        }
      }
    }
  }
#pragma unroll
  for (int dr = 0; dr < 2; ++dr)
#pragma unroll
    for (int dc = 0; dc < 2; ++dc)
      if ((dr == 0 || hr) && (dc == 0 || hc))
#pragma unroll
        for (int b = 0; b < BT; ++b)
          if (b0 + b < g.B) {
            const size_t at = (static_cast<size_t>(b0 + b) * g.H + r0 + dr) * g.W + c0 + dc;
            out[at] = epi_value(ep, acc[dr][dc][b], at);
          }
}

// Contributions of view v to the run of NP consecutive pixels ib .. ib+NP-1
// of one line, of which positions pf .. pl are real pixels: one first-bin
// search for pixel ib + pf, then one walk over the bins that reach any of
// them, each sinogram tile loaded once.  Per pixel the same bins in the same
// order as pixel_one_view: equal bit for bit.
template <int BT, int WM, int NP>
__device__ __forceinline__ void pixel_run_view(
    const Geo2& g, const ViewGeo& vg, int v, const double* __restrict__ edges,
    const LineGeo& lg, float pl, const float* __restrict__ sino, int ib, int pf, int plst,
    int N, int b0, const float (&gw)[NP], float (&acc)[NP][BT]) {
  const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
  const int ifirst = ib + pf, ilast = ib + plst;
  int k = min(max(guess_bin_fast(g, vg, lg.a0 + (ifirst + 0.5) * static_cast<double>(pl)), 0),
              g.D - 1);
  EdgePos p1 = edge_pos(ev[k], lg.a0, lg.inv_p, N);
  while (k > 0 && p1.c >= ifirst) p1 = edge_pos(ev[--k], lg.a0, lg.inv_p, N);
  EdgePos p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  while (p2.c < ifirst) {
    if (++k >= g.D) return;
    p1 = p2;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
  const float* __restrict__ sv = sino_view<BT>(g, sino, v, b0);
  while (p1.c <= ilast) {
    float q[BT];
    load_sino<BT>(g, sv, k, q);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int i = ib + p;
      if (p >= pf && p <= plst && p1.c <= i && i <= p2.c) {
        const float w = overlap_in_pixel(p1, p2, i) * gw[p];
#pragma unroll
        for (int b = 0; b < BT; ++b) acc[p][b] = fmaf(w, q[b], acc[p][b]);
      }
    }
    if (++k >= g.D) break;
    p1 = p2;
    if (p1.c > ilast) break;
    p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
  }
}

// pixel_run_view with the bins read as one window: the K + 1 edges and K
// sinogram values from one bin below the first pixel's guess are loaded at
// once (independent loads, one round trip instead of one per bin), every
// bin of the window is tested against every pixel of the run, and the walk
// continues one bin at a time only if the run reaches past the window.  The
// window is used only when nothing below it can reach the run (its first
// edge lies before the run, or it starts at bin 0); otherwise the serial
// search decides.  Per pixel the same bins in the same (ascending) order,
// with the same overlaps: equal bit for bit.
template <int BT, int WM, int NP, int K>
__device__ __forceinline__ void pixel_run_view_win(
    const Geo2& g, const ViewGeo& vg, int v, const double* __restrict__ edges,
    const LineGeo& lg, float pl, const float* __restrict__ sino, int ib, int pf, int plst,
    int N, int b0, const float (&gw)[NP], float (&acc)[NP][BT]) {
  const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
  const int ifirst = ib + pf, ilast = ib + plst;
  if (g.D < K + 1) {
    pixel_run_view<BT, WM, NP>(g, vg, v, edges, lg, pl, sino, ib, pf, plst, N, b0, gw, acc);
    return;
  }
  static_assert(K % 4 == 0, "window of whole 4-bin groups");
  const int kg = guess_bin_fast(g, vg, lg.a0 + (ifirst + 0.5) * static_cast<double>(pl));
  // edges kw .. kw + K + 1 are read in pairs, so the window ends one edge
  // short of the table's last (index D)
  const int kw = min(max(kg - 1, 0), g.D - K - 1);
  double e[K + 2];
#pragma unroll
  for (int j = 0; j < K + 2; j += 2) {
    // 16-byte loads of edge pairs (8-byte aligned)
    double2 pr;
    __builtin_memcpy(&pr, ev + kw + j, sizeof(pr));
    e[j] = pr.x;
    e[j + 1] = pr.y;
  }
  const float* __restrict__ sv = sino_view<BT>(g, sino, v, b0);
  float q[K][BT];
  if constexpr (BT == 1 && !sino_quad_major<BT>()) {
    if (g.Bp == 1) {
      // 16-byte loads of four bins (4-byte aligned)
#pragma unroll
      for (int j = 0; j < K; j += 4) {
        float4 t;
        __builtin_memcpy(&t, sv + kw + j, sizeof(t));
        q[j][0] = t.x;
        q[j + 1][0] = t.y;
        q[j + 2][0] = t.z;
        q[j + 3][0] = t.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j) load_sino<BT>(g, sv, kw + j, q[j]);
    }
  } else {
#pragma unroll
    for (int j = 0; j < K; ++j) load_sino<BT>(g, sv, kw + j, q[j]);
  }
  EdgePos p1 = edge_pos(e[0], lg.a0, lg.inv_p, N);
  if (kw > 0 && p1.c >= ifirst) {
    // a bin below the window may reach the run: the serial search decides
    pixel_run_view<BT, WM, NP>(g, vg, v, edges, lg, pl, sino, ib, pf, plst, N, b0, gw, acc);
    return;
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const EdgePos p2 = edge_pos(e[j + 1], lg.a0, lg.inv_p, N);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int i = ib + p;
      if (p >= pf && p <= plst && p1.c <= i && i <= p2.c) {
        const float w = overlap_in_pixel(p1, p2, i) * gw[p];
#pragma unroll
        for (int b = 0; b < BT; ++b) acc[p][b] = fmaf(w, q[j][b], acc[p][b]);
      }
    }
    p1 = p2;
  }
  // past the window, one bin at a time (the walk's own loop)
  int k = kw + K;
  if (k >= g.D) return;
  while (p1.c <= ilast) {
    const EdgePos p2 = edge_pos(ev[k + 1], lg.a0, lg.inv_p, N);
    float qk[BT];
    load_sino<BT>(g, sv, k, qk);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int i = ib + p;
      if (p >= pf && p <= plst && p1.c <= i && i <= p2.c) {
        const float w = overlap_in_pixel(p1, p2, i) * gw[p];
#pragma unroll
        for (int b = 0; b < BT; ++b) acc[p][b] = fmaf(w, qk[b], acc[p][b]);
      }
    }
    if (++k >= g.D) break;
    p1 = p2;
  }
}

// sinogram -> image for the Q x Q pixel block at (r0, c0), all views: per
// view the block is Q lines of Q consecutive pixels whichever the axis (rows
// over columns c0 .. c0+Q-1, or columns read bottom-up).  Equal bit for bit
// to pixels_body on each pixel.  KW > 0: each line's bins are read as one
// window of KW bins (pixel_run_view_win).
template <int BT, int WM, int Q, int KW = 0>
__device__ __forceinline__ void blocks_body(
    const Geo2& g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out, int r0, int c0, int b0) {
  const int nr = min(Q, g.H - r0), nc = min(Q, g.W - c0);
  // pixel centres once, before the view loop, as pixels_body takes them
  float xc[Q][Q], yc[Q][Q];
  if constexpr (WM == W_BPW) {
#pragma unroll
    for (int dr = 0; dr < Q; ++dr)
#pragma unroll
      for (int dc = 0; dc < Q; ++dc)
        pixel_centre(g, 1, min(r0 + dr, g.H - 1), min(c0 + dc, g.W - 1), &xc[dr][dc],
                     &yc[dr][dc]);
  }
  float acc[Q][Q][BT];  // [row offset][column offset][instance]
#pragma unroll
  for (int dr = 0; dr < Q; ++dr)
#pragma unroll
    for (int dc = 0; dc < Q; ++dc)
#pragma unroll
      for (int b = 0; b < BT; ++b) acc[dr][dc][b] = 0.0f;
  const float kbp = bp_const(g);
  for (int v = 0; v < g.V; ++v) {
    const ViewGeo vg = views[v];
    const LineGeo* __restrict__ lv = lines + static_cast<size_t>(v) * g.Lmax;
    const float* __restrict__ pv = pint + static_cast<size_t>(v) * g.Lmax;
    if (vg.xaxis) {
      // lines = image rows r0 + dr; run position p = column c0 + p
#pragma unroll
      for (int dr = 0; dr < Q; ++dr) {
        if (dr >= nr) break;
        const int l = r0 + dr;
        const float pl = pv[l];
        float gw[Q];
#pragma unroll
        for (int p = 0; p < Q; ++p) {
          gw[p] = (WM == W_FP) ? pl : kbp;
          if constexpr (WM == W_BPW) gw[p] *= fbp_weight(g, vg, xc[dr][p], yc[dr][p]);
        }
        if constexpr (KW > 0)
          pixel_run_view_win<BT, WM, Q, KW>(g, vg, v, edges, lv[l], pl, sino, c0, 0, nc - 1,
                                            g.W, b0, gw, acc[dr]);
        else
          pixel_run_view<BT, WM, Q>(g, vg, v, edges, lv[l], pl, sino, c0, 0, nc - 1, g.W, b0,
                                    gw, acc[dr]);
      }
    } else {
      // lines = image columns c0 + dc; pixel i = H - 1 - r, so run position p
      // (pixel H - r0 - Q + p) is row r0 + Q - 1 - p, real when p >= Q - nr
      float run[Q][BT];
#pragma unroll
      for (int dc = 0; dc < Q; ++dc) {
        if (dc >= nc) break;
        const int l = c0 + dc;
        const float pl = pv[l];
        float gw[Q];
#pragma unroll
        for (int p = 0; p < Q; ++p) {
          gw[p] = (WM == W_FP) ? pl : kbp;
          if constexpr (WM == W_BPW)
            gw[p] *= fbp_weight(g, vg, xc[Q - 1 - p][dc], yc[Q - 1 - p][dc]);
#pragma unroll
          for (int b = 0; b < BT; ++b) run[p][b] = acc[Q - 1 - p][dc][b];
        }
        if constexpr (KW > 0)
          pixel_run_view_win<BT, WM, Q, KW>(g, vg, v, edges, lv[l], pl, sino, g.H - r0 - Q,
                                            Q - nr, Q - 1, g.H, b0, gw, run);
        else
          pixel_run_view<BT, WM, Q>(g, vg, v, edges, lv[l], pl, sino, g.H - r0 - Q, Q - nr,
                                    Q - 1, g.H, b0, gw, run);
#pragma unroll
        for (int p = 0; p < Q; ++p)
#pragma unroll
          for (int b = 0; b < BT; ++b) acc[Q - 1 - p][dc][b] = run[p][b];
      }
    }
  }
#pragma unroll
  for (int dr = 0; dr < Q; ++dr)
#pragma unroll
    for (int dc = 0; dc < Q; ++dc)
      if (dr < nr && dc < nc)
#pragma unroll
        for (int b = 0; b < BT; ++b)
          if (b0 + b < g.B)
            out[(static_cast<size_t>(b0 + b) * g.H + r0 + dr) * g.W + c0 + dc] = acc[dr][dc][b];
}

// ---------------------------------------------------------------------------
// 3b'. sinogram -> image through running sums over the bins (one instance
// per geometry, no per-pixel weight)
// ---------------------------------------------------------------------------
// A pixel's value from one view is (G(i + 1) - G(i)) / P, G the running
// integral of the walk-ordered sinogram s (FP weight folded in) along the
// common axis, piecewise linear between the knots (e_k, R_k), R_k =
// sum_{j<k} s_j (e_{j+1} - e_j), k = 0 .. D.  The knots are FP64 (the running
// sum is kept as it is accumulated) and a view's table is {e_k, R_k}: one
// aligned 16-byte load per bin the walk visits, and no per-(bin, pixel)
// overlap terms at all.

// Knots of view v of one instance's sinogram p [V][D] (detector order) into
// that view's table tab[0 .. D] (walk order, s scaled by the FP weight when
// FOLD_W), by one workgroup of blockDim.x threads (a power of two <= 1024).
template <bool FOLD_W>
__device__ __forceinline__ void prefix_sino_view(
    const Geo2& g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const float* __restrict__ wfp, const float* __restrict__ p, int v, double2* __restrict__ tab,
    double* __restrict__ lds) {
  const int t = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
  const int D = g.D;
  const int flip = views[v].flip;
  const double* __restrict__ ev = edges + static_cast<size_t>(v) * (D + 1);
  const float* __restrict__ pv = p + static_cast<size_t>(v) * D;
  auto term = [&](int k) {
    const float wk = FOLD_W ? wfp[static_cast<size_t>(v) * D + k] : 1.0f;
    return static_cast<double>(pv[flip ? D - 1 - k : k] * wk) * (ev[k + 1] - ev[k]);
  };
  const int seg = (D + nt - 1) / nt;
  const int k0 = min(t * seg, D), k1 = min(k0 + seg, D);
  double s = 0.0;
  for (int k = k0; k < k1; ++k) s += term(k);
  lds[t] = s;
  __syncthreads();
  for (int d = 1; d < nt; d *= 2) {
    const double add = (t >= d) ? lds[t - d] : 0.0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  double c = lds[t] - s;
  for (int k = k0; k < k1; ++k) {
    tab[k] = make_double2(ev[k], c);
    c += term(k);
  }
  if (t == nt - 1) tab[D] = make_double2(ev[D], lds[nt - 1]);
  __syncthreads();
}

// Adds view v's values of the pixels ib + pf .. ib + plst of one line to
// acc[pf .. plst] (times gw, the line's weight): one first-bin search, then
// the run's pixel edges and the bin edges merged in one monotone walk, G at
// each pixel edge interpolated between the two knots around it.  Knot k is
// kn[k - kmin]; the walk stays within bins kmin .. kmax - 1, which must hold
// every bin the run's pixel edges fall in (then the result does not depend
// on the range).  KP is a global or an LDS pointer to the knots as pairs of
// doubles {e, R}.
template <typename KP>
__device__ __forceinline__ double2 knot_at(KP kn, int i) {
  return make_double2(kn[2 * i], kn[2 * i + 1]);
}

template <int NP, typename KP>
__device__ __forceinline__ void pixel_run_view_prefix(
    const Geo2& g, const ViewGeo& vg, const LineGeo& lg, float pl, float gw, KP kn, int kmin,
    int kmax, int ib, int pf, int plst, float (&acc)[NP]) {
  // Positions relative to the run's edge ib (tau = e / P - a0 / P - ib) and
  // running sums relative to the first knot read (rho = R - rbase) are
  // formed in FP64 once per knot and kept in FP32: over a run of NP pixels
  // both stay of the size of a few pixels and bins, so FP32 holds them to
  // its precision relative to the values the pixels receive.
  const double bb = fma(lg.a0, lg.inv_p, static_cast<double>(ib));
  auto tau = [&](double e) { return static_cast<float>(fma(e, lg.inv_p, -bb)); };
  // the bin whose lower knot lies at or before the run's first pixel edge
  // (guessed at that edge)
  int k = min(max(guess_bin_fast(g, vg, lg.a0 + (ib + pf) * static_cast<double>(pl)), kmin),
              kmax - 1);
  double2 q = knot_at(kn, k - kmin);
  const double rbase = q.y;
  float t1 = tau(q.x), r1 = 0.0f;
  const float fpf = static_cast<float>(pf);
  while (k > kmin && t1 > fpf) {
    --k;
    q = knot_at(kn, k - kmin);
    t1 = tau(q.x);
    r1 = static_cast<float>(q.y - rbase);
  }
  q = knot_at(kn, k + 1 - kmin);
  float t2 = tau(q.x), r2 = static_cast<float>(q.y - rbase);
  float rw = __builtin_amdgcn_rcpf(t2 - t1);
  const float scale = static_cast<float>(lg.inv_p) * gw;
  float gprev = 0.0f;
#pragma unroll
  for (int j = 0; j <= NP; ++j) {
    if (j < pf || j > plst + 1) continue;
    const float dj = static_cast<float>(j);
    while (k < kmax - 1 && t2 <= dj) {
      ++k;
      t1 = t2;
      r1 = r2;
      q = knot_at(kn, k + 1 - kmin);
      t2 = tau(q.x);
      r2 = static_cast<float>(q.y - rbase);
      rw = __builtin_amdgcn_rcpf(t2 - t1);
    }
    // fraction of bin k below the edge, clamped: 0 below the first knot,
    // 1 past the last
    const float fr = __builtin_amdgcn_fmed3f((dj - t1) * rw, 0.0f, 1.0f);
    const float gv = fmaf(r2 - r1, fr, r1);
    if (j > pf) acc[j - 1] = fmaf(gv - gprev, scale, acc[j - 1]);
    gprev = gv;
  }
}

// One view's contribution to the Q x Q pixel block at (r0, c0) (nr x nc of
// it inside the image) through the view's knots (pixel_run_view_prefix);
// the block's lines as in blocks_body.  WM is W_FP or W_BP.
template <int WM, int Q, typename KP>
__device__ __forceinline__ void block_view_prefix(
    const Geo2& g, const ViewGeo& vg, const LineGeo* __restrict__ lv,
    const float* __restrict__ pv, KP kn, int kmin, int kmax, int r0, int c0, int nr, int nc,
    float (&acc)[Q][Q]) {
  static_assert(WM != W_BPW, "per-pixel weights need the bin walk");
  const float kbp = bp_const(g);
  if (vg.xaxis) {
#pragma unroll
    for (int dr = 0; dr < Q; ++dr) {
      if (dr >= nr) break;
      const int l = r0 + dr;
      const float pl = const_load(pv + l);
      pixel_run_view_prefix<Q>(g, vg, lv[l], pl, (WM == W_FP) ? pl : kbp, kn, kmin, kmax, c0,
                               0, nc - 1, acc[dr]);
    }
  } else {
    // lines = image columns c0 + dc, pixels read bottom-up: run position p
    // (pixel H - r0 - Q + p) is row r0 + Q - 1 - p, real when p >= Q - nr
#pragma unroll
    for (int dc = 0; dc < Q; ++dc) {
      if (dc >= nc) break;
      const int l = c0 + dc;
      const float pl = const_load(pv + l);
      float run[Q];
#pragma unroll
      for (int p = 0; p < Q; ++p) run[p] = acc[Q - 1 - p][dc];
      pixel_run_view_prefix<Q>(g, vg, lv[l], pl, (WM == W_FP) ? pl : kbp, kn, kmin, kmax,
                               g.H - r0 - Q, Q - nr, Q - 1, run);
#pragma unroll
      for (int p = 0; p < Q; ++p) acc[Q - 1 - p][dc] = run[p];
    }
  }
}

// Knot range [kb, ke] of view v that holds every pixel edge of the pixel
// tile [r0, r1) x [c0, c1) (image rows x columns): the tile's axis extent is
// spanned by its corners (a pixel edge's axis point is monotone along and
// across the lines), widened by one knot each side against rounding.
__device__ __forceinline__ int2 knot_range(const Geo2& g, const ViewGeo& vg,
                                           const double* __restrict__ ev,
                                           const LineGeo* __restrict__ lv, int r0, int r1,
                                           int c0, int c1) {
  double amin = 1e300, amax = -1e300;
  auto corner = [&](int l, int i) {
    const LineGeo lg = lv[l];
    const double a = lg.a0 + i / lg.inv_p;
    amin = fmin(amin, a);
    amax = fmax(amax, a);
  };
  if (vg.xaxis) {  // lines = rows, pixel edges = columns
    corner(r0, c0), corner(r0, c1), corner(r1 - 1, c0), corner(r1 - 1, c1);
  } else {  // lines = columns, pixel edges from the bottom
    corner(c0, g.H - r1), corner(c0, g.H - r0), corner(c1 - 1, g.H - r1),
        corner(c1 - 1, g.H - r0);
  }
  // last knot <= amin, first knot >= amax (edges ascend in walk order)
  int lo = 0, hi = g.D;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (ev[mid] <= amin) lo = mid; else hi = mid - 1;
  }
  const int kb = max(lo - 1, 0);
  lo = 0, hi = g.D;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (ev[mid] >= amax) hi = mid; else lo = mid + 1;
  }
  const int ke = min(lo + 1, g.D);
  return make_int2(kb, max(ke, kb + 1));
}

// sinogram -> image for one pixel, all views.  Each view's search is
// started (line map, guess, first two edges loaded) before the previous
// view's walk, so those loads overlap the walk's instead of opening the
// view; per view the same bins in the same order as pixel_one_view.
template <int BT, int WM>
__device__ __forceinline__ void pixels_body(
    const Geo2& g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out, int r, int c, int b0,
    const Epi& ep = Epi{}) {
  float xcw = 0.0f, ycw = 0.0f;
  pixel_centre(g, 1, r, c, &xcw, &ycw);
  float acc[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) acc[b] = 0.0f;
  ViewGeo vg = views[0];
  PixViewStart cur = pixel_view_start(g, vg, 0, edges, lines, pint, r, c, xcw, ycw);
  for (int v = 0; v < g.V; ++v) {
    const int vn = min(v + 1, g.V - 1);
    const ViewGeo vgn = views[vn];
    const PixViewStart nxt = pixel_view_start(g, vgn, vn, edges, lines, pint, r, c, xcw, ycw);
    pixel_view_from<BT, WM>(g, vg, v, edges, cur, sino, r, c, b0, xcw, ycw, acc);
    vg = vgn;
    cur = nxt;
  }
#pragma unroll
  for (int b = 0; b < BT; ++b)
    if (b0 + b < g.B) {
      const size_t at = (static_cast<size_t>(b0 + b) * g.H + r) * g.W + c;
      out[at] = epi_value(ep, acc[b], at);
    }
}

// Edge positions staged per (line of a pixel tile, view).
template <int T>
struct PixelTileLds {
  static constexpr int kCap = 4 * T;  // staged edges per line
  int klo[T];                 // first staged edge of each tile line
  int n[T];                   // staged edges of each tile line
  int first[T][T + 1];        // first staged edge with c >= i0 + j
  int ec[T][kCap];            // edge_pos().c of staged edges
  float ef[T][kCap];          // edge_pos().f of staged edges
  int fallback[2];            // per view parity: window did not cover
};

// sinogram -> image for one kTile x kTile pixel tile, all views; every thread
// of the block must call it (barriers).  Per view, the tile's lines get
// their bin-edge positions computed once into LDS (instead of once per
// pixel), a first-edge table is scattered from them, and each pixel reads
// its bin range and overlaps from LDS.  The overlaps are the same edge_pos
// values the forward kernel uses, so A^T stays the exact transpose of A.
// A line whose staged window misses part of the tile takes the per-pixel
// path for that view (same arithmetic, slower).
template <int BT, int WM, int kTile>
__device__ __forceinline__ void pixels_tile_body(
    const Geo2& g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out, int tile_c, int tile_r,
    int b0, PixelTileLds<kTile>& s,
    const Epi& ep = Epi{}) {
  constexpr int kEdgeCap = PixelTileLds<kTile>::kCap;
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int c0 = tile_c * kTile, r0 = tile_r * kTile;
  const int c = c0 + tx, r = r0 + ty;
  const bool inside = (c < g.W) && (r < g.H);
  float xcw = 0.0f, ycw = 0.0f;
  if constexpr (WM == W_BPW) pixel_centre(g, 1, r, c, &xcw, &ycw);
  const float kbp = bp_const(g);
  float acc[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) acc[b] = 0.0f;
  if (threadIdx.x < 2) s.fallback[threadIdx.x] = 0;
  __syncthreads();

  for (int v = 0; v < g.V; ++v) {
    const ViewGeo vg = views[v];
    const int xaxis = vg.xaxis;
    const int L = xaxis ? g.H : g.W;
    const int N = xaxis ? g.W : g.H;
    const int lbase = xaxis ? r0 : c0;
    const int i0 = xaxis ? c0 : g.H - kTile - r0;
    const int par = v & 1;
    const double* __restrict__ ev = edges + static_cast<size_t>(v) * (g.D + 1);
    const LineGeo* __restrict__ lv = lines + static_cast<size_t>(v) * g.Lmax;
    // A: staged edge window of each tile line
    if (threadIdx.x == 0) s.fallback[par ^ 1] = 0;
    if (threadIdx.x < kTile) {
      const int la = lbase + threadIdx.x;
      int klo = 0, n = 0;
      if (la < L) {
        const LineGeo lg = lv[la];
        const double p = static_cast<double>(pint[static_cast<size_t>(v) * g.Lmax + la]);
        const int k0 = guess_bin_fast(g, vg, lg.a0 + i0 * p) - 2;
        const int k1 = guess_bin_fast(g, vg, lg.a0 + (i0 + kTile) * p) + 3;
        klo = min(max(k0, 0), g.D);
        const int khi = min(max(k1, klo), g.D);
        n = khi - klo + 1;
        if (n > kEdgeCap) {
          n = kEdgeCap;
          s.fallback[par] = 1;
        }
      }
      s.klo[threadIdx.x] = klo;
      s.n[threadIdx.x] = n;
#pragma unroll
      for (int j = 0; j <= kTile; ++j) s.first[threadIdx.x][j] = n;
    }
    __syncthreads();
    // B: edge positions of every staged edge, and the window checks
    for (int idx = threadIdx.x; idx < kTile * kEdgeCap; idx += blockDim.x) {
      const int ll = idx / kEdgeCap, e = idx % kEdgeCap;
      const int n = s.n[ll];
      if (e >= n) continue;
      const int klo = s.klo[ll];
      const LineGeo lg = lv[lbase + ll];
      const EdgePos ep = edge_pos(ev[klo + e], lg.a0, lg.inv_p, N);
      s.ec[ll][e] = ep.c;
      s.ef[ll][e] = ep.f;
      // the window must start before the tile's first real pixel and end
      // after its last one, unless it reaches the detector's end
      if (e == 0 && klo > 0 && ep.c >= max(i0, 0)) s.fallback[par] = 1;
      if (e == n - 1 && klo + e < g.D && ep.c < min(i0 + kTile, N)) s.fallback[par] = 1;
    }
    __syncthreads();
    // C: first staged edge at or after each pixel edge of the tile
    for (int idx = threadIdx.x; idx < kTile * kEdgeCap; idx += blockDim.x) {
      const int ll = idx / kEdgeCap, e = idx % kEdgeCap;
      if (e >= s.n[ll]) continue;
      const int ce = s.ec[ll][e];
      const int cp = (e == 0) ? -(1 << 30) : s.ec[ll][e - 1];
      const int jlo = max(cp + 1 - i0, 0), jhi = min(ce - i0, kTile);
      for (int j = jlo; j <= jhi; ++j) s.first[ll][j] = e;
    }
    __syncthreads();
    // D: accumulate
    if (inside) {
      if (!s.fallback[par]) {
        const int ll = xaxis ? ty : tx;
        const int j = xaxis ? tx : (kTile - 1 - ty);
        const int i = i0 + j;
        const int klo = s.klo[ll], n = s.n[ll];
        const int e0 = max(s.first[ll][j] - 1, 0);
        const int e1 = min(s.first[ll][j + 1] - 1, n - 2);
        float gw = (WM == W_FP) ? pint[static_cast<size_t>(v) * g.Lmax + lbase + ll] : kbp;
        if constexpr (WM == W_BPW) gw *= fbp_weight(g, vg, xcw, ycw);
        const float* __restrict__ sv = sino_view<BT>(g, sino, v, b0);
        for (int e = e0; e <= e1; ++e) {
          const EdgePos p1{s.ec[ll][e], s.ef[ll][e]};
          const EdgePos p2{s.ec[ll][e + 1], s.ef[ll][e + 1]};
          const float w = overlap_in_pixel(p1, p2, i) * gw;
          float q[BT];
          load_sino<BT>(g, sv, klo + e, q);
#pragma unroll
          for (int b = 0; b < BT; ++b) acc[b] = fmaf(w, q[b], acc[b]);
        }
      } else {
        pixel_one_view<BT, WM>(g, vg, v, edges, lines, pint, sino, r, c, b0, xcw, ycw, acc);
      }
    }
    __syncthreads();
  }
  if (inside) {
#pragma unroll
    for (int b = 0; b < BT; ++b)
      if (b0 + b < g.B) {
        const size_t at = (static_cast<size_t>(b0 + b) * g.H + r) * g.W + c;
        out[at] = epi_value(ep, acc[b], at);
      }
  }
}

}  // namespace body2d
}  // namespace dd
