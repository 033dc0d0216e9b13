// 2D distance-driven projectors (parallel, fan-flat, fan-arc) for gfx950.
//
// Behaviour followed: deeprecon/cpu/dis2d/FanArcDisCPU2d.cpp (fan arc),
// FanFlatDisCPU2d.cpp (fan flat), ParaDisCPU2d.cpp (parallel), and the op
// table of deeprecon/interface/RadonTorch.cpp:17-568.
//
// One call = four or five stream-ordered launches, no host synchronisation:
//   1. k_setup2d        per view: axis/flip choice, bin edges on the common
//                       axis (FP64), FP weights, per-line pixel-edge maps.
//   2. k_interleave_*   copies the input into an instance-interleaved layout
//                       ([line][pixel][b] or [view][bin][b]) so one geometry
//                       evaluation feeds BT instances through vector loads.
//   3. k_bins_split     image -> sinogram ops: one thread per detector bin and
//                       chunk of image lines gathers along its lines (no
//                       atomics); k_bins_reduce adds the chunks in order.
//      k_pixels         sinogram -> image ops: one thread per pixel gathers
//                       over every view (no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <mutex>

#include "dd.h"
#include "dd2d_body.h"
#include "dd_common.h"
#include "dd_profile.h"

namespace dd {
namespace {

using namespace body2d;

std::atomic<int> g_forced_tile{0};
// sinogram -> image kernel of dd2d_run and of plans without a variant:
// -1 = automatic: 2 x 2 quads for up to 4 instances per launch when they
// still give at least 4 workgroups per compute unit, per pixel otherwise.
std::atomic<int> g_adjoint_variant{-1};

// Compute units of the first device the library runs on (thread-safe static;
// every device of one process is the same part here).
int compute_units() {
  static const int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        n > 0)
      return n;
    return 256;
  }();
  return cus;
}

// A variant is a sinogram -> image kernel (DD_ADJ_*), optionally OR-ed with
// DD_FWD_PREFIX (image -> sinogram through running sums, tile 1).
bool known_variant(int v) {
  if (v < 0) return false;
  const int adj = v & ~DD_FWD_PREFIX;
  return (adj >= DD_ADJ_PER_PIXEL && adj <= DD_ADJ_QUAD) || adj == DD_ADJ_BLOCK_PREFIX;
}

// The automatic choice keeps the walks, so a slice run alone gives the bits
// it gets in any batch; the running-sum adjoint (DD_ADJ_BLOCK_PREFIX, tile 1
// only) is chosen explicitly or by the autotuner.
int adjoint_variant(int requested, const Geo2& g) {
  if (!known_variant(requested)) requested = g_adjoint_variant;
  // the running sums are per instance: instance tile 1 only
  int flags = (requested >= 0 && g.BT == 1) ? (requested & DD_FWD_PREFIX) : 0;
  int adj = requested >= 0 ? (requested & ~DD_FWD_PREFIX) : -1;
  if (adj == DD_ADJ_BLOCK_PREFIX && g.BT != 1) adj = -1;
  if (adj < 0) {
    const long long quad_groups = 1LL * ((g.W + 31) / 32) * ((g.H + 31) / 32) * (g.Bp / g.BT);
    adj = (g.BT <= 4 && quad_groups >= 4LL * compute_units()) ? DD_ADJ_QUAD : DD_ADJ_PER_PIXEL;
  }
  return adj | flags;
}

// Whether op runs through the running-sum adjoint (its weights per pixel
// keep BackProjWeighted on the walk).
bool prefix_path(const Geo2& g, int op, int variant) {
  return (variant & ~DD_FWD_PREFIX) == DD_ADJ_BLOCK_PREFIX && g.BT == 1 &&
         (op == DD_PROJ_T || op == DD_BPROJ);
}

// Whether op runs through the running-sum forward (BackProjTransWeighted
// keeps the walk).
bool fwd_prefix_path(const Geo2& g, int op, int variant) {
  return (variant & DD_FWD_PREFIX) != 0 && g.BT == 1 && (op == DD_PROJ || op == DD_BPROJ_T);
}

// Running-sum buffers a layout holds.
enum PrefixRegions { kPrefNone = 0, kPrefAdjoint = 1, kPrefForward = 2 };

int prefix_regions(const Geo2& g, int op, int variant) {
  return (prefix_path(g, op, variant) ? kPrefAdjoint : 0) |
         (fwd_prefix_path(g, op, variant) ? kPrefForward : 0);
}

// Pixel workgroup width of the per-pixel sinogram -> image gather (256
// threads): 64 x 4 pixels when 8+ instances share a launch (a wave is one
// image row, one line of every view on the row axis), 16 x 16 otherwise.
template <int BT>
constexpr int pix_tile_w() {
  return BT >= 8 ? 64 : 16;
}

// Workgroup shape of the planned gathers (dd_shape, validated): forward
// waves per workgroup and views per workgroup; per-pixel adjoint workgroup
// width (0 = pix_tile_w).  None of them changes a result bit.
struct Shape {
  int fwd_waves = kThreads / 64;
  int fwd_views = 1;
  int pix_width = 0;
};

bool make_shape(const dd_shape* in, Shape* out) {
  Shape s;
  if (in) {
    if (in->fwd_waves) {
      if (in->fwd_waves != 1 && in->fwd_waves != 2 && in->fwd_waves != 4) return false;
      s.fwd_waves = in->fwd_waves;
    }
    if (in->fwd_views) {
      if (in->fwd_views < 1 || in->fwd_views > 8) return false;
      s.fwd_views = in->fwd_views;
    }
    if (in->pix_width) {
      if (in->pix_width != 16 && in->pix_width != 32 && in->pix_width != 64) return false;
      s.pix_width = in->pix_width;
    }
  }
  *out = s;
  return true;
}

// ---------------------------------------------------------------------------
// one-call kernels: thin grids over the shared bodies
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_setup2d(
    Geo2 g, const float* __restrict__ angles, ViewGeo* __restrict__ views,
    double* __restrict__ edges, float* __restrict__ wfp,
    LineGeo* __restrict__ lines, float* __restrict__ pint) {
  setup2d_view(g, angles, views, edges, wfp, lines, pint, blockIdx.x);
}

// x [B][H][W] -> the row copy and the column copy, by 16 x 16 pixel tiles.
// The column copy runs down the image's columns, so each tile passes through
// LDS and is written transposed: a wave's lanes then write contiguous pixels
// of both copies (written row by row, neighbouring lanes' column-copy pixels
// lie H pixels apart).
constexpr int kItile = 16;

template <int BT>
__global__ __launch_bounds__(kItile * kItile) void k_interleave_image(
    Geo2 g, const float* __restrict__ x, float* __restrict__ xrow,
    float* __restrict__ xcol) {
  // pixel (r, c) of the tile, value b, at t[r * kRowPitch + c * (BT + 1) + b]:
  // odd strides, so neither the row-wise writes nor the column-wise reads
  // below pile onto one LDS bank
  constexpr int kRowPitch = kItile * (BT + 1) + 1;
  __shared__ float t[kItile * kRowPitch];
  const int tx = static_cast<int>(threadIdx.x) % kItile;
  const int ty = static_cast<int>(threadIdx.x) / kItile;
  const int h0 = blockIdx.y * kItile, w0 = blockIdx.x * kItile, b0 = blockIdx.z * BT;
  if (h0 + ty < g.H && w0 + tx < g.W) {
    float v[BT];
    load_image_px<BT>(g, x, h0 + ty, w0 + tx, b0, v);
    store_image_px<BT, false>(g, xrow, h0 + ty, w0 + tx, b0, v);
#pragma unroll
    for (int b = 0; b < BT; ++b) t[ty * kRowPitch + tx * (BT + 1) + b] = v[b];
  }
  __syncthreads();
  // transposed: consecutive lanes take consecutive rows of one image
  // column, which are consecutive pixels of the column copy
  if (h0 + tx < g.H && w0 + ty < g.W) {
    float v[BT];
#pragma unroll
    for (int b = 0; b < BT; ++b) v[b] = t[tx * kRowPitch + ty * (BT + 1) + b];
    store_image_px<BT, true>(g, xcol, h0 + tx, w0 + ty, b0, v);
  }
}

template <int BT, bool FOLD_W>
__global__ __launch_bounds__(kThreads) void k_interleave_sino(
    Geo2 g, const ViewGeo* __restrict__ views, const float* __restrict__ wfp,
    const float* __restrict__ p, float* __restrict__ s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.D) return;
  interleave_sino_bin<BT, FOLD_W>(g, views, wfp, p, s, blockIdx.y, k, blockIdx.z * BT);
}

// image -> sinogram: one thread per (view, bin, chunk of lines), BT
// instances.  Work item i has chunk i % kSplit: workgroups are dealt to the 8
// XCDs in turn, so each XCD's L2 only holds its chunk's lines of the image.
// Writes unweighted partial sums part[chunk][b][v][k] (walk order);
// k_bins_reduce adds them in chunk order.

// One view of k_bins_split: bin k's unweighted sum over the chunk's lines.
// Every lane stays to the end (the edge exchange needs the whole wave);
// lanes past the last bin work on clamped edges and write nothing.
template <int BT, int WM>
__device__ __forceinline__ void bins_split_view(
    const Geo2& g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ xrow, const float* __restrict__ xcol, float* __restrict__ part,
    int v, int chunk, int k, int b0, bool edge_lane) {
  const ViewGeo vg = views[v];
  const int L = vg.xaxis ? g.H : g.W;
  const int N = vg.xaxis ? g.W : g.H;
  const int l0 = L * chunk / kSplit, l1 = L * (chunk + 1) / kSplit;
  const double e1 = edges[static_cast<size_t>(v) * (g.D + 1) + min(k, g.D)];
  float acc[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) acc[b] = 0.0f;
  bins_accumulate<BT, WM>(g, vg, e1, lines + static_cast<size_t>(v) * g.Lmax,
                          pint + static_cast<size_t>(v) * g.Lmax,
                          (vg.xaxis ? xrow : xcol) +
                              (image_quad_major<BT>()
                                   ? static_cast<size_t>(b0 / 4) * g.W * g.H * 4
                                   : static_cast<size_t>(b0)),
                          bp_const(g), l0, l1, N, acc, edge_lane);
  if (edge_lane || k >= g.D) return;
#pragma unroll
  for (int b = 0; b < BT; ++b)
    if (b0 + b < g.B)
      part[((static_cast<size_t>(chunk) * g.Bp + b0 + b) * g.V + v) * g.D + k] = acc[b];
}

// The workgroup is a (views x bins x lines) tile: blockDim.x / 64 waves of 63
// bins, `vpb` consecutive views one after the other, the lines of one chunk.
template <int BT, int WM>
__global__ __launch_bounds__(kThreads) void k_bins_split(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ xrow, const float* __restrict__ xcol,
    float* __restrict__ part, int vpb) {
  const int bins_per_block = static_cast<int>(blockDim.x / 64) * kBinsPerWave;
  const int kb_n = (g.D + bins_per_block - 1) / bins_per_block;
  const int vb_n = (g.V + vpb - 1) / vpb;
  const int chunk = blockIdx.x % kSplit;
  int rest = blockIdx.x / kSplit;
  const int kb = rest % kb_n;
  rest /= kb_n;
  const int vb = rest % vb_n;
  const int b0 = (rest / vb_n) * BT;
  const int lane = static_cast<int>(threadIdx.x % 64);
  const int k = kb * bins_per_block + static_cast<int>(threadIdx.x / 64) * kBinsPerWave + lane;
  const bool edge_lane = lane == 63;
  for (int v = vb * vpb; v < min((vb + 1) * vpb, g.V); ++v)
    bins_split_view<BT, WM>(g, views, edges, lines, pint, xrow, xcol, part, v, chunk, k, b0,
                            edge_lane);
}


// part [kSplit][Bp][V][D] -> out [B][V][D]: chunk sums in order, times the
// bin's FP weight, detector order.
template <int WM>
__global__ __launch_bounds__(kThreads) void k_bins_reduce(
    Geo2 g, const ViewGeo* __restrict__ views, const float* __restrict__ wfp,
    const float* __restrict__ part, float* __restrict__ out, Epi ep) {
  const int kb_n = (g.D + kThreads - 1) / kThreads;
  const int k = (blockIdx.x % kb_n) * kThreads + threadIdx.x;
  const int v = (blockIdx.x / kb_n) % g.V, b = blockIdx.x / (kb_n * g.V);
  if (k >= g.D) return;
  const size_t plane = static_cast<size_t>(g.Bp) * g.V * g.D;
  const size_t at = (static_cast<size_t>(b) * g.V + v) * g.D + k;
  float sum = part[at];
#pragma unroll
  for (int c = 1; c < kSplit; ++c) sum += part[c * plane + at];
  const ViewGeo vg = views[v];
  const int kk = vg.flip ? g.D - 1 - k : k;
  const float wk = (WM == W_FP) ? wfp[static_cast<size_t>(v) * g.D + k] : 1.0f;
  const size_t o = (static_cast<size_t>(b) * g.V + v) * g.D + kk;
  out[o] = epi_value(ep, sum * wk, o);
}

// sinogram -> image: one thread per pixel, BT instances, all views.
// Variant DD_ADJ_PER_PIXEL: every pixel finds its bins on its own.
template <int BT, int WM, int TW>
__global__ __launch_bounds__(kThreads) void k_pixels(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out, Epi ep) {
  constexpr int TH = kThreads / TW;
  const int c = blockIdx.x * TW + (threadIdx.x % TW);
  // with 64-wide rows a wave is one image row: say so, for scalar row loads
  const int ty = TW == 64 ? __builtin_amdgcn_readfirstlane(threadIdx.x / TW)
                          : static_cast<int>(threadIdx.x / TW);
  const int r = blockIdx.y * TH + ty;
  if (c >= g.W || r >= g.H) return;
  pixels_body<BT, WM>(g, views, edges, lines, pint, sino, out, r, c, blockIdx.z * BT, ep);
}

// View-resolved sinogram -> image: out[b][v] is view v's contribution
// alone (the per-view loop of deeprecon's bprj_sv_fun as one launch).
template <int BT, int WM>
__global__ __launch_bounds__(kThreads) void k_pixels_views(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out) {
  const int ntx = (g.W + kTile - 1) / kTile;
  const int c = (blockIdx.x % ntx) * kTile + (threadIdx.x % kTile);
  const int r = (blockIdx.x / ntx) * kTile + (threadIdx.x / kTile);
  const int v = blockIdx.y;
  const int b0 = blockIdx.z * BT;
  if (c >= g.W || r >= g.H) return;
  float xcw = 0.0f, ycw = 0.0f;
  if constexpr (WM == W_BPW) pixel_centre(g, 1, r, c, &xcw, &ycw);
  float acc[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) acc[b] = 0.0f;
  pixel_one_view<BT, WM>(g, views[v], v, edges, lines, pint, sino, r, c, b0, xcw, ycw, acc);
  const size_t plane = static_cast<size_t>(g.H) * g.W;
#pragma unroll
  for (int b = 0; b < BT; ++b)
    if (b0 + b < g.B)
      out[(static_cast<size_t>(b0 + b) * g.V + v) * plane + static_cast<size_t>(r) * g.W + c] =
          acc[b];
}

// Variant DD_ADJ_QUAD: one thread per 2 x 2 pixel block, a 32 x 32 pixel
// tile per workgroup.
template <int BT, int WM>
__global__ __launch_bounds__(kThreads) void k_pixels_quad(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out, Epi ep) {
  const int c0 = blockIdx.x * (2 * kTile) + 2 * (threadIdx.x % kTile);
  const int r0 = blockIdx.y * (2 * kTile) + 2 * (threadIdx.x / kTile);
  if (c0 >= g.W || r0 >= g.H) return;
  quads_body<BT, WM>(g, views, edges, lines, pint, sino, out, r0, c0, blockIdx.z * BT, ep);
}

// Variants DD_ADJ_TILE16 / DD_ADJ_TILE8: a T x T tile stages its lines' edge
// positions per view in LDS.
template <int BT, int WM, int T>
__global__ __launch_bounds__(T * T) void k_pixels_tile(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float* __restrict__ sino, float* __restrict__ out, Epi ep) {
  __shared__ PixelTileLds<T> lds;
  pixels_tile_body<BT, WM, T>(g, views, edges, lines, pint, sino, out, blockIdx.x,
                              blockIdx.y, blockIdx.z * BT, lds, ep);
}

// ---------------------------------------------------------------------------
// Variant DD_ADJ_BLOCK_PREFIX at instance tile 1: sinogram -> image through
// running sums over the bins, with the views split over workgroups
// ---------------------------------------------------------------------------
// A single slice gives 64 x 64-pixel tiles for only 64 workgroups at 512^2,
// so each tile's views are split into chunks: workgroup (tile, chunk,
// instance) adds its chunk's views into a partial image, and k_sum_chunks
// adds the partial images in chunk order.  Per view the same evaluation as
// the batch path's running-sum adjoint (block_view_prefix over FP64 knots
// staged in LDS).
constexpr int kPrefQ = 4;                          // pixels per thread side
constexpr int kPrefSide = 16;                      // threads per tile side
constexpr int kPrefThreads = kPrefSide * kPrefSide;
constexpr int kPrefTile = kPrefSide * kPrefQ;      // 64 x 64 pixels
constexpr int kPrefKnots = kPrefThreads;           // knots staged per view

__host__ __device__ inline int prefix_tiles(const Geo2& g) {
  return ((g.W + kPrefTile - 1) / kPrefTile) * ((g.H + kPrefTile - 1) / kPrefTile);
}

// View chunks: about 1024 workgroups in all, at least 16 views per chunk.
int prefix_chunks(const Geo2& g) {
  const long long items = 1LL * prefix_tiles(g) * g.B;
  long long c = (1024 + items - 1) / items;
  c = std::min<long long>(c, std::max(1, g.V / 16));
  return static_cast<int>(std::max<long long>(1, std::min<long long>(c, 64)));
}

// Knots of view v of instance b: one workgroup per (view, instance).
template <bool FOLD_W>
__global__ __launch_bounds__(kThreads) void k_prefix_sino(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const float* __restrict__ wfp, const float* __restrict__ sino,
    double2* __restrict__ psino) {
  __shared__ double lds[kThreads];
  const int v = blockIdx.x, b = blockIdx.y;
  prefix_sino_view<FOLD_W>(g, views, edges, wfp, sino + static_cast<size_t>(b) * g.V * g.D, v,
                           psino + (static_cast<size_t>(b) * g.V + v) * (g.D + 1), lds);
}

// Knot range of every (view, 64 x 64 tile): kr[v * tiles + t].
__global__ __launch_bounds__(kThreads) void k_prefix_ranges(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, int2* __restrict__ kr) {
  const int tx = (g.W + kPrefTile - 1) / kPrefTile;
  const int tiles = prefix_tiles(g);
  const long long idx = 1LL * blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 1LL * tiles * g.V) return;
  const int v = static_cast<int>(idx / tiles), t = static_cast<int>(idx % tiles);
  const int c0 = (t % tx) * kPrefTile, r0 = (t / tx) * kPrefTile;
  kr[idx] = knot_range(g, views[v], edges + static_cast<size_t>(v) * (g.D + 1),
                       lines + static_cast<size_t>(v) * g.Lmax, r0, min(r0 + kPrefTile, g.H),
                       c0, min(c0 + kPrefTile, g.W));
}

// Workgroup (tile, chunk, instance): views [chunk * vchunk, +vchunk).  Each
// view's knots for the tile are staged in LDS (double-buffered: the next
// view's copy is loaded while this one is walked, one barrier per view); a
// view whose range exceeds the buffer is walked from global memory.
template <int WM>
__global__ __launch_bounds__(kPrefThreads)
__attribute__((amdgpu_waves_per_eu(5))) void k_pixels_prefix(
    Geo2 g, const ViewGeo* __restrict__ views, const LineGeo* __restrict__ lines,
    const float* __restrict__ pint, const double2* __restrict__ psino,
    const int2* __restrict__ kr, float* __restrict__ part, int vchunk) {
  __shared__ double2 knots[2][kPrefKnots];
  const int tid = static_cast<int>(threadIdx.x);
  const int tx = (g.W + kPrefTile - 1) / kPrefTile;
  const int tiles = prefix_tiles(g);
  const int t = blockIdx.x, b = blockIdx.z;
  const int v0 = blockIdx.y * vchunk, v1 = min(v0 + vchunk, g.V);
  const int c0 = (t % tx) * kPrefTile + kPrefQ * (tid % kPrefSide);
  const int r0 = (t / tx) * kPrefTile + kPrefQ * (tid / kPrefSide);
  const bool active = c0 < g.W && r0 < g.H;
  const int nr = min(kPrefQ, g.H - r0), nc = min(kPrefQ, g.W - c0);
  const double2* __restrict__ tab = psino + static_cast<size_t>(b) * g.V * (g.D + 1);
  const int2* __restrict__ rng = kr + t;  // view v at rng[v * tiles]
  float acc[kPrefQ][kPrefQ];
#pragma unroll
  for (int dr = 0; dr < kPrefQ; ++dr)
#pragma unroll
    for (int dc = 0; dc < kPrefQ; ++dc) acc[dr][dc] = 0.0f;
  if (v0 < v1) {
    int2 r = rng[static_cast<size_t>(v0) * tiles];
    bool staged = r.y - r.x + 1 <= kPrefKnots;
    if (staged && tid <= r.y - r.x)
      knots[v0 & 1][tid] = tab[static_cast<size_t>(v0) * (g.D + 1) + r.x + tid];
    __syncthreads();
    for (int v = v0; v < v1; ++v) {
      // the next view's knot for this thread, in flight during this view's walks
      const int vn = min(v + 1, v1 - 1);
      const int2 rn = rng[static_cast<size_t>(vn) * tiles];
      const bool staged_n = rn.y - rn.x + 1 <= kPrefKnots;
      double2 kn = make_double2(0.0, 0.0);
      const bool mine = v + 1 < v1 && staged_n && tid <= rn.y - rn.x;
      if (mine) kn = tab[static_cast<size_t>(vn) * (g.D + 1) + rn.x + tid];
      if (active) {
        const ViewGeo vg = views[v];
        const LineGeo* __restrict__ lv = lines + static_cast<size_t>(v) * g.Lmax;
        const float* __restrict__ pv = pint + static_cast<size_t>(v) * g.Lmax;
        if (staged)
          block_view_prefix<WM, kPrefQ>(
              g, vg, lv, pv, (const __attribute__((address_space(3))) double*)(knots[v & 1]),
              r.x, r.y, r0, c0, nr, nc, acc);
        else
          block_view_prefix<WM, kPrefQ>(
              g, vg, lv, pv,
              reinterpret_cast<const double*>(tab + static_cast<size_t>(v) * (g.D + 1)), 0,
              g.D, r0, c0, nr, nc, acc);
      }
      if (mine) knots[(v + 1) & 1][tid] = kn;
      __syncthreads();
      r = rn;
      staged = staged_n;
    }
  }
  if (active) {
    float* __restrict__ o =
        part + (static_cast<size_t>(blockIdx.y) * g.B + b) * g.H * g.W;
#pragma unroll
    for (int dr = 0; dr < kPrefQ; ++dr)
#pragma unroll
      for (int dc = 0; dc < kPrefQ; ++dc)
        if (dr < nr && dc < nc) o[static_cast<size_t>(r0 + dr) * g.W + c0 + dc] = acc[dr][dc];
  }
}

// Running-sum forward at instance tile 1 (DD_FWD_PREFIX): per (line,
// instance) the line's running pixel sums as FP32 (hi, lo) pairs of an FP64
// scan, {hi, lo, x, 0} per pixel edge; the H rows first, then the W columns
// read bottom-up.
__global__ __launch_bounds__(kThreads) void k_prefix_lines(
    Geo2 g, const float* __restrict__ img, float4* __restrict__ prow,
    float4* __restrict__ pcol) {
  __shared__ double lds[kThreads];
  const int line = blockIdx.x, b = blockIdx.y;
  const int H = g.H, W = g.W;
  const float* __restrict__ x = img + static_cast<size_t>(b) * H * W;
  if (line < H)
    prefix_line(x + static_cast<size_t>(line) * W, 1, W,
                prow + (static_cast<size_t>(b) * H + line) * prefix_entries(W), lds);
  else
    prefix_line(x + static_cast<size_t>(H - 1) * W + (line - H), -static_cast<long long>(W), H,
                pcol + (static_cast<size_t>(b) * W + line - H) * prefix_entries(H), lds);
}

// One wave per (line chunk, block of kBinsPerWave bins, view, instance),
// chunk fastest: each bin's integral over each line of the chunk as a
// difference of the line's running sums (bins_accumulate_prefix), into the
// chunk's partial sinogram, which k_bins_reduce adds in chunk order.
template <int WM>
__global__ __launch_bounds__(64) void k_bins_prefix(
    Geo2 g, const ViewGeo* __restrict__ views, const double* __restrict__ edges,
    const LineGeo* __restrict__ lines, const float* __restrict__ pint,
    const float4* __restrict__ prow, const float4* __restrict__ pcol,
    float* __restrict__ part) {
  const int kb_n = (g.D + kBinsPerWave - 1) / kBinsPerWave;
  const int chunk = blockIdx.x % kSplit;
  int rest = blockIdx.x / kSplit;
  const int lane = static_cast<int>(threadIdx.x);
  const int k = (rest % kb_n) * kBinsPerWave + lane;
  rest /= kb_n;
  const int v = rest % g.V, b = rest / g.V;
  // the whole wave works (the upper edge comes from the next lane); lanes
  // past the last bin use clamped edges and write nothing
  const ViewGeo vg = views[v];
  const int L = vg.xaxis ? g.H : g.W;
  const int N = vg.xaxis ? g.W : g.H;
  const double e1 = edges[static_cast<size_t>(v) * (g.D + 1) + min(k, g.D)];
  const float4* __restrict__ tab =
      (vg.xaxis ? prow : pcol) + static_cast<size_t>(b) * L * prefix_entries(N);
  const float acc = bins_accumulate_prefix<WM>(
      e1, lines + static_cast<size_t>(v) * g.Lmax, pint + static_cast<size_t>(v) * g.Lmax, tab,
      bp_const(g), L * chunk / kSplit, L * (chunk + 1) / kSplit, N);
  if (lane == 63 || k >= g.D) return;
  part[((static_cast<size_t>(chunk) * g.Bp + b) * g.V + v) * g.D + k] = acc;
}

// out = epilogue(sum of the chunks' partial images, in chunk order).
__global__ __launch_bounds__(kThreads) void k_sum_chunks(
    Geo2 g, const float* __restrict__ part, int nchunk, float* __restrict__ out, Epi ep) {
  const size_t n = static_cast<size_t>(g.B) * g.H * g.W;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = part[i];
  for (int c = 1; c < nchunk; ++c) s += part[static_cast<size_t>(c) * n + i];
  out[i] = epi_value(ep, s, i);
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
  size_t buf0, buf1, part, data_total;                // data buffers
  size_t psino, krange, ppart;                        // running-sum adjoint
  size_t prow, pcol;                                  // running-sum forward
};

constexpr size_t kAlign = 256;

int pick_tile(int batch) {
  const int forced = g_forced_tile.load();
  if (forced > 0) return forced;
  int t = 1;
  while (t < batch && t < 16) t *= 2;
  return t;
}

bool to_sino(int op) {
  return op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW;
}

// prefix (PrefixRegions): also the running-sum adjoint's knots [B][V][D+1]
// (16 bytes each), knot ranges [V][tiles] and partial images
// [chunks][B][H][W]; or the running-sum forward's line tables, rows
// [B][H][W+1] and columns [B][W][H+1] of 16-byte entries.
Layout make_layout(const Geo2& g, int op, int prefix = kPrefNone) {
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
    L.part = take(sizeof(float) * static_cast<size_t>(kSplit) * g.Bp * g.V * g.D);
    if (prefix & kPrefForward) {
      L.prow = take(16 * static_cast<size_t>(g.B) * g.H * prefix_entries(g.W));
      L.pcol = take(16 * static_cast<size_t>(g.B) * g.W * prefix_entries(g.H));
    }
  } else {
    L.buf0 = take(sizeof(float) * static_cast<size_t>(g.V) * g.D * g.Bp);
    L.buf1 = L.buf0;
    if (prefix & kPrefAdjoint) {
      L.psino = take(16 * static_cast<size_t>(g.B) * g.V * (g.D + 1));
      L.krange = take(8 * static_cast<size_t>(g.V) * prefix_tiles(g));
      L.ppart = take(sizeof(float) * static_cast<size_t>(prefix_chunks(g)) * g.B * g.H * g.W);
    }
  }
  L.data_total = off;
  return L;
}

// Planned workspaces hold the running-sum buffers whenever they can be used
// (instance tile 1), so any variant may be chosen after planning.
int planned_prefix(const Geo2& g, int op) {
  if (g.BT != 1) return kPrefNone;
  if (op == DD_PROJ_T || op == DD_BPROJ) return kPrefAdjoint;
  if (op == DD_PROJ || op == DD_BPROJ_T) return kPrefForward;
  return kPrefNone;
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

// Per-pixel adjoint of one workgroup width (256 threads: TW x 256 / TW pixels).
template <int BT, int TW>
void launch_pixels(const Geo2& g, int op, const GeoPtrs& G, const float* buf0, float* output,
                   hipStream_t st, const Epi& ep) {
  constexpr int TH = kThreads / TW;
  dim3 grid((g.W + TW - 1) / TW, (g.H + TH - 1) / TH, g.Bp / BT);
  if (op == DD_PROJ_T)
    hipLaunchKernelGGL((k_pixels<BT, W_FP, TW>), grid, dim3(kThreads), 0, st, g, G.views,
                       G.edges, G.lines, G.pint, buf0, output, ep);
  else if (op == DD_BPROJ)
    hipLaunchKernelGGL((k_pixels<BT, W_BP, TW>), grid, dim3(kThreads), 0, st, g, G.views,
                       G.edges, G.lines, G.pint, buf0, output, ep);
  else
    hipLaunchKernelGGL((k_pixels<BT, W_BPW, TW>), grid, dim3(kThreads), 0, st, g, G.views,
                       G.edges, G.lines, G.pint, buf0, output, ep);
}

template <int BT>
dd_status launch_tile(const Geo2& g, int op, int variant, const float* input,
                      float* output, const GeoPtrs& G, char* ws, const Layout& L,
                      hipStream_t st, const Epi& ep, const Shape& sh) {
  auto* buf0 = reinterpret_cast<float*>(ws + L.buf0);
  auto* buf1 = reinterpret_cast<float*>(ws + L.buf1);
  const int ntile = g.Bp / BT;
  const int adj = variant & ~DD_FWD_PREFIX;
  if (BT == 1 && fwd_prefix_path(g, op, variant)) {
    auto* prow = reinterpret_cast<float4*>(ws + L.prow);
    auto* pcol = reinterpret_cast<float4*>(ws + L.pcol);
    auto* part = reinterpret_cast<float*>(ws + L.part);
    const int kb_n = (g.D + kBinsPerWave - 1) / kBinsPerWave;
    const long long items = 1LL * kSplit * kb_n * g.V * g.B;
    dim3 rgrid(((g.D + kThreads - 1) / kThreads) * g.V * g.B);
    ProfileScope prof(DD_KERNEL_IMAGE_TO_SINO, st);
    hipLaunchKernelGGL(k_prefix_lines, dim3(g.H + g.W, g.B), dim3(kThreads), 0, st, g, input,
                       prow, pcol);
    if (op == DD_PROJ) {
      hipLaunchKernelGGL((k_bins_prefix<W_FP>), dim3(static_cast<unsigned>(items)), dim3(64), 0,
                         st, g, G.views, G.edges, G.lines, G.pint, prow, pcol, part);
      hipLaunchKernelGGL((k_bins_reduce<W_FP>), rgrid, dim3(kThreads), 0, st, g, G.views,
                         G.wfp, part, output, ep);
    } else {
      hipLaunchKernelGGL((k_bins_prefix<W_BP>), dim3(static_cast<unsigned>(items)), dim3(64), 0,
                         st, g, G.views, G.edges, G.lines, G.pint, prow, pcol, part);
      hipLaunchKernelGGL((k_bins_reduce<W_BP>), rgrid, dim3(kThreads), 0, st, g, G.views,
                         G.wfp, part, output, ep);
    }
  } else if (to_sino(op)) {
    hipLaunchKernelGGL(k_interleave_image<BT>,
                       dim3((g.W + kItile - 1) / kItile, (g.H + kItile - 1) / kItile, ntile),
                       dim3(kItile * kItile), 0, st, g, input, buf0, buf1);
    auto* part = reinterpret_cast<float*>(ws + L.part);
    const int kb_n = (g.D + kThreads - 1) / kThreads;
    const int waves = sh.fwd_waves, vpb = sh.fwd_views;
    const int kbs_n = (g.D + waves * kBinsPerWave - 1) / (waves * kBinsPerWave);
    dim3 pgrid(kSplit * kbs_n * ((g.V + vpb - 1) / vpb) * ntile);
    dim3 pblock(64 * waves);
    dim3 rgrid(kb_n * g.V * g.B);
    ProfileScope prof(DD_KERNEL_IMAGE_TO_SINO, st);
    if (op == DD_PROJ) {
      hipLaunchKernelGGL((k_bins_split<BT, W_FP>), pgrid, pblock, 0, st, g, G.views,
                         G.edges, G.lines, G.pint, buf0, buf1, part, vpb);
      hipLaunchKernelGGL((k_bins_reduce<W_FP>), rgrid, dim3(kThreads), 0, st, g, G.views,
                         G.wfp, part, output, ep);
    } else if (op == DD_BPROJ_T) {
      hipLaunchKernelGGL((k_bins_split<BT, W_BP>), pgrid, pblock, 0, st, g, G.views,
                         G.edges, G.lines, G.pint, buf0, buf1, part, vpb);
      hipLaunchKernelGGL((k_bins_reduce<W_BP>), rgrid, dim3(kThreads), 0, st, g, G.views,
                         G.wfp, part, output, ep);
    } else {
      hipLaunchKernelGGL((k_bins_split<BT, W_BPW>), pgrid, pblock, 0, st, g, G.views,
                         G.edges, G.lines, G.pint, buf0, buf1, part, vpb);
      hipLaunchKernelGGL((k_bins_reduce<W_BPW>), rgrid, dim3(kThreads), 0, st, g, G.views,
                         G.wfp, part, output, ep);
    }
  } else if (BT == 1 && prefix_path(g, op, variant)) {
    auto* psino = reinterpret_cast<double2*>(ws + L.psino);
    auto* kr = reinterpret_cast<int2*>(ws + L.krange);
    auto* pp = reinterpret_cast<float*>(ws + L.ppart);
    const int tiles = prefix_tiles(g);
    const int vchunk = (g.V + prefix_chunks(g) - 1) / prefix_chunks(g);
    const int nchunk = (g.V + vchunk - 1) / vchunk;
    const long long nr = 1LL * tiles * g.V;
    const size_t npix = static_cast<size_t>(g.B) * g.H * g.W;
    ProfileScope prof(DD_KERNEL_SINO_TO_IMAGE, st);
    if (op == DD_PROJ_T)
      hipLaunchKernelGGL((k_prefix_sino<true>), dim3(g.V, g.B), dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.wfp, input, psino);
    else
      hipLaunchKernelGGL((k_prefix_sino<false>), dim3(g.V, g.B), dim3(kThreads), 0, st, g,
                         G.views, G.edges, G.wfp, input, psino);
    hipLaunchKernelGGL(k_prefix_ranges, dim3(static_cast<unsigned>((nr + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, st, g, G.views, G.edges, G.lines, kr);
    const dim3 grid(tiles, nchunk, g.B);
    if (op == DD_PROJ_T)
      hipLaunchKernelGGL((k_pixels_prefix<W_FP>), grid, dim3(kPrefThreads), 0, st, g, G.views,
                         G.lines, G.pint, psino, kr, pp, vchunk);
    else
      hipLaunchKernelGGL((k_pixels_prefix<W_BP>), grid, dim3(kPrefThreads), 0, st, g, G.views,
                         G.lines, G.pint, psino, kr, pp, vchunk);
    hipLaunchKernelGGL(k_sum_chunks, dim3(static_cast<unsigned>((npix + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, st, g, pp, nchunk, output, ep);
  } else {
    dim3 sgrid((g.D + kThreads - 1) / kThreads, g.V, ntile);
    if (op == DD_PROJ_T)
      hipLaunchKernelGGL((k_interleave_sino<BT, true>), sgrid, dim3(kThreads),
                         0, st, g, G.views, G.wfp, input, buf0);
    else
      hipLaunchKernelGGL((k_interleave_sino<BT, false>), sgrid, dim3(kThreads),
                         0, st, g, G.views, G.wfp, input, buf0);
    ProfileScope prof(DD_KERNEL_SINO_TO_IMAGE, st);
    if (adj == DD_ADJ_QUAD) {
      dim3 grid((g.W + 2 * kTile - 1) / (2 * kTile), (g.H + 2 * kTile - 1) / (2 * kTile), ntile);
      if (op == DD_PROJ_T)
        hipLaunchKernelGGL((k_pixels_quad<BT, W_FP>), grid, dim3(kThreads), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
      else if (op == DD_BPROJ)
        hipLaunchKernelGGL((k_pixels_quad<BT, W_BP>), grid, dim3(kThreads), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
      else
        hipLaunchKernelGGL((k_pixels_quad<BT, W_BPW>), grid, dim3(kThreads), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
    } else if (adj == DD_ADJ_TILE16) {
      dim3 grid((g.W + 15) / 16, (g.H + 15) / 16, ntile);
      if (op == DD_PROJ_T)
        hipLaunchKernelGGL((k_pixels_tile<BT, W_FP, 16>), grid, dim3(256), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
      else if (op == DD_BPROJ)
        hipLaunchKernelGGL((k_pixels_tile<BT, W_BP, 16>), grid, dim3(256), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
      else
        hipLaunchKernelGGL((k_pixels_tile<BT, W_BPW, 16>), grid, dim3(256), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
    } else if (adj == DD_ADJ_TILE8) {
      dim3 grid((g.W + 7) / 8, (g.H + 7) / 8, ntile);
      if (op == DD_PROJ_T)
        hipLaunchKernelGGL((k_pixels_tile<BT, W_FP, 8>), grid, dim3(64), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
      else if (op == DD_BPROJ)
        hipLaunchKernelGGL((k_pixels_tile<BT, W_BP, 8>), grid, dim3(64), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
      else
        hipLaunchKernelGGL((k_pixels_tile<BT, W_BPW, 8>), grid, dim3(64), 0, st, g,
                           G.views, G.edges, G.lines, G.pint, buf0, output, ep);
    } else {
      const int tw = sh.pix_width > 0 ? sh.pix_width : pix_tile_w<BT>();
      if (tw == 64)
        launch_pixels<BT, 64>(g, op, G, buf0, output, st, ep);
      else if (tw == 32)
        launch_pixels<BT, 32>(g, op, G, buf0, output, st, ep);
      else
        launch_pixels<BT, 16>(g, op, G, buf0, output, st, ep);
    }
  }
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

dd_status run_tiles(const Geo2& q, int op, int variant, const float* input,
                    float* output, const GeoPtrs& G, char* w, const Layout& L,
                    hipStream_t st, const Epi& ep = Epi{}, const Shape& sh = Shape{}) {
  switch (q.BT) {
    case 1: return launch_tile<1>(q, op, variant, input, output, G, w, L, st, ep, sh);
    case 2: return launch_tile<2>(q, op, variant, input, output, G, w, L, st, ep, sh);
    case 4: return launch_tile<4>(q, op, variant, input, output, G, w, L, st, ep, sh);
    case 8: return launch_tile<8>(q, op, variant, input, output, G, w, L, st, ep, sh);
    default: return launch_tile<16>(q, op, variant, input, output, G, w, L, st, ep, sh);
  }
}

template <int BT>
dd_status launch_views(const Geo2& g, int op, const float* input, float* output,
                       const GeoPtrs& G, char* ws, const Layout& L, hipStream_t st) {
  auto* buf0 = reinterpret_cast<float*>(ws + L.buf0);
  const int ntile = g.Bp / BT;
  dim3 sgrid((g.D + kThreads - 1) / kThreads, g.V, ntile);
  if (op == DD_PROJ_T)
    hipLaunchKernelGGL((k_interleave_sino<BT, true>), sgrid, dim3(kThreads), 0, st, g,
                       G.views, G.wfp, input, buf0);
  else
    hipLaunchKernelGGL((k_interleave_sino<BT, false>), sgrid, dim3(kThreads), 0, st, g,
                       G.views, G.wfp, input, buf0);
  dim3 grid(((g.W + kTile - 1) / kTile) * ((g.H + kTile - 1) / kTile), g.V, ntile);
  ProfileScope prof(DD_KERNEL_SINO_TO_IMAGE, st);
  if (op == DD_PROJ_T)
    hipLaunchKernelGGL((k_pixels_views<BT, W_FP>), grid, dim3(kThreads), 0, st, g, G.views,
                       G.edges, G.lines, G.pint, buf0, output);
  else if (op == DD_BPROJ)
    hipLaunchKernelGGL((k_pixels_views<BT, W_BP>), grid, dim3(kThreads), 0, st, g, G.views,
                       G.edges, G.lines, G.pint, buf0, output);
  else
    hipLaunchKernelGGL((k_pixels_views<BT, W_BPW>), grid, dim3(kThreads), 0, st, g, G.views,
                       G.edges, G.lines, G.pint, buf0, output);
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

// Fixed scratch of dd2d_run_noalloc: one copy per device, resident until the
// process ends.  Host code never uses the symbol's name as an address (it is a
// host-side stand-in); it uses the per-device address that
// hipGetSymbolAddress returns, taken once per device with that device current.
__device__ __align__(256) char scratch[DD_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
std::atomic<void*> g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (nullptr on failure).
void* scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (p) return p;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
  if (!p) {
    void* q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return nullptr;
    g_scratch_ptr[dev].store(q, std::memory_order_release);
    p = q;
  }
  return p;
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

void dd2d_set_adjoint_variant(int variant) {
  g_adjoint_variant = known_variant(variant) ? variant : -1;
}

void dd2d_set_instance_tile(int tile) {
  g_forced_tile = (tile == 1 || tile == 2 || tile == 4 || tile == 8 || tile == 16) ? tile : 0;
}

size_t dd2d_workspace_size(int scan, int op, const dd_geom2d* g, int batch,
                           int views) {
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, pick_tile(batch))) return 0;
  const Layout L = make_layout(q, op, prefix_regions(q, op, adjoint_variant(-1, q)));
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
  const int var = adjoint_variant(-1, q);
  const Layout L = make_layout(q, op, prefix_regions(q, op, var));
  if (!ws || ws_bytes < L.geo_total + L.data_total) return DD_NO_WORKSPACE;
  auto st = static_cast<hipStream_t>(stream);
  char* w = static_cast<char*>(ws);
  launch_setup(q, angles, w, L, st);
  return run_tiles(q, op, var, input, output, geo_ptrs(w, L), w + L.geo_total, L, st);
}

dd_status dd2d_run_noalloc(int scan, int op, const dd_geom2d* g, const float* input,
                           float* output, const float* angles, int batch, int views,
                           void* stream) {
  dd_status rc = check_op(scan, op);
  if (rc != DD_OK) return rc;
  if (!input || !output || !angles) return DD_BAD_ARG;
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, pick_tile(batch))) return DD_BAD_ARG;
  const int var = adjoint_variant(-1, q);
  const Layout L = make_layout(q, op, prefix_regions(q, op, var));
  if (L.geo_total + L.data_total > DD_NOALLOC_SCRATCH_BYTES) return DD_TOO_LARGE;
  char* w = static_cast<char*>(scratch_ptr());
  if (!w) return DD_LAUNCH_FAIL;
  // every table and buffer in the scratch is fully written by this call's
  // own launches before it is read, so nothing needs clearing between calls
  auto st = static_cast<hipStream_t>(stream);
  launch_setup(q, angles, w, L, st);
  return run_tiles(q, op, var, input, output, geo_ptrs(w, L), w + L.geo_total, L, st);
}

dd_status dd2d_run_views(int scan, int op, const dd_geom2d* g, const float* input,
                         float* output, const float* angles, int batch, int views,
                         void* ws, size_t ws_bytes, void* stream) {
  dd_status rc = check_op(scan, op);
  if (rc != DD_OK) return rc;
  if (op != DD_PROJ_T && op != DD_BPROJ && op != DD_BPROJ_W) return DD_UNSUPPORTED;
  if (!input || !output || !angles) return DD_BAD_ARG;
  if (views > 65535) return DD_UNSUPPORTED;  // one grid row per view
  Geo2 q;
  const int tile = std::min(pick_tile(batch), 4);
  if (!fill_geo(&q, scan, g, batch, views, tile)) return DD_BAD_ARG;
  const Layout L = make_layout(q, op);
  if (!ws || ws_bytes < L.geo_total + L.data_total) return DD_NO_WORKSPACE;
  auto st = static_cast<hipStream_t>(stream);
  char* w = static_cast<char*>(ws);
  launch_setup(q, angles, w, L, st);
  const GeoPtrs G = geo_ptrs(w, L);
  switch (q.BT) {
    case 1: return launch_views<1>(q, op, input, output, G, w + L.geo_total, L, st);
    case 2: return launch_views<2>(q, op, input, output, G, w + L.geo_total, L, st);
    default: return launch_views<4>(q, op, input, output, G, w + L.geo_total, L, st);
  }
}

size_t dd2d_views_workspace_size(int scan, int op, const dd_geom2d* g, int batch,
                                 int views) {
  Geo2 q;
  if (check_op(scan, op) != DD_OK) return 0;
  if (!fill_geo(&q, scan, g, batch, views, std::min(pick_tile(batch), 4))) return 0;
  const Layout L = make_layout(q, op);
  return L.geo_total + L.data_total;
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

static int valid_tile(int tile, int batch) {
  return (tile == 1 || tile == 2 || tile == 4 || tile == 8 || tile == 16) ? tile
                                                                          : pick_tile(batch);
}

size_t dd2d_planned_workspace_size(int scan, int op, const dd_geom2d* g,
                                   int batch, int views, int tile) {
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, valid_tile(tile, batch))) return 0;
  return make_layout(q, op, planned_prefix(q, op)).data_total;
}

dd_status dd2d_run_planned(int scan, int op, const dd_geom2d* g,
                           const void* geo, const float* input, float* output,
                           int batch, int views, int tile, int variant, void* ws,
                           size_t ws_bytes, void* stream) {
  dd_status rc = check_op(scan, op);
  if (rc != DD_OK) return rc;
  if (!geo || !input || !output) return DD_BAD_ARG;
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, valid_tile(tile, batch))) return DD_BAD_ARG;
  const Layout L = make_layout(q, op, planned_prefix(q, op));
  if (!ws || ws_bytes < L.data_total) return DD_NO_WORKSPACE;
  return run_tiles(q, op, adjoint_variant(variant, q), input, output, geo_ptrs(geo, L),
                   static_cast<char*>(ws), L, static_cast<hipStream_t>(stream));
}

dd_status dd2d_run_planned_ep(int scan, int op, const dd_geom2d* g,
                              const void* geo, const float* input, float* output,
                              int batch, int views, int tile, int variant,
                              const dd_shape* shape, const dd_epilogue* ep, void* ws,
                              size_t ws_bytes, void* stream) {
  dd_status rc = check_op(scan, op);
  if (rc != DD_OK) return rc;
  if (!geo || !input || !output) return DD_BAD_ARG;
  if (ep && ep->c == output) return DD_BAD_ARG;  // c is read after other threads write
  Shape sh;
  if (!make_shape(shape, &sh)) return DD_BAD_ARG;
  Geo2 q;
  if (!fill_geo(&q, scan, g, batch, views, valid_tile(tile, batch))) return DD_BAD_ARG;
  const Layout L = make_layout(q, op, planned_prefix(q, op));
  if (!ws || ws_bytes < L.data_total) return DD_NO_WORKSPACE;
  Epi e;
  if (ep) {
    e.c = ep->c;
    e.scale = ep->scale;
    e.alpha = ep->alpha;
    e.beta = ep->beta;
  }
  return run_tiles(q, op, adjoint_variant(variant, q), input, output, geo_ptrs(geo, L),
                   static_cast<char*>(ws), L, static_cast<hipStream_t>(stream), e, sh);
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
