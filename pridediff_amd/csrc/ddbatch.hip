// Batched 2D projection of many instances of uneven size in one launch per
// phase.
//
// A batch is a list of instances, each with its own geometry (image size,
// detector, view count and angles) and its own slot in one packed image
// buffer and one packed sinogram buffer.  setup() builds every instance's
// FP64 geometry tables once and uploads a device descriptor table; run() is
// then two or three launches per op for the whole batch — a layout phase, a
// gather phase and, for image -> sinogram ops, the chunk reduction — each a
// grid over a static list of work items (instance, view, bin block, line
// chunk), (instance, view, bin block) or (instance, pixel tile), heaviest
// instances first: by default one workgroup per item, or a set number of
// workgroups that stride over the list.  No counters, no atomics, no host work: a run is capturable and
// replays bit-identically.
//
// The per-item arithmetic is the one of the one-call kernels
// (dd2d_body.h), so a batched result equals the per-instance call exactly.
// No counterpart in the reference (one launch per instance and op,
// deeprecon/cuda/dis2d/FanArcDisCUDA2d.cu:758-1016).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "dd.h"
#include "dd2d_body.h"
#include "dd_common.h"
#include "dd_profile.h"

namespace dd {
namespace {

using namespace body2d;

// Device descriptor of one instance, as stored in the descriptor table at the
// start of the batch's device block: its tables and scratch are byte offsets
// into that block.  Kernels take the block's base as an argument and rebuild
// the pointers from it (BatchInst), so every access is a global-memory one
// (pointers read from memory would otherwise be generic: flat loads).
enum Region {
  kViews, kEdges, kWfp, kLines, kPint, kXrow, kXcol, kSino, kPart, kPrefRow, kPrefCol,
  kPrefSino, kKnotRange, kRegions
};

enum Phase {
  kLayoutImage = 0,
  kGatherBins = 1,
  kLayoutSino = 2,
  kGatherPixels = 3,
  kReduceBins = 4,
  kGatherQuads = 5,   // sinogram -> image with one thread per 4 x 4 pixels
  kLayoutPrefix = 6,  // running sums of every image row and column
  kPhases = 7
};

struct BatchDesc {
  Geo2 g;
  unsigned long long off[kRegions];
  long long img_off;   // floats, into the packed image buffer
  long long sino_off;  // floats, into the packed sinogram buffer
  int item0[kPhases];  // first work item of each phase
  int nitems[kPhases];
};

// One instance inside a kernel: the descriptor with typed pointers.
struct BatchInst {
  Geo2 g;
  ViewGeo* views;
  double* edges;
  float* wfp;
  LineGeo* lines;
  float* pint;
  float* xrow;         // layout scratch: image rows, interleaved
  float* xcol;         // layout scratch: image columns, bottom-up
  float* sino;         // layout scratch: sinogram in walk order
  float* part;         // partial bin sums per line chunk [kSplit][B][V][D]
  float4* prow;        // running sums of the rows [B][H][W + 1]
  float4* pcol;        // running sums of the columns, bottom-up [B][W][H + 1]
  double2* psino;      // running-sum knots over the bins per view [B][V][D + 1]
  int2* krange;        // knots each 64 x 64 pixel tile needs, per view [V][tiles]
  long long img_off;
  long long sino_off;
};

constexpr int kBT = 1;  // one geometry per instance

// Block items of the DD_ADJ_QUAD variant: one thread per kBlockQ x kBlockQ
// pixel block (one first-bin search and one bin walk per line serve kBlockQ
// pixels), kQuadSide x kQuadSide threads per item.  Measured on 64 uneven
// instances: adjoint 17.4 / 16.4 / 14.7 / 17.6 / 21.0 ms for blocks of 2 / 3
// / 4 / 5 / 6 (register pressure past 4); items of 8 x 8 threads (32 x 32
// pixels) instead of 16 x 16: 13.9 vs 13.6 ms.
constexpr int kBlockQ = 4;
constexpr int kQuadSide = 16;
constexpr int kQuadThreads = kQuadSide * kQuadSide;
constexpr int kQuadTile = kQuadSide * kBlockQ;  // pixels per quad item edge

__device__ __forceinline__ int find_inst(const char* __restrict__ mem, int n, int phase,
                                         int item) {
  const BatchDesc* __restrict__ desc = reinterpret_cast<const BatchDesc*>(mem);
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (desc[mid].item0[phase] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ BatchInst load_inst(char* __restrict__ mem, int i) {
  const BatchDesc& s = reinterpret_cast<const BatchDesc*>(mem)[i];
  BatchInst d;
  d.g = s.g;
  d.views = reinterpret_cast<ViewGeo*>(mem + s.off[kViews]);
  d.edges = reinterpret_cast<double*>(mem + s.off[kEdges]);
  d.wfp = reinterpret_cast<float*>(mem + s.off[kWfp]);
  d.lines = reinterpret_cast<LineGeo*>(mem + s.off[kLines]);
  d.pint = reinterpret_cast<float*>(mem + s.off[kPint]);
  d.xrow = reinterpret_cast<float*>(mem + s.off[kXrow]);
  d.xcol = reinterpret_cast<float*>(mem + s.off[kXcol]);
  d.sino = reinterpret_cast<float*>(mem + s.off[kSino]);
  d.part = reinterpret_cast<float*>(mem + s.off[kPart]);
  d.prow = reinterpret_cast<float4*>(mem + s.off[kPrefRow]);
  d.pcol = reinterpret_cast<float4*>(mem + s.off[kPrefCol]);
  d.psino = reinterpret_cast<double2*>(mem + s.off[kPrefSino]);
  d.krange = reinterpret_cast<int2*>(mem + s.off[kKnotRange]);
  d.img_off = s.img_off;
  d.sino_off = s.sino_off;
  return d;
}

// First work item of instance i in a phase (read from the table, not the
// copy, so a runtime phase index does not put the copy in scratch memory).
__device__ __forceinline__ int item0_of(const char* __restrict__ mem, int i, int phase) {
  return reinterpret_cast<const BatchDesc*>(mem)[i].item0[phase];
}

__global__ __launch_bounds__(kThreads) void k_batch_setup(
    char* __restrict__ mem, const float* const* __restrict__ angles, int n_inst, int total) {
  const BatchDesc* __restrict__ desc = reinterpret_cast<const BatchDesc*>(mem);
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    // setup items: one per (instance, view); item0[kLayoutImage] is unused here,
    // so the per-instance view offsets are recomputed by a linear scan
    int i = 0, base = 0;
    while (i < n_inst - 1 && item >= base + desc[i].g.V) {
      base += desc[i].g.V;
      ++i;
    }
    const BatchInst d = load_inst(mem, i);
    setup2d_view(d.g, angles[i], d.views, d.edges, d.wfp, d.lines, d.pint, item - base);
    __syncthreads();
  }
}

// Per (instance, view): the knot range every 64 x 64 pixel tile of the
// block items needs (knot_range), for the LDS staging of k_batch_quads_prefix.
__global__ __launch_bounds__(kThreads) void k_batch_ranges(char* __restrict__ mem, int n_inst,
                                                           int total) {
  const BatchDesc* __restrict__ desc = reinterpret_cast<const BatchDesc*>(mem);
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    int i = 0, base = 0;
    while (i < n_inst - 1 && item >= base + desc[i].g.V) {
      base += desc[i].g.V;
      ++i;
    }
    const BatchInst d = load_inst(mem, i);
    const int v = item - base;
    const int tx = (d.g.W + kQuadTile - 1) / kQuadTile;
    const int tiles = tx * ((d.g.H + kQuadTile - 1) / kQuadTile);
    const ViewGeo vg = d.views[v];
    for (int t = threadIdx.x; t < tiles; t += blockDim.x) {
      const int c0 = (t % tx) * kQuadTile, r0 = (t / tx) * kQuadTile;
      d.krange[static_cast<size_t>(v) * tiles + t] = knot_range(
          d.g, vg, d.edges + static_cast<size_t>(v) * (d.g.D + 1),
          d.lines + static_cast<size_t>(v) * d.g.Lmax, r0, min(r0 + kQuadTile, d.g.H), c0,
          min(c0 + kQuadTile, d.g.W));
    }
  }
}

template <bool FOLD_W>
__global__ __launch_bounds__(kThreads) void k_batch_layout(
    char* __restrict__ mem, int n_inst, int phase, int total,
    const float* __restrict__ in) {
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, phase, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, phase);
    if (phase == kLayoutImage) {
      // item = (image row, instance tile); threads stride over the row
      const int h = local % d.g.H;
      const int b0 = (local / d.g.H) * kBT;
      for (int w = threadIdx.x; w < d.g.W; w += blockDim.x)
        interleave_image_px<kBT>(d.g, in + d.img_off, d.xrow, d.xcol, h, w, b0);
    } else {
      // item = (view, instance tile); threads stride over the bins
      const int v = local % d.g.V;
      const int b0 = (local / d.g.V) * kBT;
      for (int k = threadIdx.x; k < d.g.D; k += blockDim.x)
        interleave_sino_bin<kBT, FOLD_W>(d.g, d.views, d.wfp, in + d.sino_off, d.sino, v, k, b0);
    }
  }
}

// Running-sum items: (line, instance), the H rows first, then the W columns
// read bottom-up; one workgroup per line.
__global__ __launch_bounds__(kThreads) void k_batch_prefix(
    char* __restrict__ mem, int n_inst, int total, const float* __restrict__ in) {
  __shared__ double lds[kThreads];
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, kLayoutPrefix, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, kLayoutPrefix);
    const int H = d.g.H, W = d.g.W;
    const int line = local % (H + W);
    const int b = local / (H + W);
    const float* __restrict__ img = in + d.img_off + static_cast<size_t>(b) * H * W;
    if (line < H)
      prefix_line(img + static_cast<size_t>(line) * W, 1, W,
                  d.prow + (static_cast<size_t>(b) * H + line) * prefix_entries(W), lds);
    else
      prefix_line(img + static_cast<size_t>(H - 1) * W + (line - H), -static_cast<long long>(W),
                  H, d.pcol + (static_cast<size_t>(b) * W + line - H) * prefix_entries(H), lds);
  }
}

// Gather-bins items: (block of kBinsPerWave bins, view, instance, line
// chunk), chunk fastest, one wavefront per item (a 64-thread workgroup), so
// instances whose bin count is not a multiple of 256 keep their lanes busy.
// Every instance's item count and first item are multiples of kSplit, so with
// a grid that is a multiple of 8 the chunk follows the XCD.
constexpr int kBinsWave = 64;  // threads per gather-bins workgroup

// PREFIX: the bins' line integrals come from the running-sum tables
// (bins_accumulate_prefix) instead of the pixel walk.
template <int WM, bool PREFIX>
__global__ __launch_bounds__(kBinsWave) void k_batch_bins(
    char* __restrict__ mem, int n_inst, int total) {
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, kGatherBins, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, kGatherBins);
    const int chunk = local % kSplit;
    int rest = local / kSplit;
    const int kb_n = (d.g.D + kBinsPerWave - 1) / kBinsPerWave;
    const int lane = static_cast<int>(threadIdx.x);
    const int k = (rest % kb_n) * kBinsPerWave + lane;
    const bool edge_lane = lane == 63;
    rest /= kb_n;
    const int v = rest % d.g.V;
    const int b0 = (rest / d.g.V) * kBT;
    // the whole wave works every item (the edge exchange needs it); lanes
    // past the last bin use clamped edges and write nothing
    const ViewGeo vg = d.views[v];
    const int L = vg.xaxis ? d.g.H : d.g.W;
    const int N = vg.xaxis ? d.g.W : d.g.H;
    const double e1 = d.edges[static_cast<size_t>(v) * (d.g.D + 1) + min(k, d.g.D)];
    float acc[kBT];
#pragma unroll
    for (int b = 0; b < kBT; ++b) acc[b] = 0.0f;
    if constexpr (PREFIX) {
      const float4* __restrict__ tab =
          (vg.xaxis ? d.prow : d.pcol) + static_cast<size_t>(b0) * L * prefix_entries(N);
      acc[0] = bins_accumulate_prefix<WM>(e1, d.lines + static_cast<size_t>(v) * d.g.Lmax,
                                          d.pint + static_cast<size_t>(v) * d.g.Lmax, tab,
                                          bp_const(d.g), L * chunk / kSplit,
                                          L * (chunk + 1) / kSplit, N);
    } else {
      bins_accumulate<kBT, WM>(d.g, vg, e1, d.lines + static_cast<size_t>(v) * d.g.Lmax,
                               d.pint + static_cast<size_t>(v) * d.g.Lmax,
                               (vg.xaxis ? d.xrow : d.xcol) + b0, bp_const(d.g),
                               L * chunk / kSplit, L * (chunk + 1) / kSplit, N, acc, edge_lane);
    }
    if (edge_lane || k >= d.g.D) continue;
#pragma unroll
    for (int b = 0; b < kBT; ++b)
      d.part[((static_cast<size_t>(chunk) * d.g.Bp + b0 + b) * d.g.V + v) * d.g.D + k] = acc[b];
  }
}

// Reduce items: (view, bin block, instance): chunk sums in chunk order, times
// the bin weight, detector order (the arithmetic of k_bins_reduce).
template <int WM>
__global__ __launch_bounds__(kThreads) void k_batch_reduce(
    char* __restrict__ mem, int n_inst, int total, float* __restrict__ out) {
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, kReduceBins, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, kReduceBins);
    const int kb_n = (d.g.D + kThreads - 1) / kThreads;
    const int k = (local % kb_n) * kThreads + threadIdx.x;
    const int v = (local / kb_n) % d.g.V;
    const int b = local / (kb_n * d.g.V);
    if (k >= d.g.D) continue;
    const size_t plane = static_cast<size_t>(d.g.Bp) * d.g.V * d.g.D;
    const size_t at = (static_cast<size_t>(b) * d.g.V + v) * d.g.D + k;
    float sum = d.part[at];
#pragma unroll
    for (int c = 1; c < kSplit; ++c) sum += d.part[c * plane + at];
    const ViewGeo vg = d.views[v];
    const int kk = vg.flip ? d.g.D - 1 - k : k;
    const float wk = (WM == W_FP) ? d.wfp[static_cast<size_t>(v) * d.g.D + k] : 1.0f;
    out[d.sino_off + (static_cast<size_t>(b) * d.g.V + v) * d.g.D + kk] = sum * wk;
  }
}

template <int WM>
__global__ __launch_bounds__(kThreads) void k_batch_pixels(
    char* __restrict__ mem, int n_inst, int total, float* __restrict__ out,
    int per_pixel) {
  __shared__ PixelTileLds<kTile> lds;
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, kGatherPixels, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, kGatherPixels);
    const int tx = (d.g.W + kTile - 1) / kTile, ty = (d.g.H + kTile - 1) / kTile;
    const int t = local % (tx * ty);
    const int b0 = (local / (tx * ty)) * kBT;
    if (per_pixel) {
      const int c = (t % tx) * kTile + (threadIdx.x % kTile);
      const int r = (t / tx) * kTile + (threadIdx.x / kTile);
      if (c < d.g.W && r < d.g.H)
        pixels_body<kBT, WM>(d.g, d.views, d.edges, d.lines, d.pint, d.sino, out + d.img_off,
                             r, c, b0);
    } else {
      pixels_tile_body<kBT, WM, kTile>(d.g, d.views, d.edges, d.lines, d.pint, d.sino,
                                       out + d.img_off, t % tx, t / tx, b0, lds);
    }
  }
}

// Sinogram running-sum items: the layout items (view, instance), one
// workgroup per view.
template <bool FOLD_W>
__global__ __launch_bounds__(kThreads) void k_batch_prefix_sino(
    char* __restrict__ mem, int n_inst, int total, const float* __restrict__ in) {
  __shared__ double lds[kThreads];
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, kLayoutSino, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, kLayoutSino);
    const int v = local % d.g.V, b = local / d.g.V;
    prefix_sino_view<FOLD_W>(
        d.g, d.views, d.edges, d.wfp, in + d.sino_off + static_cast<size_t>(b) * d.g.V * d.g.D,
        v, d.psino + (static_cast<size_t>(b) * d.g.V + v) * (d.g.D + 1), lds);
  }
}

// Block items through the running sums over the bins: per view, the knots the
// item's 64 x 64 pixel tile needs are staged in LDS (one 16-byte load per
// thread, double-buffered: view v + 1's load is in flight while view v is
// walked, one barrier per view), and every 4-pixel run's merge walk reads
// them there.  A view whose range exceeds the buffer is walked from global
// memory; both give the same bits (the walk never leaves the range).  The
// walk keeps its positions and sums in FP32 relative to the run, so the
// kernel fits 5 waves per SIMD (96 VGPRs); at 4 (101) it measured 8 % slower.
constexpr int kKnotCap = kQuadThreads;  // knots staged per view

template <int WM>
__global__ __launch_bounds__(kQuadThreads)
__attribute__((amdgpu_waves_per_eu(5))) void k_batch_quads_prefix(
    char* __restrict__ mem, int n_inst, int total, float* __restrict__ out) {
  __shared__ double2 knots[2][kKnotCap];
  const int tid = static_cast<int>(threadIdx.x);
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, kGatherQuads, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, kGatherQuads);
    const int tx = (d.g.W + kQuadTile - 1) / kQuadTile;
    const int tiles = tx * ((d.g.H + kQuadTile - 1) / kQuadTile);
    const int t = local % tiles;
    const int b = local / tiles;
    const int c0 = (t % tx) * kQuadTile + kBlockQ * (tid % kQuadSide);
    const int r0 = (t / tx) * kQuadTile + kBlockQ * (tid / kQuadSide);
    const bool active = c0 < d.g.W && r0 < d.g.H;
    const int nr = min(kBlockQ, d.g.H - r0), nc = min(kBlockQ, d.g.W - c0);
    const double2* __restrict__ tab = d.psino + static_cast<size_t>(b) * d.g.V * (d.g.D + 1);
    const int2* __restrict__ rng = d.krange + t;  // view v at rng[v * tiles]
    float acc[kBlockQ][kBlockQ];
#pragma unroll
    for (int dr = 0; dr < kBlockQ; ++dr)
#pragma unroll
      for (int dc = 0; dc < kBlockQ; ++dc) acc[dr][dc] = 0.0f;
    // stage view 0
    int2 r = rng[0];
    bool staged = r.y - r.x + 1 <= kKnotCap;
    if (staged && tid <= r.y - r.x) knots[0][tid] = tab[r.x + tid];
    __syncthreads();
    for (int v = 0; v < d.g.V; ++v) {
      // the next view's knot for this thread, in flight during this view's walks
      const int vn = min(v + 1, d.g.V - 1);
      const int2 rn = rng[static_cast<size_t>(vn) * tiles];
      const bool staged_n = rn.y - rn.x + 1 <= kKnotCap;
      double2 kn = make_double2(0.0, 0.0);
      const bool mine = v + 1 < d.g.V && staged_n && tid <= rn.y - rn.x;
      if (mine) kn = tab[static_cast<size_t>(vn) * (d.g.D + 1) + rn.x + tid];
      if (active) {
        const ViewGeo vg = d.views[v];
        const LineGeo* __restrict__ lv = d.lines + static_cast<size_t>(v) * d.g.Lmax;
        const float* __restrict__ pv = d.pint + static_cast<size_t>(v) * d.g.Lmax;
        if (staged)
          block_view_prefix<WM, kBlockQ>(
              d.g, vg, lv, pv, (const __attribute__((address_space(3))) double*)(knots[v & 1]),
              r.x, r.y, r0, c0, nr, nc, acc);
        else
          block_view_prefix<WM, kBlockQ>(
              d.g, vg, lv, pv,
              reinterpret_cast<const double*>(tab + static_cast<size_t>(v) * (d.g.D + 1)), 0,
              d.g.D, r0, c0, nr, nc, acc);
      }
      if (mine) knots[(v + 1) & 1][tid] = kn;
      __syncthreads();
      r = rn;
      staged = staged_n;
    }
    if (active) {
      float* __restrict__ o = out + d.img_off + static_cast<size_t>(b) * d.g.H * d.g.W;
#pragma unroll
      for (int dr = 0; dr < kBlockQ; ++dr)
#pragma unroll
        for (int dc = 0; dc < kBlockQ; ++dc)
          if (dr < nr && dc < nc) o[static_cast<size_t>(r0 + dr) * d.g.W + c0 + dc] = acc[dr][dc];
    }
  }
}

// Block items: (instance, 64 x 64 pixel tile), one thread per 4 x 4 pixels;
// KW > 0 reads each line's bins as one window of KW bins.
template <int WM, int KW>
__global__ __launch_bounds__(kQuadThreads)
__attribute__((amdgpu_waves_per_eu(KW > 0 ? 4 : 1))) void k_batch_quads(
    char* __restrict__ mem, int n_inst, int total, float* __restrict__ out) {
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int i = find_inst(mem, n_inst, kGatherQuads, item);
    const BatchInst d = load_inst(mem, i);
    const int local = item - item0_of(mem, i, kGatherQuads);
    const int tx = (d.g.W + kQuadTile - 1) / kQuadTile;
    const int ty = (d.g.H + kQuadTile - 1) / kQuadTile;
    const int t = local % (tx * ty);
    const int b0 = (local / (tx * ty)) * kBT;
    const int c0 = (t % tx) * kQuadTile + kBlockQ * (threadIdx.x % kQuadSide);
    const int r0 = (t / tx) * kQuadTile + kBlockQ * (threadIdx.x / kQuadSide);
    if (c0 < d.g.W && r0 < d.g.H)
      blocks_body<kBT, WM, kBlockQ, KW>(d.g, d.views, d.edges, d.lines, d.pint, d.sino,
                                        out + d.img_off, r0, c0, b0);
  }
}

}  // namespace
}  // namespace dd

using namespace dd;

struct dd2d_batch {
  std::vector<dd2d_batch_item> items;
  std::vector<BatchDesc> inst;   // host copy of the descriptors
  std::vector<int> order;        // instance order (heaviest first)
  size_t desc_off = 0, angles_off = 0, bytes = 0;
  size_t total_views = 0;
  int total[kPhases] = {};
  int grid = 0;                  // workgroups per launch (0 = one per work item)
  int variant = DD_ADJ_QUAD;     // sinogram -> image kernel (per pixel, 16x16 tiles, 4x4 blocks)
  bool fwd_prefix = false;       // image -> sinogram through running sums (DD_FWD_PREFIX)
  char* dev = nullptr;           // device memory given to setup()
};

namespace {

size_t take(size_t* off, size_t bytes) {
  const size_t at = *off;
  *off = align_up(*off + bytes, static_cast<size_t>(256));
  return at;
}

// Workgroups of one launch: the set grid (each workgroup strides over the
// items), or by default one workgroup per item, which measured fastest on
// the 64-instance fan-arc batch (21.3 ms per FP + adjoint vs 23.5 ms at 16
// workgroups per compute unit and 24.8 ms at 8).
int grid_for(const dd2d_batch* b, int total) {
  const int g = b->grid > 0 ? b->grid : total;
  return std::max(1, std::min(g, total));
}

// Bins per window of DD_ADJ_BLOCK_WINDOW: a 4-pixel run meets up to ~6 bins
// of these geometries, so 8 rarely leaves a bin to the serial tail.
constexpr int kBlockWindow = 8;

template <int KW>
void launch_quads(const dd2d_batch* b, int op, char* inst, int n, int qt, float* output,
                  hipStream_t st) {
  if (op == DD_PROJ_T)
    hipLaunchKernelGGL((k_batch_quads<W_FP, KW>), dim3(grid_for(b, qt)), dim3(kQuadThreads), 0,
                       st, inst, n, qt, output);
  else if (op == DD_BPROJ)
    hipLaunchKernelGGL((k_batch_quads<W_BP, KW>), dim3(grid_for(b, qt)), dim3(kQuadThreads), 0,
                       st, inst, n, qt, output);
  else
    hipLaunchKernelGGL((k_batch_quads<W_BPW, KW>), dim3(grid_for(b, qt)), dim3(kQuadThreads), 0,
                       st, inst, n, qt, output);
}

}  // namespace

extern "C" {

dd2d_batch* dd2d_batch_create(const dd2d_batch_item* items, int n) {
  if (!items || n <= 0) return nullptr;
  auto* b = new (std::nothrow) dd2d_batch();
  if (!b) return nullptr;
  b->items.assign(items, items + n);
  for (int i = 0; i < n; ++i) {
    const auto& it = items[i];
    if (it.scan < DD_PARA || it.scan > DD_FAN_ARC || it.views <= 0 || it.batch <= 0 ||
        it.geom.width <= 0 || it.geom.height <= 0 || it.geom.num_det <= 0 || !it.angles ||
        it.image_offset < 0 || it.sino_offset < 0) {
      delete b;
      return nullptr;
    }
  }
  // heaviest instances first, so the static item lists balance
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
    const auto& a = items[x];
    const auto& c = items[y];
    return 1.0 * a.views * a.geom.width * a.geom.height * a.batch >
           1.0 * c.views * c.geom.width * c.geom.height * c.batch;
  });
  size_t off = 0;
  b->desc_off = take(&off, sizeof(BatchDesc) * n);  // 0: kernels find it at the block base
  b->angles_off = take(&off, sizeof(float*) * n);
  b->inst.resize(n);
  int acc[kPhases] = {};
  for (int j = 0; j < n; ++j) {
    const auto& it = items[b->order[j]];
    BatchDesc d{};
    Geo2& g = d.g;
    g.W = it.geom.width;
    g.H = it.geom.height;
    g.D = it.geom.num_det;
    g.V = it.views;
    g.B = it.batch;
    g.BT = kBT;
    g.Bp = it.batch;
    g.Lmax = std::max(g.W, g.H);
    g.scan = it.scan;
    g.dx = it.geom.pix_x;
    g.dy = it.geom.pix_y;
    g.det = it.geom.det;
    g.sid = it.geom.iso_source;
    g.sdd = it.geom.source_detector;
    g.shx = it.geom.shift_x;
    g.shy = it.geom.shift_y;
    g.bin = it.geom.bin_shift;
    // geometry tables and layout scratch: byte offsets into the device block
    d.off[kViews] = take(&off, sizeof(ViewGeo) * g.V);
    d.off[kEdges] = take(&off, sizeof(double) * g.V * (g.D + 1));
    d.off[kWfp] = take(&off, sizeof(float) * g.V * g.D);
    d.off[kLines] = take(&off, sizeof(LineGeo) * g.V * g.Lmax);
    d.off[kPint] = take(&off, sizeof(float) * g.V * g.Lmax);
    const size_t img = static_cast<size_t>(g.W) * g.H * g.Bp;
    d.off[kXrow] = take(&off, sizeof(float) * img);
    d.off[kXcol] = take(&off, sizeof(float) * img);
    d.off[kSino] = take(&off, sizeof(float) * g.V * g.D * g.Bp);
    d.off[kPart] =
        take(&off, sizeof(float) * static_cast<size_t>(kSplit) * g.Bp * g.V * g.D);
    d.off[kPrefRow] = take(&off, sizeof(float4) * static_cast<size_t>(g.Bp) * g.H *
                                     prefix_entries(g.W));
    d.off[kPrefCol] = take(&off, sizeof(float4) * static_cast<size_t>(g.Bp) * g.W *
                                     prefix_entries(g.H));
    d.off[kPrefSino] =
        take(&off, sizeof(double2) * static_cast<size_t>(g.Bp) * g.V * (g.D + 1));
    d.off[kKnotRange] = take(&off, sizeof(int2) * static_cast<size_t>(g.V) *
                                       ((g.W + kQuadTile - 1) / kQuadTile) *
                                       ((g.H + kQuadTile - 1) / kQuadTile));
    d.img_off = it.image_offset;
    d.sino_off = it.sino_offset;
    const int tiles = it.batch;  // kBT == 1
    d.nitems[kLayoutImage] = g.H * tiles;
    d.nitems[kGatherBins] = kSplit * g.V * ((g.D + kBinsPerWave - 1) / kBinsPerWave) * tiles;
    d.nitems[kReduceBins] = g.V * ((g.D + kThreads - 1) / kThreads) * it.batch;
    d.nitems[kGatherQuads] =
        ((g.W + kQuadTile - 1) / kQuadTile) * ((g.H + kQuadTile - 1) / kQuadTile) * tiles;
    d.nitems[kLayoutSino] = g.V * tiles;
    d.nitems[kGatherPixels] = ((g.W + kTile - 1) / kTile) * ((g.H + kTile - 1) / kTile) * tiles;
    d.nitems[kLayoutPrefix] = (g.H + g.W) * tiles;
    for (int ph = 0; ph < kPhases; ++ph) {
      d.item0[ph] = acc[ph];
      acc[ph] += d.nitems[ph];
    }
    b->total_views += g.V;
    b->inst[j] = d;
  }
  for (int ph = 0; ph < kPhases; ++ph) b->total[ph] = acc[ph];
  b->bytes = off;
  return b;
}

size_t dd2d_batch_device_bytes(const dd2d_batch* b) { return b ? b->bytes : 0; }

void dd2d_batch_set_grid(dd2d_batch* b, int workgroups) {
  if (b) b->grid = workgroups > 0 ? workgroups : 0;
}

void dd2d_batch_set_variant(dd2d_batch* b, int variant) {
  if (!b) return;
  const int adj = variant & ~DD_FWD_PREFIX;
  b->variant = (adj == DD_ADJ_TILE16 || adj == DD_ADJ_QUAD || adj == DD_ADJ_BLOCK_WINDOW ||
                adj == DD_ADJ_BLOCK_PREFIX)
                   ? adj
                   : DD_ADJ_PER_PIXEL;
  b->fwd_prefix = (variant & DD_FWD_PREFIX) != 0;
}

dd_status dd2d_batch_setup(dd2d_batch* b, void* device_mem, size_t bytes, void* stream) {
  if (!b || !device_mem || bytes < b->bytes) return DD_NO_WORKSPACE;
  b->dev = static_cast<char*>(device_mem);
  auto st = static_cast<hipStream_t>(stream);
  const int n = static_cast<int>(b->inst.size());
  std::vector<const float*> angles(n);
  for (int j = 0; j < n; ++j) angles[j] = b->items[b->order[j]].angles;
  if (hipMemcpyAsync(b->dev + b->desc_off, b->inst.data(), sizeof(BatchDesc) * n,
                     hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(b->dev + b->angles_off, angles.data(), sizeof(float*) * n,
                     hipMemcpyHostToDevice, st) != hipSuccess)
    return DD_LAUNCH_FAIL;
  const int total = static_cast<int>(b->total_views);
  hipLaunchKernelGGL(k_batch_setup, dim3(grid_for(b, total)), dim3(kThreads), 0, st, b->dev,
                     reinterpret_cast<const float* const*>(b->dev + b->angles_off), n, total);
  hipLaunchKernelGGL(k_batch_ranges, dim3(grid_for(b, total)), dim3(kThreads), 0, st, b->dev, n,
                     total);
  if (hipGetLastError() != hipSuccess) return DD_LAUNCH_FAIL;
  // the host staging vectors die with this call: wait for the uploads
  return hipStreamSynchronize(st) == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

dd_status dd2d_batch_run(const dd2d_batch* b, int op, const float* input, float* output,
                         void* stream) {
  if (!b || !b->dev || !input || !output) return DD_BAD_ARG;
  if (op < DD_PROJ || op > DD_BPROJ_TW) return DD_BAD_ARG;
  for (const auto& it : b->items)
    if (it.scan == DD_PARA && op >= DD_BPROJ_W) return DD_UNSUPPORTED;
  auto st = static_cast<hipStream_t>(stream);
  const int n = static_cast<int>(b->inst.size());
  char* inst = b->dev;  // descriptor table at the block base
  const bool sino_out = (op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW);
  if (sino_out) {
    const int tot = b->total[kGatherBins], rtot = b->total[kReduceBins];
    // one-wave workgroups: four times the workgroups of the 256-thread phases
    const int bgrid = std::max(1, std::min(4 * grid_for(b, tot), tot));
    if (b->fwd_prefix && op != DD_BPROJ_TW) {
      // running sums of every line, then one lookup per (bin, line); the
      // weighted transpose keeps the pixel walk (its weight varies per pixel)
      ProfileScope prof(DD_KERNEL_IMAGE_TO_SINO, st);
      const int ptot = b->total[kLayoutPrefix];
      hipLaunchKernelGGL(k_batch_prefix, dim3(grid_for(b, ptot)), dim3(kThreads), 0, st, inst,
                         n, ptot, input);
      if (op == DD_PROJ) {
        hipLaunchKernelGGL((k_batch_bins<W_FP, true>), dim3(bgrid), dim3(kBinsWave), 0, st,
                           inst, n, tot);
        hipLaunchKernelGGL((k_batch_reduce<W_FP>), dim3(grid_for(b, rtot)), dim3(kThreads), 0,
                           st, inst, n, rtot, output);
      } else {
        hipLaunchKernelGGL((k_batch_bins<W_BP, true>), dim3(bgrid), dim3(kBinsWave), 0, st,
                           inst, n, tot);
        hipLaunchKernelGGL((k_batch_reduce<W_BP>), dim3(grid_for(b, rtot)), dim3(kThreads), 0,
                           st, inst, n, rtot, output);
      }
      return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
    }
    hipLaunchKernelGGL((k_batch_layout<false>), dim3(grid_for(b, b->total[kLayoutImage])),
                       dim3(kThreads), 0, st, inst, n, static_cast<int>(kLayoutImage),
                       b->total[kLayoutImage], input);
    ProfileScope prof(DD_KERNEL_IMAGE_TO_SINO, st);
    if (op == DD_PROJ) {
      hipLaunchKernelGGL((k_batch_bins<W_FP, false>), dim3(bgrid), dim3(kBinsWave), 0, st, inst, n,
                         tot);
      hipLaunchKernelGGL((k_batch_reduce<W_FP>), dim3(grid_for(b, rtot)), dim3(kThreads), 0,
                         st, inst, n, rtot, output);
    } else if (op == DD_BPROJ_T) {
      hipLaunchKernelGGL((k_batch_bins<W_BP, false>), dim3(bgrid), dim3(kBinsWave), 0, st, inst, n,
                         tot);
      hipLaunchKernelGGL((k_batch_reduce<W_BP>), dim3(grid_for(b, rtot)), dim3(kThreads), 0,
                         st, inst, n, rtot, output);
    } else {
      hipLaunchKernelGGL((k_batch_bins<W_BPW, false>), dim3(bgrid), dim3(kBinsWave), 0, st, inst, n,
                         tot);
      hipLaunchKernelGGL((k_batch_reduce<W_BPW>), dim3(grid_for(b, rtot)), dim3(kThreads), 0,
                         st, inst, n, rtot, output);
    }
  } else {
    if (b->variant == DD_ADJ_BLOCK_PREFIX && op != DD_BPROJ_W) {
      // running sums of every view, then one merge walk per 4-pixel run
      ProfileScope prof(DD_KERNEL_SINO_TO_IMAGE, st);
      const int lt = b->total[kLayoutSino], qt = b->total[kGatherQuads];
      if (op == DD_PROJ_T)
        hipLaunchKernelGGL(k_batch_prefix_sino<true>, dim3(grid_for(b, lt)), dim3(kThreads), 0,
                           st, inst, n, lt, input);
      else
        hipLaunchKernelGGL(k_batch_prefix_sino<false>, dim3(grid_for(b, lt)), dim3(kThreads), 0,
                           st, inst, n, lt, input);
      if (op == DD_PROJ_T)
        hipLaunchKernelGGL((k_batch_quads_prefix<W_FP>), dim3(grid_for(b, qt)),
                           dim3(kQuadThreads), 0, st, inst, n, qt, output);
      else
        hipLaunchKernelGGL((k_batch_quads_prefix<W_BP>), dim3(grid_for(b, qt)),
                           dim3(kQuadThreads), 0, st, inst, n, qt, output);
      return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
    }
    if (op == DD_PROJ_T)
      hipLaunchKernelGGL((k_batch_layout<true>), dim3(grid_for(b, b->total[kLayoutSino])),
                         dim3(kThreads), 0, st, inst, n, static_cast<int>(kLayoutSino),
                         b->total[kLayoutSino], input);
    else
      hipLaunchKernelGGL((k_batch_layout<false>), dim3(grid_for(b, b->total[kLayoutSino])),
                         dim3(kThreads), 0, st, inst, n, static_cast<int>(kLayoutSino),
                         b->total[kLayoutSino], input);
    ProfileScope prof(DD_KERNEL_SINO_TO_IMAGE, st);
    if (b->variant == DD_ADJ_QUAD || b->variant == DD_ADJ_BLOCK_WINDOW ||
        b->variant == DD_ADJ_BLOCK_PREFIX) {
      const int qt = b->total[kGatherQuads];
      if (b->variant == DD_ADJ_QUAD)
        launch_quads<0>(b, op, inst, n, qt, output, st);
      else  // DD_ADJ_BLOCK_WINDOW, and BPROJ_W of DD_ADJ_BLOCK_PREFIX
        launch_quads<kBlockWindow>(b, op, inst, n, qt, output, st);
      return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
    }
    const int tot = b->total[kGatherPixels];
    if (op == DD_PROJ_T)
      hipLaunchKernelGGL((k_batch_pixels<W_FP>), dim3(grid_for(b, tot)), dim3(kThreads), 0, st,
                         inst, n, tot, output, b->variant != DD_ADJ_TILE16);
    else if (op == DD_BPROJ)
      hipLaunchKernelGGL((k_batch_pixels<W_BP>), dim3(grid_for(b, tot)), dim3(kThreads), 0, st,
                         inst, n, tot, output, b->variant != DD_ADJ_TILE16);
    else
      hipLaunchKernelGGL((k_batch_pixels<W_BPW>), dim3(grid_for(b, tot)), dim3(kThreads), 0, st,
                         inst, n, tot, output, b->variant != DD_ADJ_TILE16);
  }
  return hipGetLastError() == hipSuccess ? DD_OK : DD_LAUNCH_FAIL;
}

void dd2d_batch_destroy(dd2d_batch* b) { delete b; }

}  // extern "C"
