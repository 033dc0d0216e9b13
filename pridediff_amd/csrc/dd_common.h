// Shared device-side pieces of the distance-driven kernels (gfx950, wave64).
//
// Geometry convention (deeprecon/cpu/dis2d/FanArcDisCPU2d.cpp:68-223):
// every view projects pixel edges and detector-bin edges onto one common
// axis (y = 0 when cos^2 > 1/2, else x = 0).  Along an image line parallel to
// that axis the pixel edges map affinely, a(c) = A0 + c * P, so a bin edge e
// sits at the line coordinate t = (e - A0) / P.  All ops are then built from
// the overlap of [t_k, t_k+1] with pixel [c, c+1].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dd {

// Per-view geometry shared by all kernels of one call.
struct ViewGeo {
  double sx, sy;     // source position (fan); unused by parallel beam
  double s, c;       // sin, cos of the view angle (double)
  double theta_ref;  // theta + bin shift (arc start angle)
  int xaxis;         // 1: common axis y = 0, lines are image rows
  int flip;          // detector index runs against the walk order
  float sxf, syf;    // source position in float, for FBP distance weights
  float cref, sref;  // cos, sin of theta_ref in float (first-bin guesses)
};

// Per-(view, line) pixel-edge map on the common axis.
struct LineGeo {
  double a0;    // axis coordinate of the line's first pixel edge
  double inv_p; // 1 / pixel interval on the axis
};

// Load of a table the running kernel never writes, through the constant
// address space: at a wave-uniform address it becomes a scalar load even
// where the kernel stores floats through pointers the compiler cannot tell
// apart from the table's (the batch kernels rebuild every pointer from one
// device block).
template <typename T>
__device__ __forceinline__ T const_load(const T* p) {
  return *(const __attribute__((address_space(4))) T*)(p);
}

// Integer pixel + fractional position of an axis point on one line.
// The single definition used by the forward and the adjoint kernels: both
// see bit-identical overlaps, so the two operators are exact transposes.
struct EdgePos {
  int c;
  float f;
};

// Position t (in pixel units from the first pixel edge) split into pixel
// index and fraction, clamped to [-2, n + 2].
__device__ __forceinline__ EdgePos edge_pos_t(double t, int n) {
  t = fmin(fmax(t, -2.0), static_cast<double>(n) + 2.0);
  double fl = floor(t);
  EdgePos r;
  r.c = static_cast<int>(fl);
  r.f = static_cast<float>(t - fl);
  return r;
}

__device__ __forceinline__ EdgePos edge_pos(double e, double a0, double inv_p,
                                            int n) {
  return edge_pos_t((e - a0) * inv_p, n);
}

// Overlap, in pixel units, of the bin [p1, p2] with pixel c (c1 <= c <= c2).
__device__ __forceinline__ float overlap_in_pixel(EdgePos p1, EdgePos p2,
                                                  int c) {
  float lo = (c == p1.c) ? p1.f : 0.0f;
  float hi = (c == p2.c) ? p2.f : 1.0f;
  return hi - lo;
}

// Axis choice exactly as the reference takes it: FP32 cosine squared in FP32
// (`cosVal * cosVal > 0.5`, FanArcDisCPU2d.cpp:76; flips :78-84, :148-154).
__device__ __forceinline__ void choose_axis(float theta, int* xaxis,
                                            int* flip) {
  const float cf = static_cast<float>(cos(static_cast<double>(theta)));
  const float sf = static_cast<float>(sin(static_cast<double>(theta)));
  const float c2 = cf * cf;
  *xaxis = (c2 > 0.5f) ? 1 : 0;
  *flip = *xaxis ? !(cf >= 0.0f) : !(sf >= 0.0f);
}

template <typename T>
__host__ __device__ __forceinline__ T align_up(T x, T a) {
  return (x + a - 1) / a * a;
}

// BT contiguous floats, loaded with the widest aligned vector loads.
template <int BT>
__device__ __forceinline__ void load_tile(const float* __restrict__ p,
                                          float (&v)[BT]) {
  if constexpr (BT % 4 == 0) {
#pragma unroll
    for (int i = 0; i < BT / 4; ++i) {
      float4 q = reinterpret_cast<const float4*>(p)[i];
      v[4 * i + 0] = q.x;
      v[4 * i + 1] = q.y;
      v[4 * i + 2] = q.z;
      v[4 * i + 3] = q.w;
    }
  } else if constexpr (BT == 2) {
    float2 q = *reinterpret_cast<const float2*>(p);
    v[0] = q.x;
    v[1] = q.y;
  } else {
    v[0] = p[0];
  }
}

// Two neighbouring pixels (or voxels) of an interleaved layout whose pixel
// stride is `stride` floats: BT instances at p and BT at p + stride.  When the
// two tiles are adjacent (stride == BT, always so for BT <= 2, where the
// padded batch equals the tile) they come in one vector load; p is then only
// 4-byte aligned.
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
template <int BT>
__device__ __forceinline__ void load_pair(const float* __restrict__ p, int stride,
                                          float (&x0)[BT], float (&x1)[BT]) {
  if constexpr (BT == 1) {
    if (stride == 1) {
      const f2u q = *reinterpret_cast<const f2u*>(p);
      x0[0] = q.x;
      x1[0] = q.y;
      return;
    }
  } else if constexpr (BT == 2) {
    if (stride == 2) {
      const f4u q = *reinterpret_cast<const f4u*>(p);
      x0[0] = q.x;
      x0[1] = q.y;
      x1[0] = q.z;
      x1[1] = q.w;
      return;
    }
  }
  load_tile<BT>(p, x0);
  load_tile<BT>(p + stride, x1);
}

template <int BT>
__device__ __forceinline__ void store_tile(float* __restrict__ p,
                                           const float (&v)[BT]) {
  if constexpr (BT % 4 == 0) {
#pragma unroll
    for (int i = 0; i < BT / 4; ++i)
      reinterpret_cast<float4*>(p)[i] =
          make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
  } else if constexpr (BT == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
    p[0] = v[0];
  }
}

}  // namespace dd
