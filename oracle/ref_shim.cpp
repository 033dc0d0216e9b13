// extern "C" entry points over the reference's own CPU projector functions.
//
// Test infrastructure only: this file is compiled together with the
// reference's CPU sources (deeprecon/cpu/dis2d/{FanArc,FanFlat,Para}DisCPU2d.cpp,
// read in place under the reference checkout) by oracle/Makefile into
// oracle/_ref/libdd_ref2d.so.  Nothing in the product links or loads it; the
// tests use it as the checker for the C restatement in dd_oracle.cpp and
// bench.py times it as the reference CPU baseline.
//
// Signatures followed: deeprecon/include/dis2d/FanArcDisCPU2d.h:3-23,
// FanFlatDisCPU2d.h:3-23, ParaDisCPU2d.h:4-17.
#include <cstdint>
#include <cstring>
#include <vector>

#include "FanArcDisCPU2d.h"
#include "FanFlatDisCPU2d.h"
#include "ParaDisCPU2d.h"

namespace {
using Fn = void (*)(float*, float*, float*, int, int, int, int, int, float,
                    float, float, float, float, float, float, float);

// [geometry][op]; geometry 0 = para, 1 = fan flat, 2 = fan arc;
// op 0..5 = Proj, ProjTrans, BackProj, BackProjTrans, BackProjWeighted,
// BackProjTransWeighted (para has no weighted pair).
const Fn kTable[3][6] = {
    {ProjParaDisCPU2d, ProjTransParaDisCPU2d, BackProjParaDisCPU2d,
     BackProjTransParaDisCPU2d, nullptr, nullptr},
    {ProjFanFlatDisCPU2d, ProjTransFanFlatDisCPU2d, BackProjFanFlatDisCPU2d,
     BackProjTransFanFlatDisCPU2d, BackProjWeightedFanFlatDisCPU2d,
     BackProjTransWeightedFanFlatDisCPU2d},
    {ProjFanArcDisCPU2d, ProjTransFanArcDisCPU2d, BackProjFanArcDisCPU2d,
     BackProjTransFanArcDisCPU2d, BackProjWeightedFanArcDisCPU2d,
     BackProjTransWeightedFanArcDisCPU2d},
};

std::vector<int64_t> g_i, g_j;
std::vector<float> g_v;
}  // namespace

extern "C" {

// Runs one reference op.  `image` is [B,H,W], `proj` is [B,V,D]; the output
// side must be zero-filled by the caller, as RadonTorch.cpp:68 does.
// Returns 0 on success, -1 for an unknown geometry/op pair.
int ddref_op2d(int geom, int op, float* image, float* proj, float* angles,
               int batch, int width, int height, int views, int dets,
               float dx, float dy, float det, float sid, float sdd,
               float shift_x, float shift_y, float bin_shift) {
  if (geom < 0 || geom > 2 || op < 0 || op > 5) return -1;
  Fn f = kTable[geom][op];
  if (!f) return -1;
  f(image, proj, angles, batch, width, height, views, dets, dx, dy, det, sid,
    sdd, shift_x, shift_y, bin_shift);
  return 0;
}

// Builds the reference COO system matrix and keeps it until the next call;
// returns the number of non-zeros.
int64_t ddref_sysmatrix2d(int geom, const float* angles, int width, int height,
                          int views, int dets, float dx, float dy, float det,
                          float sid, float sdd, float shift_x, float shift_y,
                          float bin_shift) {
  g_i.clear();
  g_j.clear();
  g_v.clear();
  if (geom == 0)
    GetSysMatrixPara2dKernel(g_i, g_j, g_v, angles, width, height, views, dets,
                             dx, dy, det, sid, sdd, shift_x, shift_y,
                             bin_shift);
  else if (geom == 1)
    GetSysMatrixFanFlat2dKernel(g_i, g_j, g_v, angles, width, height, views,
                                dets, dx, dy, det, sid, sdd, shift_x, shift_y,
                                bin_shift);
  else if (geom == 2)
    GetSysMatrixFanArc2dKernel(g_i, g_j, g_v, angles, width, height, views,
                               dets, dx, dy, det, sid, sdd, shift_x, shift_y,
                               bin_shift);
  else
    return -1;
  return static_cast<int64_t>(g_v.size());
}

// Copies the matrix built by the last ddref_sysmatrix2d call.
void ddref_sysmatrix_copy(int64_t* rows, int64_t* cols, float* vals) {
  std::memcpy(rows, g_i.data(), g_i.size() * sizeof(int64_t));
  std::memcpy(cols, g_j.data(), g_j.size() * sizeof(int64_t));
  std::memcpy(vals, g_v.data(), g_v.size() * sizeof(float));
}

}  // extern "C"
