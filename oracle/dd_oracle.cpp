// CPU restatement of the deeprecon distance-driven 2D projectors.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP kernels: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
// The product path (pridediff_amd) never links, loads or falls back to it.
//
// It restates the per-view merge walk of deeprecon/cpu/dis2d/*.cpp:
//   * fan-beam, equiangular arc detector: FanArcDisCPU2d.cpp:42-223 (Proj),
//     :226-423 (ProjTrans), :426-618 (BackProj<FBPWEIGHT>),
//     :621-797 (BackProjTrans<FBPWEIGHT>), :1057-1225 (system matrix);
//   * fan-beam, flat detector: FanFlatDisCPU2d.cpp:35-1045, :1048-1215;
//   * parallel beam: ParaDisCPU2d.cpp:20-159 (Proj<TRANSPOSE>),
//     :162-317 (BackProj<TRANSPOSE>), :320-485 (op table), :488-612.
// The reference writes the walk out once per op; here one walker per
// geometry emits (line, pixel, bin, overlap) segments and every op is a
// consumer of those segments, so all ops share one geometry by construction.
//
// Two instantiations:
//   Real = float  mirrors the reference's evaluation: FP32 stored values with
//                 double intermediates where the reference promotes
//                 (NumDet / 2.0, ::sin/::cos/::tan/::atan on double);
//   Real = double is the FP64 oracle (every quantity in double).
//
// Pinned by: tests/test_oracle_2d.py against oracle/_ref (the reference's
// own CPU code compiled in place) and the committed golden vectors under
// tests/golden/, plus closed forms (parallel-beam mass conservation, chord
// length of a constant image), adjointness of Proj/ProjTrans and a brute-force
// exact polygon-overlap matrix on tiny inputs.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

enum Geom { kPara = 0, kFanFlat = 1, kFanArc = 2 };
enum Op {
  kProj = 0,          // image -> sinogram, weight = FP coefficient
  kProjTrans = 1,     // sinogram -> image, same weight (matched adjoint)
  kBackProj = 2,      // sinogram -> image, FBP backprojection weight
  kBackProjTrans = 3, // image -> sinogram, same weight
  kBackProjW = 4,     // BackProj with the FBP distance weight
  kBackProjTransW = 5 // its transpose
};

template <typename Real>
struct Params {
  int B, W, H, V, D;
  Real dx, dy, det, sid, sdd, shx, shy, bin;
};

// Per-view state handed to the consumer by a walker.
template <typename Real>
struct ViewState {
  int v;
  bool xaxis;        // true: common axis y = 0, walk rows; false: x = 0, columns
  bool flip;         // detector index runs opposite to the walk
  Real srcx, srcy;   // source position
  const Real* coef;  // FP coefficient per walk-ordered bin (fan geometries)
  Real coef_const;   // FP coefficient of parallel beam
};

// Per-line state: the line index and the pixel interval on the axis.
template <typename Real>
struct LineState {
  int line;
  Real pint;
  Real px, py;  // centre of the current pixel, tracked as the reference does
};

// Axis and detector-order choice of a view.  The reference decides with the
// FP32 cosine squared in FP32 (`cosVal * cosVal > 0.5`, FanArcDisCPU2d.cpp:76;
// flips at :78-84, :148-154).  It is a discrete part of the algorithm, not a
// rounding: views on the 45-degree diagonals take the axis FP32 picks, so the
// FP64 instantiation decides exactly as the FP32 one does.
template <typename Real>
void choose_axis(Real view, ViewState<Real>& vs) {
  const float cf = static_cast<float>(std::cos(static_cast<double>(view)));
  const float sf = static_cast<float>(std::sin(static_cast<double>(view)));
  vs.xaxis = (cf * cf > 0.5);
  vs.flip = vs.xaxis ? !(cf >= 0) : !(sf >= 0);
}

// --- small geometry helpers (FanArcDisCPU2d.cpp:11-33) ----------------------
template <typename Real>
Real map_to_x(Real sx, Real sy, Real px, Real py) {
  return (sx * py - sy * px) / (py - sy);
}
template <typename Real>
Real map_to_y(Real sx, Real sy, Real px, Real py) {
  return (sy * px - sx * py) / (px - sx);
}
template <typename Real>
Real cos_to_normal(Real a, Real b) {
  return static_cast<Real>(std::fabs(static_cast<double>(b)) /
                           std::sqrt(static_cast<double>(a * a + b * b)));
}

// --- the consumer: turns segments into the six ops --------------------------
template <typename Real>
struct Consumer {
  int op;
  int geom;
  const Params<Real>* p;
  const Real* in;  // input array (image [B,H,W] or sinogram [B,V,D])
  Real* out;       // output array
  // system-matrix mode
  std::vector<int64_t>* si = nullptr;
  std::vector<int64_t>* sj = nullptr;
  std::vector<float>* sv = nullptr;

  Real bp_coef = 0;  // BP coefficient of the current line

  void begin_line(const ViewState<Real>& vs, const LineState<Real>& ls) {
    // BackProj weights: arc SID*dgamma / PixInterval (FanArcDisCPU2d.cpp:451,480),
    // flat VirDet / PixInterval (FanFlatDisCPU2d.cpp:478),
    // para det / PixInterval (ParaDisCPU2d.cpp:202).
    if (geom == kFanArc)
      bp_coef = (p->sid * p->det) / ls.pint;
    else if (geom == kFanFlat)
      bp_coef = (p->sid / p->sdd * p->det) / ls.pint;
    else
      bp_coef = p->det / ls.pint;
    (void)vs;
  }

  Real fbp_weight(const ViewState<Real>& vs, Real px, Real py) const {
    if (geom == kFanArc) {  // FanArcDisCPU2d.cpp:27-29, used as a divisor
      return (vs.srcx - px) * (vs.srcx - px) + (vs.srcy - py) * (vs.srcy - py);
    }
    // FanFlatDisCPU2d.cpp:19-22, used as a factor
    Real r = p->sid;
    Real d = (vs.srcx * px + vs.srcy * py) / r;
    return r * r / ((r - d) * (r - d));
  }

  void segment(const ViewState<Real>& vs, const LineState<Real>& ls, int pix,
               int k, Real overlap) {
    const Params<Real>& g = *p;
    int row, col;
    if (vs.xaxis) {
      row = ls.line;
      col = pix;
    } else {
      row = g.H - 1 - pix;
      col = ls.line;
    }
    int kk = vs.flip ? (g.D - 1 - k) : k;
    const bool fp_weight = (op == kProj || op == kProjTrans);
    Real fp_coef = (geom == kPara) ? vs.coef_const : vs.coef[k];
    Real coef = overlap * (fp_weight ? fp_coef : bp_coef);
    if (si) {
      si->push_back(static_cast<int64_t>(vs.v) * g.D + kk);
      sj->push_back(static_cast<int64_t>(row) * g.W + col);
      sv->push_back(static_cast<float>(coef));
      return;
    }
    for (int b = 0; b < g.B; ++b) {
      size_t img = (static_cast<size_t>(b) * g.H + row) * g.W + col;
      size_t sino = (static_cast<size_t>(b) * g.V + vs.v) * g.D + kk;
      switch (op) {
        case kProj:
          out[sino] += in[img] * coef;
          break;
        case kBackProjTrans:
          out[sino] += in[img] * coef;
          break;
        case kBackProjTransW: {
          Real t = in[img] * coef;
          if (geom == kFanArc)
            t /= fbp_weight(vs, ls.px, ls.py);
          else
            t *= fbp_weight(vs, ls.px, ls.py);
          out[sino] += t;
          break;
        }
        case kProjTrans:
        case kBackProj:
          out[img] += in[sino] * coef;
          break;
        case kBackProjW: {
          Real t = in[sino] * coef;
          if (geom == kFanArc)
            t /= fbp_weight(vs, ls.px, ls.py);
          else
            t *= fbp_weight(vs, ls.px, ls.py);
          out[img] += t;
          break;
        }
      }
    }
  }
};

// The merge of the uniform pixel edges of one line with the bin boundaries.
// `bound(k)` returns the right edge of walk-ordered bin k.  Segments are
// emitted until either sequence is exhausted (FanArcDisCPU2d.cpp:118-144).
template <typename Real, typename C, typename BoundFn, typename StepFn>
void merge_line(C& cons, const ViewState<Real>& vs, LineState<Real>& ls,
                int npix, int pix, int k, Real b0, Real a0, Real pix_step,
                BoundFn bound, StepFn next_bound) {
  const int D = cons.p->D;
  Real pnext = static_cast<Real>(pix + 1) * ls.pint + a0;
  Real dnext = 0;
  if (k < D) dnext = bound(k);
  while (pix < npix && k < D) {
    if (pnext < dnext) {
      if (pix >= 0) cons.segment(vs, ls, pix, k, pnext - b0);
      b0 = pnext;
      ++pix;
      if (vs.xaxis)
        ls.px += pix_step;
      else
        ls.py += pix_step;
      pnext += ls.pint;
    } else {
      if (pix >= 0) cons.segment(vs, ls, pix, k, dnext - b0);
      b0 = dnext;
      ++k;
      if (k < D) dnext = next_bound(k, dnext);
    }
  }
}

// ---- fan beam (arc and flat detector) --------------------------------------
template <typename Real, typename C>
void walk_fan(C& cons, const Real* angles, bool arc) {
  const Params<Real>& g = *cons.p;
  const int D = g.D;
  std::vector<Real> bnd(D + 1), coef(D);
  for (int v = 0; v < g.V; ++v) {
    const Real view = angles[v];
    const Real s = static_cast<Real>(std::sin(static_cast<double>(view)));
    const Real c = static_cast<Real>(std::cos(static_cast<double>(view)));
    ViewState<Real> vs;
    vs.v = v;
    vs.srcx = -s * g.sid;
    vs.srcy = c * g.sid;
    choose_axis(view, vs);
    const int flag = vs.flip ? -1 : 1;
    const Real vdet = g.sid / g.sdd * g.det;  // flat: virtual detector pitch
    const Real vbin = g.sid / g.sdd * g.bin;
    Real det0 = 0;  // flat: first detector point along the walk
    // Bin boundaries on the common axis (FanArcDisCPU2d.cpp:77-86,
    // FanFlatDisCPU2d.cpp:72-84).
    for (int k = 0; k <= D; ++k) {
      double off = vs.flip ? (D / 2.0 - k) : (k - D / 2.0);
      if (arc) {
        Real ang = static_cast<Real>(off * g.det + g.bin + view);
        double t = std::tan(static_cast<double>(ang));
        bnd[k] = vs.xaxis ? static_cast<Real>(vs.srcx + vs.srcy * t)
                          : static_cast<Real>(vs.srcy + vs.srcx / t);
      } else {
        Real ex = static_cast<Real>((off * vdet + vbin) * c);
        Real ey = static_cast<Real>((off * vdet + vbin) * s);
        if (k == 0) det0 = vs.xaxis ? ex : ey;
        bnd[k] = vs.xaxis ? map_to_x(vs.srcx, vs.srcy, ex, ey)
                          : map_to_y(vs.srcx, vs.srcy, ex, ey);
      }
    }
    // FP coefficient: perpendicular pixel size / (bin width on the axis *
    // cosine of the bin-centre ray to the axis normal) (FanArcDisCPU2d.cpp:87-93).
    for (int k = 0; k < D; ++k) {
      Real mid = static_cast<Real>((bnd[k + 1] + bnd[k]) / 2.0 -
                                   (vs.xaxis ? vs.srcx : vs.srcy));
      Real cs = vs.xaxis ? cos_to_normal(mid, vs.srcy)
                         : cos_to_normal(mid, vs.srcx);
      coef[k] = (vs.xaxis ? g.dy : g.dx) / ((bnd[k + 1] - bnd[k]) * cs);
    }
    vs.coef = coef.data();
    vs.coef_const = 0;
    auto bound = [&](int k) { return bnd[k + 1]; };
    auto next_bound = [&](int k, Real) { return bnd[k + 1]; };

    const int nline = vs.xaxis ? g.H : g.W;
    const int npix = vs.xaxis ? g.W : g.H;
    const Real e0 = vs.xaxis
                        ? static_cast<Real>(-g.W / 2.0 * g.dx + g.shx)
                        : static_cast<Real>(-g.H / 2.0 * g.dy + g.shy);
    const Real e1 = vs.xaxis
                        ? static_cast<Real>(g.W / 2.0 * g.dx + g.shx)
                        : static_cast<Real>(g.H / 2.0 * g.dy + g.shy);
    const Real det_interval = vs.xaxis
        ? static_cast<Real>(vdet * std::fabs(static_cast<double>(c)))
        : static_cast<Real>(vdet * std::fabs(static_cast<double>(s)));
    for (int line = 0; line < nline; ++line) {
      LineState<Real> ls;
      ls.line = line;
      Real a0, a1, lc;  // pixel edges of the line on the axis; line centre
      if (vs.xaxis) {
        lc = static_cast<Real>((g.H / 2.0 - line - 0.5) * g.dy + g.shy);
        a0 = map_to_x(vs.srcx, vs.srcy, e0, lc);
        a1 = map_to_x(vs.srcx, vs.srcy, e1, lc);
      } else {
        lc = static_cast<Real>((line - g.W / 2.0 + 0.5) * g.dx + g.shx);
        a0 = map_to_y(vs.srcx, vs.srcy, lc, e0);
        a1 = map_to_y(vs.srcx, vs.srcy, lc, e1);
      }
      ls.pint = (a1 - a0) / static_cast<Real>(npix);
      int k, pix;
      Real b0;
      if (arc) {
        // Start bin from the angle between the ray through the first pixel
        // edge and the central ray (FanArcDisCPU2d.cpp:102-105, :172-175).
        Real tan0, tan1;
        if (vs.xaxis) {
          tan0 = ((vs.srcx - a0) == 0)
                     ? static_cast<Real>(1e10)
                     : static_cast<Real>(vs.srcy / (vs.srcx - a0));
          Real sb = g.bin + view;
          tan1 = (sb == 0) ? static_cast<Real>(1e10)
                           : static_cast<Real>(-1 / std::tan(static_cast<double>(sb)));
        } else {
          tan0 = (vs.srcy - a0) / vs.srcx;
          Real sb = g.bin + view;
          tan1 = static_cast<Real>(-1 / std::tan(static_cast<double>(sb)));
        }
        Real q = (tan0 - tan1) / (1 + tan0 * tan1);
        Real delta = static_cast<Real>(std::atan(static_cast<double>(q)));
        k = static_cast<int>(std::floor(D / 2.0 + delta * flag / g.det));
        if (k < 0) {
          b0 = bnd[0];
          pix = static_cast<int>(std::floor((b0 - a0) / ls.pint));
          k = 0;
        } else {
          b0 = a0;
          pix = 0;
        }
      } else {
        // Flat: start where the first pixel edge or the first bin boundary
        // lies, mapping the pixel edge back onto the detector
        // (FanFlatDisCPU2d.cpp:100-110, :170-179).
        b0 = static_cast<Real>(std::fmax(static_cast<double>(a0),
                                         static_cast<double>(bnd[0])));
        if (a0 == b0) {
          Real back;
          if (vs.xaxis)
            back = vs.srcy * b0 / ((b0 - vs.srcx) * s / c + vs.srcy);
          else
            back = vs.srcx * b0 / ((b0 - vs.srcy) * c / s + vs.srcx);
          pix = 0;
          k = static_cast<int>(std::floor((back - det0) / det_interval));
        } else {
          pix = static_cast<int>(std::floor((b0 - a0) / ls.pint));
          k = 0;
        }
      }
      // Pixel centre as tracked by the BackProj kernels
      // (FanArcDisCPU2d.cpp:497,516 and :573,592).
      Real pix_step;
      if (vs.xaxis) {
        ls.py = lc;
        ls.px = static_cast<Real>((pix + 0.5) * g.dx + e0);
        pix_step = g.dx;
      } else {
        ls.px = lc;
        ls.py = static_cast<Real>((pix + 0.5) * g.dy + e0);
        pix_step = g.dy;
      }
      cons.begin_line(vs, ls);
      merge_line(cons, vs, ls, npix, pix, k, b0, a0, pix_step, bound,
                 next_bound);
    }
  }
}

// ---- parallel beam (ParaDisCPU2d.cpp:45-158) -------------------------------
template <typename Real, typename C>
void walk_para(C& cons, const Real* angles) {
  const Params<Real>& g = *cons.p;
  const int D = g.D;
  for (int v = 0; v < g.V; ++v) {
    const Real view = angles[v];
    const Real s = static_cast<Real>(std::sin(static_cast<double>(view)));
    const Real c = static_cast<Real>(std::cos(static_cast<double>(view)));
    ViewState<Real> vs;
    vs.v = v;
    vs.srcx = 0;
    vs.srcy = 0;
    vs.coef = nullptr;
    choose_axis(view, vs);
    const Real axis_cs = vs.xaxis ? c : s;
    // First bin boundary on the axis and the uniform bin pitch there.
    const Real det0 = static_cast<Real>(
        ((vs.flip ? (D / 2.0) : (-D / 2.0)) * g.det + g.bin) / axis_cs);
    const Real det_interval = static_cast<Real>(
        g.det / std::fabs(static_cast<double>(axis_cs)));
    vs.coef_const = (vs.xaxis ? g.dy : g.dx) / g.det;
    const Real slope = vs.xaxis ? s / c : c / s;
    const int nline = vs.xaxis ? g.H : g.W;
    const int npix = vs.xaxis ? g.W : g.H;
    const Real e0 = vs.xaxis
                        ? static_cast<Real>(-g.W / 2.0 * g.dx + g.shx)
                        : static_cast<Real>(-g.H / 2.0 * g.dy + g.shy);
    auto bound = [&](int k) {
      return static_cast<Real>(k + 1) * det_interval + det0;
    };
    auto next_bound = [&](int, Real prev) { return prev + det_interval; };
    for (int line = 0; line < nline; ++line) {
      LineState<Real> ls;
      ls.line = line;
      ls.pint = vs.xaxis ? g.dx : g.dy;
      Real lc = vs.xaxis
                    ? static_cast<Real>((g.H / 2.0 - line - 0.5) * g.dy + g.shy)
                    : static_cast<Real>((line - g.W / 2.0 + 0.5) * g.dx + g.shx);
      Real a0 = slope * lc + e0;
      Real b0 = static_cast<Real>(std::fmax(static_cast<double>(a0),
                                            static_cast<double>(det0)));
      int pix, k;
      if (a0 == b0) {
        pix = 0;
        k = static_cast<int>(std::floor((b0 - det0) / det_interval));
      } else {
        pix = static_cast<int>(std::floor((b0 - a0) / ls.pint));
        k = 0;
      }
      ls.px = ls.py = 0;
      cons.begin_line(vs, ls);
      merge_line(cons, vs, ls, npix, pix, k, b0, a0, static_cast<Real>(0),
                 bound, next_bound);
    }
  }
}

template <typename Real>
int run_op(int geom, int op, const Real* in, Real* out, const Real* angles,
           const Params<Real>& prm) {
  if (geom < 0 || geom > 2 || op < 0 || op > 5) return -1;
  if (geom == kPara && op >= kBackProjW) return -1;
  Consumer<Real> cons;
  cons.op = op;
  cons.geom = geom;
  cons.p = &prm;
  cons.in = in;
  cons.out = out;
  if (geom == kPara)
    walk_para(cons, angles);
  else
    walk_fan(cons, angles, geom == kFanArc);
  return 0;
}

std::vector<int64_t> g_i, g_j;
std::vector<float> g_v;

}  // namespace

extern "C" {

// One op in the FP32-mirror evaluation.  `image` is [B,H,W], `proj` is
// [B,V,D]; the output side is accumulated into, so pass it zero-filled.
// op: 0 Proj, 1 ProjTrans, 2 BackProj, 3 BackProjTrans, 4 BackProjWeighted,
// 5 BackProjTransWeighted.  geom: 0 para, 1 fan flat, 2 fan arc.
int ddo_op2d_f32(int geom, int op, float* image, float* proj,
                 const float* angles, int batch, int width, int height,
                 int views, int dets, float dx, float dy, float det, float sid,
                 float sdd, float shx, float shy, float bin) {
  Params<float> prm{batch, width, height, views, dets, dx,  dy,
                    det,   sid,   sdd,    shx,   shy,  bin};
  const bool to_sino =
      (op == kProj || op == kBackProjTrans || op == kBackProjTransW);
  return to_sino ? run_op<float>(geom, op, image, proj, angles, prm)
                 : run_op<float>(geom, op, proj, image, angles, prm);
}

// The same op with every quantity in double: the FP64 oracle.
int ddo_op2d_f64(int geom, int op, double* image, double* proj,
                 const double* angles, int batch, int width, int height,
                 int views, int dets, double dx, double dy, double det,
                 double sid, double sdd, double shx, double shy, double bin) {
  Params<double> prm{batch, width, height, views, dets, dx,  dy,
                     det,   sid,   sdd,    shx,   shy,  bin};
  const bool to_sino =
      (op == kProj || op == kBackProjTrans || op == kBackProjTransW);
  return to_sino ? run_op<double>(geom, op, image, proj, angles, prm)
                 : run_op<double>(geom, op, proj, image, angles, prm);
}

// COO system matrix with the FP weights (FanArcDisCPU2d.cpp:1057-1225):
// row = v*D + detector, column = row*W + col.  Kept until the next call.
int64_t ddo_sysmatrix2d(int geom, const float* angles, int width, int height,
                        int views, int dets, float dx, float dy, float det,
                        float sid, float sdd, float shx, float shy,
                        float bin) {
  if (geom < 0 || geom > 2) return -1;
  Params<float> prm{1, width, height, views, dets, dx,  dy,
                    det, sid,   sdd,    shx,   shy,  bin};
  g_i.clear();
  g_j.clear();
  g_v.clear();
  Consumer<float> cons;
  cons.op = kProj;
  cons.geom = geom;
  cons.p = &prm;
  cons.in = nullptr;
  cons.out = nullptr;
  cons.si = &g_i;
  cons.sj = &g_j;
  cons.sv = &g_v;
  if (geom == kPara)
    walk_para(cons, angles);
  else
    walk_fan(cons, angles, geom == kFanArc);
  return static_cast<int64_t>(g_v.size());
}

void ddo_sysmatrix_copy(int64_t* rows, int64_t* cols, float* vals) {
  std::memcpy(rows, g_i.data(), g_i.size() * sizeof(int64_t));
  std::memcpy(cols, g_j.data(), g_j.size() * sizeof(int64_t));
  std::memcpy(vals, g_v.data(), g_v.size() * sizeof(float));
}

}  // extern "C"
