// FP64 restatement of deeprecon's 3D distance-driven projectors.
//
// TEST INFRASTRUCTURE ONLY (see dd_oracle.cpp).  The reference has no CPU
// path for 3D (deeprecon/interface/RadonTorch.cpp:604 raises "CPU is not
// supported yet"), so this file restates its CUDA kernels in double:
//   cone-flat deeprecon/cuda/dis3d/ConeFlatDisCUDA3d.cu:27-281 (Proj),
//             :284-536 (ProjTrans), :539-794 (BackProj<W>), :797-1058;
//   cone-arc  deeprecon/cuda/dis3d/ConeArcDisCUDA3d.cu:34-1055;
//   spiral    deeprecon/cuda/dis3d/SpiralDisCUDA3d.cu:34-1092 (per-view source
//             axial position, radius shift, angle shift, axial shift :72-83).
// Separable footprint: per view and image line, the column walk on the 2D
// common axis (as in 2D) times, per column, a z walk on the axis plane, where
// detector rows are magnified at the column centre (:92-93) and voxel slices
// at the line centre (:120-121).
//
// Pinned by tests/test_oracle_3d.py: reductions to the pinned 2D oracle
// (one detector row and one slice that covers it), exact adjointness of each
// op pair, spiral with zero source tables equal to cone-arc with a centred
// volume, and one exact identity per free shift parameter, each shown to fail
// on a planted sign slip (ddo_plant3d):
//   PixZShift = dz           <=> the volume moved down one slice (:103-104);
//   BinRowShift = row pitch  <=> detector rows moved by one (:68);
//   SourceShiftRadius = c    <=> cone-arc with SID+c, SDD+c, column pitch and
//                                shift x SDD/(SDD+c) (SpiralDisCUDA3d.cu:72-75);
//   SourceShiftAngle = d     <=> cone-arc at view angle + d with column shift
//                                - d (SDD-SID)/SDD - d (:75-78);
//   SourceShiftAxial = s     <=> axial position + s with BinRowShift - s (:81-82).
// The CUDA kernels' own FP32 behaviour is not reproduced (parity unpinned at
// the bit level; the GPU is held to this FP64 oracle).
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

enum { kConeFlat = 3, kConeArc = 4, kSpiral = 5 };
enum { kProj = 0, kProjTrans = 1, kBackProj = 2, kBackProjTrans = 3,
       kBackProjW = 4, kBackProjTransW = 5 };

// Planted mistakes for the pinning tests (tests/test_oracle_3d.py): each bit
// flips one sign or drops one factor, so a test can show that its identity
// fails on that mistake.  Zero in every normal use.
enum {
  kPlantPixZShift = 1,      // z0 - shz instead of z0 + shz
  kPlantBinRowShift = 2,    // row edges - rsh instead of + rsh
  kPlantRadius = 4,         // SID - dr, SDD - dr
  kPlantAngleSource = 8,    // source rotated by -da
  kPlantAngleColumns = 16,  // column shift + da (SDD'-SID')/SDD'
  kPlantAxial = 32,         // source z = axial - shift
  kPlantColPitch = 64,      // column pitch not rescaled by SDD/SDD'
};
int g_plant = 0;

struct P3 {
  int scan, B, W, H, Dz, C, R, V;
  double dx, dy, dz, col, row, sid, sdd, shx, shy, shz, csh, rsh;
  const double* spos;  // [4][V] for spiral, else null
};

void run3d(const P3& g, int op, const double* angles, const double* in,
           double* out) {
  const bool to_sino = (op == kProj || op == kBackProjTrans || op == kBackProjTransW);
  const int C = g.C, R = g.R;
  std::vector<double> ce(C + 1), mag(C), zdet(R + 1);
  for (int v = 0; v < g.V; ++v) {
    // ---- per-view source and detector parameters
    double sid = g.sid, sdd = g.sdd, colp = g.col, csh = g.csh, sz = 0.0, zoff = 0.0;
    float th_src_f = static_cast<float>(angles[v]);
    const double th_det = angles[v];
    if (g.scan == kSpiral) {
      const double pa = g.spos[v], dr = g.spos[g.V + v], da = g.spos[2 * g.V + v],
                   dzs = g.spos[3 * g.V + v];
      const double rs = (g_plant & kPlantRadius) ? -dr : dr;
      sid = g.sid + rs;
      sdd = g.sdd + rs;
      colp = (g_plant & kPlantColPitch) ? g.col : g.col * g.sdd / sdd;
      csh = g.csh * g.sdd / sdd +
            ((g_plant & kPlantAngleColumns) ? da : -da) * (sdd - sid) / sdd;
      th_src_f = static_cast<float>(angles[v]) +
                 static_cast<float>((g_plant & kPlantAngleSource) ? -da : da);
      sz = pa + ((g_plant & kPlantAxial) ? -dzs : dzs);
      zoff = pa;
    }
    const double th_src = static_cast<double>(th_src_f);
    // axis choice in FP32 as the reference takes it
    const float cf = static_cast<float>(std::cos(th_src));
    const float sf = static_cast<float>(std::sin(th_src));
    const bool xaxis = cf * cf > 0.5;
    const bool flip = xaxis ? !(cf >= 0) : !(sf >= 0);
    const double s = std::sin(th_src), c = std::cos(th_src);
    const double sx = -s * sid, sy = c * sid;
    const double vcol = g.sid / g.sdd * g.col, vcsh = g.sid / g.sdd * g.csh;
    // column edges on the axis, walk order
    for (int k = 0; k <= C; ++k) {
      const double off = flip ? (C / 2.0 - k) : (k - C / 2.0);
      if (g.scan == kConeFlat) {
        const double u = off * vcol + vcsh, px = u * c, py = u * s;
        ce[k] = xaxis ? (sx * py - sy * px) / (py - sy) : (sy * px - sx * py) / (px - sx);
      } else {
        const double t = std::tan(off * colp + csh + th_det);
        ce[k] = xaxis ? sx + sy * t : sy + sx / t;
      }
    }
    // magnification of detector rows at each column centre
    for (int k = 0; k < C; ++k) {
      const double offc = flip ? (C / 2.0 - k - 0.5) : (k + 0.5 - C / 2.0);
      double dcx, dcy;
      if (g.scan == kConeFlat) {
        const double u = offc * vcol + vcsh;
        dcx = u * c;
        dcy = u * s;
      } else {
        const double gc = offc * colp + csh + th_det;
        dcx = std::sin(gc) * sdd + sx;
        dcy = -std::cos(gc) * sdd + sy;
      }
      mag[k] = xaxis ? sy / (sy - dcy) : sx / (sx - dcx);
    }
    // detector row edges (z on the detector, or on the virtual detector)
    for (int r = 0; r <= R; ++r) {
      const double z = (r - R / 2.0) * g.row + ((g_plant & kPlantBinRowShift) ? -g.rsh : g.rsh);
      zdet[r] = (g.scan == kConeFlat) ? z * g.sid / g.sdd : z + zoff;
    }
    auto zrow = [&](int k, int r) { return mag[k] * (zdet[r] - sz) + sz; };
    const double kbp = (g.scan == kConeFlat) ? vcol : sid * colp;
    const double perp = xaxis ? g.dy : g.dx;
    const double s_ax = xaxis ? sx : sy, s_pp = xaxis ? sy : sx;
    auto wfp = [&](int k, int r) {
      const double z0 = zrow(k, r), z1 = zrow(k, r + 1);
      const double a = 0.5 * (ce[k] + ce[k + 1]) - s_ax, b = 0.5 * (z0 + z1);
      const double cs = std::fabs(s_pp) / std::sqrt(a * a + b * b + s_pp * s_pp);
      return perp / ((ce[k + 1] - ce[k]) * (z1 - z0) * cs);
    };
    const int nline = xaxis ? g.H : g.W;
    const int npix = xaxis ? g.W : g.H;
    const double shz = (g_plant & kPlantPixZShift) ? -g.shz : g.shz;
    const double z0v = -g.Dz / 2.0 * g.dz + (g.scan == kSpiral ? 0.0 : shz);
    for (int l = 0; l < nline; ++l) {
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
      const double zl0 = m * (z0v - sz) + sz, pz = m * g.dz;
      // column merge: pixels [a0 + c p, a0 + (c+1) p] against bins [ce_k, ce_k+1]
      for (int cpx = 0; cpx < npix; ++cpx) {
        const double pa = a0 + cpx * p, pb = pa + p;
        for (int k = 0; k < C; ++k) {
          const double ox = std::fmin(pb, ce[k + 1]) - std::fmax(pa, ce[k]);
          if (ox <= 0) continue;
          int row_img, col_img;
          double xc, yc;
          if (xaxis) {
            row_img = l;
            col_img = cpx;
            xc = (cpx - g.W / 2.0 + 0.5) * g.dx + g.shx;
            yc = (g.H / 2.0 - l - 0.5) * g.dy + g.shy;
          } else {
            row_img = g.H - 1 - cpx;
            col_img = l;
            xc = (l - g.W / 2.0 + 0.5) * g.dx + g.shx;
            yc = (cpx - g.H / 2.0 + 0.5) * g.dy + g.shy;
          }
          double fbp = 1.0;
          if (op == kBackProjW || op == kBackProjTransW) {
            if (g.scan == kConeFlat) {
              const double d = (sx * xc + sy * yc) / g.sid;
              fbp = g.sid * g.sid / ((g.sid - d) * (g.sid - d));
            } else {
              fbp = 1.0 / ((sx - xc) * (sx - xc) + (sy - yc) * (sy - yc));
            }
          }
          const int kk = flip ? C - 1 - k : k;
          // z merge: slices [zl0 + s pz, +pz] against rows [zrow(k,r), zrow(k,r+1)],
          // both increasing (the reference's nested walk, ConeFlatDisCUDA3d.cu:131-166)
          int sl = 0, r = 0;
          while (sl < g.Dz && r < R) {
            const double za = zl0 + sl * pz, zb = za + pz;
            const double ra = zrow(k, r), rb = zrow(k, r + 1);
            const double oz = std::fmin(zb, rb) - std::fmax(za, ra);
            const int s_here = sl, r_here = r;
            if (zb < rb) ++sl; else ++r;
            if (oz <= 0) continue;
            {
              const int sl = s_here, r = r_here;
              const double wfp_v = (op == kProj || op == kProjTrans) ? wfp(k, r) * ox * oz : 0.0;
              const double wbp_v = kbp * (ox / p) * (oz / pz) * fbp;
              const double w = (op == kProj || op == kProjTrans) ? wfp_v : wbp_v;
              for (int b = 0; b < g.B; ++b) {
                const size_t img = ((static_cast<size_t>(b) * g.Dz + sl) * g.H + row_img) * g.W + col_img;
                const size_t sino = ((static_cast<size_t>(b) * g.V + v) * R + r) * C + kk;
                if (to_sino)
                  out[sino] += w * in[img];
                else
                  out[img] += w * in[sino];
              }
            }
          }
        }
      }
    }
  }
}

}  // namespace

extern "C" {

// Sets the planted-mistake mask (test infrastructure; 0 restores the oracle).
void ddo_plant3d(int mask) { g_plant = mask; }

// 3D op in FP64.  scan: 3 cone-flat, 4 cone-arc, 5 spiral.  `image` is
// [B, Dz, H, W], `proj` [B, V, rows, cols]; the output side is accumulated
// into (pass it zero-filled).  `spos` is [4, V] for the spiral, else null.
// Geometry doubles: dx, dy, dz, col, row, sid, sdd, shx, shy, shz, col_shift,
// row_shift (the reference's Options order).
int ddo_op3d_f64(int scan, int op, double* image, double* proj,
                 const double* angles, const double* spos, int batch,
                 int width, int height, int depth, int views, int cols,
                 int rows, double dx, double dy, double dz, double col,
                 double row, double sid, double sdd, double shx, double shy,
                 double shz, double csh, double rsh) {
  if (scan < kConeFlat || scan > kSpiral || op < 0 || op > 5) return -1;
  if (scan == kSpiral && !spos) return -1;
  P3 g{scan, batch, width, height, depth, cols, rows, views, dx, dy, dz, col, row,
       sid, sdd, shx, shy, shz, csh, rsh, spos};
  const bool to_sino = (op == kProj || op == kBackProjTrans || op == kBackProjTransW);
  run3d(g, op, angles, to_sino ? image : proj, to_sino ? proj : image);
  return 0;
}

}  // extern "C"
