// `radontorch`: the torch-op boundary of the projector, MI355X build.
//
// Same module name, op names, arities and tensor layouts as the reference's
// deeprecon/interface/RadonTorch.cpp (ops :17-1372, registration :1375-1418).
// Each op decodes the Options tensor, allocates the output and a workspace
// through torch's caching allocator, and calls the C ABI of libdd (include/dd.h)
// on torch's current HIP stream, so every op is stream-ordered and can be
// captured into a HIP graph.
//
// Differences from the reference, all deliberate:
//  * inputs must live on a HIP device; the reference's CPU kernels are not
//    part of this build and CPU tensors raise an error (never a CPU fallback);
//  * `options` may live on the CPU (the Python layer passes a host mirror and
//    the op reads it without a device sync).  A GPU-resident options tensor is
//    read once per tensor version (outside capture) and remembered, so later
//    calls and graph captures run without host reads (the reference issues
//    11-17 `.item()` syncs per call, RadonTorch.cpp:227-237);
//  * input shapes are checked against the geometry instead of being read out
//    of bounds, and every HIP error is reported.
#include <torch/extension.h>

#include <ATen/hip/HIPContext.h>
#include <c10/hip/HIPGraphsC10Utils.h>
#include <c10/hip/HIPGuard.h>
#include <c10/hip/HIPStream.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "dd.h"
#include "dd3d.h"

namespace {

const char* status_name(int rc) {
  switch (rc) {
    case DD_OK: return "ok";
    case DD_BAD_ARG: return "bad argument";
    case DD_UNSUPPORTED: return "unsupported geometry/op pair";
    case DD_LAUNCH_FAIL: return "HIP launch failure";
    case DD_NO_WORKSPACE: return "workspace too small";
    case DD_NO_DEVICE: return "no HIP device";
  }
  return "unknown error";
}

#define DD_CHECK(expr, what)                                              \
  do {                                                                    \
    int rc__ = (expr);                                                    \
    TORCH_CHECK(rc__ == DD_OK, what, " failed: ", status_name(rc__));     \
  } while (0)

// ---- options decoding ------------------------------------------------------
// Device-resident options are read to the host once per (tensor, version,
// storage) and remembered.  Each entry holds a reference to its tensor, so
// an entry's key (the TensorImpl address) cannot be reused by another tensor
// while it is cached; the cache is cleared when it reaches kOptCacheMax.
struct CachedOptions {
  torch::Tensor ref;
  int64_t version;
  const void* data;
  std::vector<float> values;
};
constexpr size_t kOptCacheMax = 256;
std::mutex g_opt_mu;
// never destroyed: its tensors must not be freed after the HIP runtime at exit
auto& g_opt_cache = *new std::unordered_map<const c10::TensorImpl*, CachedOptions>();

bool capturing() {
  return c10::hip::currentStreamCaptureStatusMayInitCtx() !=
         c10::hip::CaptureStatus::None;
}

std::vector<float> read_options(const torch::Tensor& options, int64_t n,
                                const char* op) {
  TORCH_CHECK(options.dim() == 1 && options.numel() >= n, op,
              ": options must be a 1D tensor with ", n, " entries");
  if (options.device().is_cpu()) {
    auto o = options.to(torch::kFloat32).contiguous();
    const float* p = o.data_ptr<float>();
    return std::vector<float>(p, p + n);
  }
  const c10::TensorImpl* key = options.unsafeGetTensorImpl();
  const int64_t version = options._version();
  const void* data = options.data_ptr();
  std::lock_guard<std::mutex> lock(g_opt_mu);
  auto it = g_opt_cache.find(key);
  if (it != g_opt_cache.end() && it->second.version == version && it->second.data == data &&
      static_cast<int64_t>(it->second.values.size()) == n)
    return it->second.values;
  TORCH_CHECK(!capturing(), op,
              ": a device-resident options tensor must be used once outside stream "
              "capture before it is captured (or pass the options on the CPU)");
  auto o = options.to(torch::kCPU, torch::kFloat32).contiguous();
  const float* p = o.data_ptr<float>();
  if (g_opt_cache.size() >= kOptCacheMax) g_opt_cache.clear();
  CachedOptions c{options, version, data, std::vector<float>(p, p + n)};
  auto& slot = g_opt_cache[key];
  slot = std::move(c);
  return slot.values;
}

dd_geom2d decode2d(const torch::Tensor& options, const char* op) {
  auto o = read_options(options, 11, op);
  dd_geom2d g;
  g.width = static_cast<int>(o[0]);
  g.height = static_cast<int>(o[1]);
  g.num_det = static_cast<int>(o[2]);
  g.pix_x = o[3];
  g.pix_y = o[4];
  g.det = o[5];
  g.iso_source = o[6];
  g.source_detector = o[7];
  g.shift_x = o[8];
  g.shift_y = o[9];
  g.bin_shift = o[10];
  return g;
}

dd_geom3d decode3d(const torch::Tensor& options, bool spiral, const char* op) {
  const int n = spiral ? 16 : 17;
  auto o = read_options(options, n, op);
  dd_geom3d g;
  g.width = static_cast<int>(o[0]);
  g.height = static_cast<int>(o[1]);
  g.depth = static_cast<int>(o[2]);
  g.num_col = static_cast<int>(o[3]);
  g.num_row = static_cast<int>(o[4]);
  g.pix_x = o[5];
  g.pix_y = o[6];
  g.pix_z = o[7];
  g.col_size = o[8];
  g.row_size = o[9];
  g.iso_source = o[10];
  g.source_detector = o[11];
  g.shift_x = o[12];
  g.shift_y = o[13];
  if (spiral) {
    g.shift_z = 0.0f;
    g.col_shift = o[14];
    g.row_shift = o[15];
  } else {
    g.shift_z = o[14];
    g.col_shift = o[15];
    g.row_shift = o[16];
  }
  return g;
}

void check_data(const torch::Tensor& t, const torch::Tensor& angles,
                const char* op) {
  TORCH_CHECK(t.is_cuda(), op, ": this build runs on a HIP device only; got a ",
              t.device().str(), " tensor");
  TORCH_CHECK(t.scalar_type() == torch::kFloat32, op,
              ": expected a float32 input, got ", t.scalar_type());
  TORCH_CHECK(t.is_contiguous(), op, ": expected a contiguous input");
  TORCH_CHECK(angles.device() == t.device(), op,
              ": angles must be on the input's device");
  TORCH_CHECK(angles.scalar_type() == torch::kFloat32 && angles.dim() == 1 &&
                  angles.is_contiguous(),
              op, ": angles must be a contiguous 1D float32 tensor");
}

torch::Tensor workspace(size_t bytes, const torch::Tensor& like) {
  return torch::empty({static_cast<int64_t>(std::max<size_t>(bytes, 1))},
                      like.options().dtype(torch::kUInt8));
}

void* cur_stream(const torch::Tensor& t) {
  return static_cast<void*>(
      c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

// ---- 2D ----------------------------------------------------------------------
torch::Tensor op2d(int scan, int op, const torch::Tensor& input,
                   const torch::Tensor& options, const torch::Tensor& angles,
                   const char* name) {
  check_data(input, angles, name);
  const dd_geom2d g = decode2d(options, name);
  TORCH_CHECK(input.dim() == 4, name, ": expected a 4D [B, C, ., .] input");
  const int64_t B = input.size(0), C = input.size(1);
  const int64_t V = angles.size(0);
  const bool to_sino = (op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW);
  if (to_sino) {
    TORCH_CHECK(input.size(2) == g.height && input.size(3) == g.width, name,
                ": image is ", input.sizes(), " but the geometry says H=",
                g.height, " W=", g.width);
  } else {
    TORCH_CHECK(input.size(2) == V && input.size(3) == g.num_det, name,
                ": sinogram is ", input.sizes(), " but there are ", V,
                " angles and ", g.num_det, " detector bins");
  }
  c10::hip::HIPGuard guard(input.device().index());
  auto out = to_sino ? torch::empty({B, C, V, g.num_det}, input.options())
                     : torch::empty({B, C, g.height, g.width}, input.options());
  if (B * C == 0 || V == 0) return out.zero_();
  const size_t wsb = dd2d_workspace_size(scan, op, &g, static_cast<int>(B * C),
                                         static_cast<int>(V));
  TORCH_CHECK(wsb > 0, name, ": invalid geometry");
  auto ws = workspace(wsb, input);
  DD_CHECK(dd2d_run(scan, op, &g, input.data_ptr<float>(), out.data_ptr<float>(),
                    angles.data_ptr<float>(), static_cast<int>(B * C),
                    static_cast<int>(V), ws.data_ptr(), wsb, cur_stream(input)),
           name);
  return out;
}

// View-resolved sinogram -> image op: [B, C, V, D] -> [B, C*V, H, W], entry
// (b, c*V + v) holding view v's contribution alone.
torch::Tensor views2d(int64_t scan, int64_t op, const torch::Tensor& input,
                      const torch::Tensor& options, const torch::Tensor& angles) {
  const char* name = "views2d";
  TORCH_CHECK(scan >= DD_PARA && scan <= DD_FAN_ARC, name, ": 2D scan expected");
  TORCH_CHECK(op == DD_PROJ_T || op == DD_BPROJ || op == DD_BPROJ_W, name,
              ": op must be PROJ_T, BPROJ or BPROJ_W");
  check_data(input, angles, name);
  const dd_geom2d g = decode2d(options, name);
  TORCH_CHECK(input.dim() == 4, name, ": expected a 4D [B, C, V, D] sinogram");
  const int64_t B = input.size(0), C = input.size(1), V = angles.size(0);
  TORCH_CHECK(input.size(2) == V && input.size(3) == g.num_det, name, ": sinogram is ",
              input.sizes(), " but there are ", V, " angles and ", g.num_det,
              " detector bins");
  c10::hip::HIPGuard guard(input.device().index());
  auto out = torch::empty({B, C * V, g.height, g.width}, input.options());
  if (B * C == 0 || V == 0) return out.zero_();
  const size_t wsb = dd2d_views_workspace_size(static_cast<int>(scan), static_cast<int>(op), &g,
                                               static_cast<int>(B * C), static_cast<int>(V));
  TORCH_CHECK(wsb > 0, name, ": invalid geometry or op");
  auto ws = workspace(wsb, input);
  DD_CHECK(dd2d_run_views(static_cast<int>(scan), static_cast<int>(op), &g,
                          input.data_ptr<float>(), out.data_ptr<float>(),
                          angles.data_ptr<float>(), static_cast<int>(B * C),
                          static_cast<int>(V), ws.data_ptr(), wsb, cur_stream(input)),
           name);
  return out;
}

torch::Tensor sysmatrix2d(int scan, const torch::Tensor& options,
                          const torch::Tensor& angles, const char* name) {
  const dd_geom2d g = decode2d(options, name);
  TORCH_CHECK(angles.dim() == 1, name, ": angles must be 1D");
  TORCH_CHECK(torch::cuda::is_available(), name,
              ": building the system matrix needs a HIP device in this build");
  auto dev = torch::Device(torch::kCUDA, c10::hip::current_device());
  c10::hip::HIPGuard guard(dev.index());
  auto ang = angles.to(dev, torch::kFloat32).contiguous();
  const int V = static_cast<int>(ang.size(0));
  auto f = torch::TensorOptions().device(dev);
  const size_t wsb = dd2d_sysmatrix_workspace_size(scan, &g, V);
  auto ws = torch::empty({static_cast<int64_t>(std::max<size_t>(wsb, 1))},
                         f.dtype(torch::kUInt8));
  auto counts = torch::empty({static_cast<int64_t>(V) * g.num_det},
                             f.dtype(torch::kInt64));
  void* st = static_cast<void*>(c10::hip::getCurrentHIPStream(dev.index()).stream());
  DD_CHECK(dd2d_sysmatrix_count(scan, &g, ang.data_ptr<float>(), V,
                                counts.data_ptr<int64_t>(), ws.data_ptr(), wsb, st),
           name);
  auto offsets = torch::cumsum(counts, 0) - counts;
  const int64_t nnz = counts.sum().item<int64_t>();
  auto rows = torch::empty({nnz}, f.dtype(torch::kInt64));
  auto cols = torch::empty({nnz}, f.dtype(torch::kInt64));
  auto vals = torch::empty({nnz}, f.dtype(torch::kFloat32));
  if (nnz > 0)
    DD_CHECK(dd2d_sysmatrix_fill(scan, &g, ang.data_ptr<float>(), V,
                                 offsets.data_ptr<int64_t>(),
                                 rows.data_ptr<int64_t>(), cols.data_ptr<int64_t>(),
                                 vals.data_ptr<float>(), ws.data_ptr(), wsb, st),
             name);
  auto idx = torch::stack({rows, cols}).to(torch::kCPU);
  const int64_t M = static_cast<int64_t>(V) * g.num_det;
  const int64_t N = static_cast<int64_t>(g.width) * g.height;
  return torch::sparse_coo_tensor(
      idx, vals.to(torch::kCPU), {M, N},
      torch::TensorOptions().dtype(torch::kFloat32).layout(torch::kSparse));
}

// ---- 3D ----------------------------------------------------------------------
torch::Tensor op3d(int scan, int op, const torch::Tensor& input,
                   const torch::Tensor& options, const torch::Tensor& angles,
                   const torch::Tensor* source_pos, const char* name) {
  check_data(input, angles, name);
  const bool spiral = (scan == DD_SPIRAL);
  const dd_geom3d g = decode3d(options, spiral, name);
  TORCH_CHECK(input.dim() == 5, name, ": expected a 5D [B, C, ., ., .] input");
  const int64_t B = input.size(0), C = input.size(1);
  const int64_t V = angles.size(0);
  const float* sp = nullptr;
  torch::Tensor spc;
  if (spiral) {
    TORCH_CHECK(source_pos != nullptr, name, ": source_pos required");
    spc = *source_pos;
    TORCH_CHECK(spc.device() == input.device(), name,
                ": source_pos must be on the input's device");
    TORCH_CHECK(spc.scalar_type() == torch::kFloat32 && spc.dim() == 2 &&
                    spc.size(0) == 4 && spc.size(1) == V && spc.is_contiguous(),
                name, ": source_pos must be a contiguous float32 [4, V] tensor");
    sp = spc.data_ptr<float>();
  }
  const bool to_sino = (op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW);
  if (to_sino) {
    TORCH_CHECK(input.size(2) == g.depth && input.size(3) == g.height &&
                    input.size(4) == g.width,
                name, ": volume is ", input.sizes(), " but the geometry says D=",
                g.depth, " H=", g.height, " W=", g.width);
  } else {
    TORCH_CHECK(input.size(2) == V && input.size(3) == g.num_row &&
                    input.size(4) == g.num_col,
                name, ": sinogram is ", input.sizes(), " but the geometry says V=",
                V, " rows=", g.num_row, " cols=", g.num_col);
  }
  c10::hip::HIPGuard guard(input.device().index());
  auto out = to_sino
                 ? torch::empty({B, C, V, g.num_row, g.num_col}, input.options())
                 : torch::empty({B, C, g.depth, g.height, g.width}, input.options());
  if (B * C == 0 || V == 0) return out.zero_();
  const size_t wsb = dd3d_workspace_size(scan, op, &g, static_cast<int>(B * C),
                                         static_cast<int>(V));
  TORCH_CHECK(wsb > 0, name, ": invalid geometry");
  auto ws = workspace(wsb, input);
  DD_CHECK(dd3d_run(scan, op, &g, input.data_ptr<float>(), out.data_ptr<float>(),
                    angles.data_ptr<float>(), sp, static_cast<int>(B * C),
                    static_cast<int>(V), ws.data_ptr(), wsb, cur_stream(input)),
           name);
  return out;
}

// ---- planned 2D operator (batched path) ---------------------------------------
// Geometry tables are built once at construction; every call only moves
// data (interleave + gather), on torch's current stream, with workspaces
// owned by the plan, so a step replays inside a captured HIP graph without
// host work.  No counterpart in the reference (which rebuilds its texture
// and geometry on every call, FanArcDisCUDA2d.cu:758-814).
class Plan2d {
 public:
  Plan2d(int scan, const torch::Tensor& options, const torch::Tensor& angles,
         int64_t batch, int64_t tile, int64_t variant, int64_t fwd_waves, int64_t fwd_views,
         int64_t pix_width)
      : scan_(scan), batch_(batch), tile_(static_cast<int>(tile)),
        variant_(static_cast<int>(variant)),
        shape_{static_cast<int>(fwd_waves), static_cast<int>(fwd_views),
               static_cast<int>(pix_width)} {
    const int64_t adj = variant < 0 ? variant : (variant & ~static_cast<int64_t>(DD_FWD_PREFIX));
    TORCH_CHECK(variant == -1 || (variant >= 0 && ((adj >= DD_ADJ_PER_PIXEL && adj <= DD_ADJ_QUAD) ||
                                                   adj == DD_ADJ_BLOCK_PREFIX)),
                "Plan2d: unknown kernel variant");
    TORCH_CHECK(fwd_waves == 0 || fwd_waves == 1 || fwd_waves == 2 || fwd_waves == 4,
                "Plan2d: fwd_waves must be 1, 2 or 4 (0 = default)");
    TORCH_CHECK(fwd_views >= 0 && fwd_views <= 8, "Plan2d: fwd_views must be 1..8 (0 = default)");
    TORCH_CHECK(pix_width == 0 || pix_width == 16 || pix_width == 32 || pix_width == 64,
                "Plan2d: pix_width must be 16, 32 or 64 (0 = default)");
    TORCH_CHECK(scan >= DD_PARA && scan <= DD_FAN_ARC, "Plan2d: 2D scan expected");
    TORCH_CHECK(angles.is_cuda() && angles.scalar_type() == torch::kFloat32 &&
                    angles.dim() == 1,
                "Plan2d: angles must be a 1D float32 device tensor");
    TORCH_CHECK(batch > 0, "Plan2d: batch must be positive");
    g_ = decode2d(options, "Plan2d");
    views_ = angles.size(0);
    angles_ = angles.contiguous();
    c10::hip::HIPGuard guard(angles_.device().index());
    const size_t gb = dd2d_geometry_size(scan, &g_, static_cast<int>(views_));
    TORCH_CHECK(gb > 0, "Plan2d: invalid geometry");
    geo_ = workspace(gb, angles_);
    DD_CHECK(dd2d_geometry(scan, &g_, angles_.data_ptr<float>(),
                           static_cast<int>(views_), geo_.data_ptr(), gb,
                           cur_stream(angles_)),
             "Plan2d geometry");
    for (int op = DD_PROJ; op <= DD_BPROJ_TW; ++op) {
      if (scan == DD_PARA && op >= DD_BPROJ_W) break;
      ws_bytes_[op] = dd2d_planned_workspace_size(scan, op, &g_, static_cast<int>(batch),
                                                  static_cast<int>(views_), tile_);
      // allocated up front so no allocation happens inside a graph capture
      ws_[op] = workspace(ws_bytes_[op], angles_);
    }
  }

  // Runs op `op` (dd_op) on a [batch, C, ...] input whose B*C equals the
  // planned batch; returns a new output tensor.
  torch::Tensor run(int64_t op, const torch::Tensor& input) {
    return run_ep(op, input, c10::nullopt, c10::nullopt, 1.0, 0.0);
  }

  // run() with the fused epilogue  out = alpha * scale * op(input) + beta * c
  // (scale: an optional one-element device tensor, read when the op runs;
  // c: an optional tensor of the output's shape).
  torch::Tensor run_ep(int64_t op, const torch::Tensor& input,
                       const c10::optional<torch::Tensor>& c,
                       const c10::optional<torch::Tensor>& scale, double alpha,
                       double beta) {
    TORCH_CHECK(op >= DD_PROJ && op <= DD_BPROJ_TW && ws_bytes_[op] > 0,
                "Plan2d: op not available for this scan");
    check_data(input, angles_, "Plan2d");
    TORCH_CHECK(input.dim() == 4 && input.size(0) * input.size(1) == batch_,
                "Plan2d: expected a 4D input with B*C == ", batch_);
    const bool to_sino = (op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW);
    if (to_sino) {
      TORCH_CHECK(input.size(2) == g_.height && input.size(3) == g_.width,
                  "Plan2d: image shape does not match the geometry");
    } else {
      TORCH_CHECK(input.size(2) == views_ && input.size(3) == g_.num_det,
                  "Plan2d: sinogram shape does not match the geometry");
    }
    c10::hip::HIPGuard guard(input.device().index());
    auto out = to_sino
                   ? torch::empty({input.size(0), input.size(1), views_, g_.num_det},
                                  input.options())
                   : torch::empty({input.size(0), input.size(1), g_.height, g_.width},
                                  input.options());
    dd_epilogue ep{nullptr, nullptr, static_cast<float>(alpha), static_cast<float>(beta)};
    if (c.has_value()) {
      check_data(*c, angles_, "Plan2d epilogue term");
      TORCH_CHECK(c->sizes() == out.sizes(), "Plan2d: epilogue term must have the output's shape");
      ep.c = c->data_ptr<float>();
    }
    if (scale.has_value()) {
      check_data(*scale, angles_, "Plan2d epilogue scale");
      TORCH_CHECK(scale->numel() == 1, "Plan2d: epilogue scale must hold one value");
      ep.scale = scale->data_ptr<float>();
    }
    DD_CHECK(dd2d_run_planned_ep(scan_, static_cast<int>(op), &g_, geo_.data_ptr(),
                                 input.data_ptr<float>(), out.data_ptr<float>(),
                                 static_cast<int>(batch_), static_cast<int>(views_), tile_,
                                 variant_, &shape_, &ep, ws_[op].data_ptr(), ws_bytes_[op],
                                 cur_stream(input)),
             "Plan2d run");
    return out;
  }

  int64_t batch() const { return batch_; }
  int64_t tile() const { return tile_; }
  const dd_shape& shape() const { return shape_; }
  int64_t variant() const { return variant_; }
  int64_t views() const { return views_; }

 private:
  int scan_;
  int64_t batch_;
  int tile_ = 0;
  int variant_ = -1;
  dd_shape shape_{0, 0, 0};
  int64_t views_ = 0;
  dd_geom2d g_{};
  torch::Tensor angles_, geo_;
  torch::Tensor ws_[6];
  size_t ws_bytes_[6] = {0, 0, 0, 0, 0, 0};
};

// ---- batched uneven instances ----------------------------------------------
// Instances of different image size, detector and view count, packed into
// one flat image buffer and one flat sinogram buffer; every op is two or
// three launches for the whole batch (ddbatch.hip).
class PlanBatch2d {
 public:
  PlanBatch2d(int scan, const std::vector<torch::Tensor>& options,
              const std::vector<torch::Tensor>& angles, int64_t grid) {
    TORCH_CHECK(scan >= DD_PARA && scan <= DD_FAN_ARC, "PlanBatch2d: 2D scan expected");
    TORCH_CHECK(!options.empty() && options.size() == angles.size(),
                "PlanBatch2d: one options and one angles tensor per instance");
    std::vector<dd2d_batch_item> items(options.size());
    int64_t img = 0, sino = 0;
    for (size_t i = 0; i < options.size(); ++i) {
      TORCH_CHECK(angles[i].is_cuda() && angles[i].scalar_type() == torch::kFloat32 &&
                      angles[i].dim() == 1 && angles[i].is_contiguous(),
                  "PlanBatch2d: angles must be contiguous 1D float32 device tensors");
      angles_.push_back(angles[i]);
      auto& it = items[i];
      it.scan = scan;
      it.geom = decode2d(options[i], "PlanBatch2d");
      it.views = static_cast<int>(angles[i].size(0));
      it.batch = 1;
      it.angles = angles[i].data_ptr<float>();
      it.image_offset = img;
      it.sino_offset = sino;
      image_off_.push_back(img);
      sino_off_.push_back(sino);
      img += static_cast<int64_t>(it.geom.width) * it.geom.height;
      sino += static_cast<int64_t>(it.views) * it.geom.num_det;
      shapes_.push_back({it.geom.height, it.geom.width, it.views, it.geom.num_det});
    }
    image_size_ = img;
    sino_size_ = sino;
    b_ = dd2d_batch_create(items.data(), static_cast<int>(items.size()));
    TORCH_CHECK(b_ != nullptr, "PlanBatch2d: invalid instance description");
    if (grid > 0) dd2d_batch_set_grid(b_, static_cast<int>(grid));
    const auto& a0 = angles_.front();
    c10::hip::HIPGuard guard(a0.device().index());
    mem_ = workspace(dd2d_batch_device_bytes(b_), a0);
    DD_CHECK(dd2d_batch_setup(b_, mem_.data_ptr(), dd2d_batch_device_bytes(b_), cur_stream(a0)),
             "PlanBatch2d setup");
  }
  ~PlanBatch2d() { dd2d_batch_destroy(b_); }
  PlanBatch2d(const PlanBatch2d&) = delete;
  PlanBatch2d& operator=(const PlanBatch2d&) = delete;

  // op on the packed input (flat float32 device tensor); returns a new
  // packed output.
  torch::Tensor run(int64_t op, const torch::Tensor& input) {
    TORCH_CHECK(input.is_cuda() && input.scalar_type() == torch::kFloat32 &&
                    input.is_contiguous(),
                "PlanBatch2d: packed input must be a contiguous float32 device tensor");
    const bool to_sino = (op == DD_PROJ || op == DD_BPROJ_T || op == DD_BPROJ_TW);
    TORCH_CHECK(input.numel() == (to_sino ? image_size_ : sino_size_),
                "PlanBatch2d: packed input has ", input.numel(), " values, expected ",
                to_sino ? image_size_ : sino_size_);
    c10::hip::HIPGuard guard(input.device().index());
    auto out = torch::empty({to_sino ? sino_size_ : image_size_}, input.options());
    DD_CHECK(dd2d_batch_run(b_, static_cast<int>(op), input.data_ptr<float>(),
                            out.data_ptr<float>(), cur_stream(input)),
             "PlanBatch2d run");
    return out;
  }
  void set_grid(int64_t g) { dd2d_batch_set_grid(b_, static_cast<int>(g)); }
  void set_variant(int64_t v) { dd2d_batch_set_variant(b_, static_cast<int>(v)); }
  std::vector<int64_t> image_offsets() const { return image_off_; }
  std::vector<int64_t> sino_offsets() const { return sino_off_; }
  std::vector<std::vector<int64_t>> shapes() const { return shapes_; }
  int64_t image_size() const { return image_size_; }
  int64_t sino_size() const { return sino_size_; }

 private:
  dd2d_batch* b_ = nullptr;
  torch::Tensor mem_;
  std::vector<torch::Tensor> angles_;
  std::vector<int64_t> image_off_, sino_off_;
  std::vector<std::vector<int64_t>> shapes_;
  int64_t image_size_ = 0, sino_size_ = 0;
};

}  // namespace

// ---- op table (RadonTorch.cpp:1375-1418) -------------------------------------
#define DEF2D(NAME, SCAN, OP)                                                   \
  torch::Tensor NAME(torch::Tensor input, torch::Tensor options,               \
                     torch::Tensor angles) {                                   \
    return op2d(SCAN, OP, input, options, angles, #NAME);                      \
  }
#define DEF3D(NAME, SCAN, OP)                                                   \
  torch::Tensor NAME(torch::Tensor input, torch::Tensor options,               \
                     torch::Tensor angles) {                                   \
    return op3d(SCAN, OP, input, options, angles, nullptr, #NAME);             \
  }
#define DEFSP(NAME, OP)                                                         \
  torch::Tensor NAME(torch::Tensor input, torch::Tensor options,               \
                     torch::Tensor angles, torch::Tensor source_pos) {         \
    return op3d(DD_SPIRAL, OP, input, options, angles, &source_pos, #NAME);    \
  }

DEF2D(ProjFanFlatDis2d, DD_FAN_FLAT, DD_PROJ)
DEF2D(ProjTransFanFlatDis2d, DD_FAN_FLAT, DD_PROJ_T)
DEF2D(BackProjFanFlatDis2d, DD_FAN_FLAT, DD_BPROJ)
DEF2D(BackProjTransFanFlatDis2d, DD_FAN_FLAT, DD_BPROJ_T)
DEF2D(BackProjWeightedFanFlatDis2d, DD_FAN_FLAT, DD_BPROJ_W)
DEF2D(BackProjTransWeightedFanFlatDis2d, DD_FAN_FLAT, DD_BPROJ_TW)
DEF2D(ProjFanArcDis2d, DD_FAN_ARC, DD_PROJ)
DEF2D(ProjTransFanArcDis2d, DD_FAN_ARC, DD_PROJ_T)
DEF2D(BackProjFanArcDis2d, DD_FAN_ARC, DD_BPROJ)
DEF2D(BackProjTransFanArcDis2d, DD_FAN_ARC, DD_BPROJ_T)
DEF2D(BackProjWeightedFanArcDis2d, DD_FAN_ARC, DD_BPROJ_W)
DEF2D(BackProjTransWeightedFanArcDis2d, DD_FAN_ARC, DD_BPROJ_TW)
DEF2D(ProjParaDis2d, DD_PARA, DD_PROJ)
DEF2D(ProjTransParaDis2d, DD_PARA, DD_PROJ_T)
DEF2D(BackProjParaDis2d, DD_PARA, DD_BPROJ)
DEF2D(BackProjTransParaDis2d, DD_PARA, DD_BPROJ_T)
DEF3D(ProjConeFlatDis3d, DD_CONE_FLAT, DD_PROJ)
DEF3D(ProjTransConeFlatDis3d, DD_CONE_FLAT, DD_PROJ_T)
DEF3D(BackProjConeFlatDis3d, DD_CONE_FLAT, DD_BPROJ)
DEF3D(BackProjTransConeFlatDis3d, DD_CONE_FLAT, DD_BPROJ_T)
DEF3D(BackProjWeightedConeFlatDis3d, DD_CONE_FLAT, DD_BPROJ_W)
DEF3D(BackProjTransWeightedConeFlatDis3d, DD_CONE_FLAT, DD_BPROJ_TW)
DEF3D(ProjConeArcDis3d, DD_CONE_ARC, DD_PROJ)
DEF3D(ProjTransConeArcDis3d, DD_CONE_ARC, DD_PROJ_T)
DEF3D(BackProjConeArcDis3d, DD_CONE_ARC, DD_BPROJ)
DEF3D(BackProjTransConeArcDis3d, DD_CONE_ARC, DD_BPROJ_T)
DEF3D(BackProjWeightedConeArcDis3d, DD_CONE_ARC, DD_BPROJ_W)
DEF3D(BackProjTransWeightedConeArcDis3d, DD_CONE_ARC, DD_BPROJ_TW)
DEFSP(ProjSpiralDis3d, DD_PROJ)
DEFSP(ProjTransSpiralDis3d, DD_PROJ_T)
DEFSP(BackProjSpiralDis3d, DD_BPROJ)
DEFSP(BackProjTransSpiralDis3d, DD_BPROJ_T)
DEFSP(BackProjWeightedSpiralDis3d, DD_BPROJ_W)
DEFSP(BackProjTransWeightedSpiralDis3d, DD_BPROJ_TW)

torch::Tensor GetSysMatrixFanFlat2d(torch::Tensor options, torch::Tensor angles) {
  return sysmatrix2d(DD_FAN_FLAT, options, angles, "GetSysMatrixFanFlat2d");
}
torch::Tensor GetSysMatrixFanArc2d(torch::Tensor options, torch::Tensor angles) {
  return sysmatrix2d(DD_FAN_ARC, options, angles, "GetSysMatrixFanArc2d");
}
torch::Tensor GetSysMatrixPara2d(torch::Tensor options, torch::Tensor angles) {
  return sysmatrix2d(DD_PARA, options, angles, "GetSysMatrixPara2d");
}

PYBIND11_MODULE(radontorch, m) {
  m.doc() = "deeprecon projector ops on MI355X (libdd)";
#define REG(NAME) m.def(#NAME, &NAME)
  REG(ProjFanFlatDis2d);
  REG(ProjTransFanFlatDis2d);
  REG(BackProjFanFlatDis2d);
  REG(BackProjTransFanFlatDis2d);
  REG(BackProjWeightedFanFlatDis2d);
  REG(BackProjTransWeightedFanFlatDis2d);
  REG(ProjFanArcDis2d);
  REG(ProjTransFanArcDis2d);
  REG(BackProjFanArcDis2d);
  REG(BackProjTransFanArcDis2d);
  REG(BackProjWeightedFanArcDis2d);
  REG(BackProjTransWeightedFanArcDis2d);
  REG(ProjParaDis2d);
  REG(ProjTransParaDis2d);
  REG(BackProjParaDis2d);
  REG(BackProjTransParaDis2d);
  REG(ProjConeFlatDis3d);
  REG(ProjTransConeFlatDis3d);
  REG(BackProjConeFlatDis3d);
  REG(BackProjTransConeFlatDis3d);
  REG(BackProjWeightedConeFlatDis3d);
  REG(BackProjTransWeightedConeFlatDis3d);
  REG(ProjConeArcDis3d);
  REG(ProjTransConeArcDis3d);
  REG(BackProjConeArcDis3d);
  REG(BackProjTransConeArcDis3d);
  REG(BackProjWeightedConeArcDis3d);
  REG(BackProjTransWeightedConeArcDis3d);
  REG(ProjSpiralDis3d);
  REG(ProjTransSpiralDis3d);
  REG(BackProjSpiralDis3d);
  REG(BackProjTransSpiralDis3d);
  REG(BackProjWeightedSpiralDis3d);
  REG(BackProjTransWeightedSpiralDis3d);
  REG(GetSysMatrixFanFlat2d);
  REG(GetSysMatrixFanArc2d);
  REG(GetSysMatrixPara2d);
#undef REG
  pybind11::class_<Plan2d>(m, "Plan2d")
      .def(pybind11::init<int, const torch::Tensor&, const torch::Tensor&, int64_t, int64_t,
                          int64_t, int64_t, int64_t, int64_t>(),
           pybind11::arg("scan"), pybind11::arg("options"), pybind11::arg("angles"),
           pybind11::arg("batch"), pybind11::arg("tile") = 0, pybind11::arg("variant") = -1,
           pybind11::arg("fwd_waves") = 0, pybind11::arg("fwd_views") = 0,
           pybind11::arg("pix_width") = 0)
      .def("run", &Plan2d::run)
      .def("run_ep", &Plan2d::run_ep, pybind11::arg("op"), pybind11::arg("input"),
           pybind11::arg("c") = pybind11::none(), pybind11::arg("scale") = pybind11::none(),
           pybind11::arg("alpha") = 1.0, pybind11::arg("beta") = 0.0)
      .def_property_readonly("shape", [](const Plan2d& p) {
        const dd_shape& sh = p.shape();
        return pybind11::make_tuple(sh.fwd_waves, sh.fwd_views, sh.pix_width);
      })
      .def_property_readonly("tile", &Plan2d::tile)
      .def_property_readonly("variant", &Plan2d::variant)
      .def_property_readonly("batch", &Plan2d::batch)
      .def_property_readonly("views", &Plan2d::views);
  pybind11::class_<PlanBatch2d>(m, "PlanBatch2d")
      .def(pybind11::init<int, const std::vector<torch::Tensor>&,
                          const std::vector<torch::Tensor>&, int64_t>(),
           pybind11::arg("scan"), pybind11::arg("options"), pybind11::arg("angles"),
           pybind11::arg("grid") = 0)
      .def("run", &PlanBatch2d::run)
      .def("set_grid", &PlanBatch2d::set_grid)
      .def("set_variant", &PlanBatch2d::set_variant)
      .def_property_readonly("image_offsets", &PlanBatch2d::image_offsets)
      .def_property_readonly("sino_offsets", &PlanBatch2d::sino_offsets)
      .def_property_readonly("shapes", &PlanBatch2d::shapes)
      .def_property_readonly("image_size", &PlanBatch2d::image_size)
      .def_property_readonly("sino_size", &PlanBatch2d::sino_size);
  m.def("profile_enable", [](bool on) { dd_profile_enable(on ? 1 : 0); });
  m.def("profile_read", [](int cls) {
    double ms = 0.0;
    long long n = 0;
    dd_profile_read(cls, &ms, &n);
    return pybind11::make_tuple(ms, n);
  });
  m.def("views2d", &views2d, pybind11::arg("scan"), pybind11::arg("op"), pybind11::arg("input"),
        pybind11::arg("options"), pybind11::arg("angles"));
  m.def("set_instance_tile", [](int t) { dd2d_set_instance_tile(t); });
  m.def("set_adjoint_variant", [](int v) { dd2d_set_adjoint_variant(v); });
  m.def("set_compensated3d", [](int mode) { dd3d_set_compensated(mode); },
        pybind11::arg("mode"));
  m.def("set_adjoint_staging3d", [](int on) { dd3d_set_adjoint_staging(on); },
        pybind11::arg("on"));
  m.def("set_forward_order3d", [](int mode) { dd3d_set_forward_order(mode); },
        pybind11::arg("mode"));
  m.attr("ADJ_PER_PIXEL") = static_cast<int>(DD_ADJ_PER_PIXEL);
  m.attr("ADJ_TILE16") = static_cast<int>(DD_ADJ_TILE16);
  m.attr("ADJ_TILE8") = static_cast<int>(DD_ADJ_TILE8);
  m.attr("ADJ_QUAD") = static_cast<int>(DD_ADJ_QUAD);
  m.attr("ADJ_BLOCK_WINDOW") = static_cast<int>(DD_ADJ_BLOCK_WINDOW);
  m.attr("FWD_PREFIX") = static_cast<int>(DD_FWD_PREFIX);
  m.attr("ADJ_BLOCK_PREFIX") = static_cast<int>(DD_ADJ_BLOCK_PREFIX);
  m.attr("PROJ") = static_cast<int>(DD_PROJ);
  m.attr("PROJ_T") = static_cast<int>(DD_PROJ_T);
  m.attr("BPROJ") = static_cast<int>(DD_BPROJ);
  m.attr("BPROJ_T") = static_cast<int>(DD_BPROJ_T);
  m.attr("BPROJ_W") = static_cast<int>(DD_BPROJ_W);
  m.attr("BPROJ_TW") = static_cast<int>(DD_BPROJ_TW);
  m.attr("SCAN") = pybind11::dict(pybind11::arg("para") = static_cast<int>(DD_PARA),
                                  pybind11::arg("flat") = static_cast<int>(DD_FAN_FLAT),
                                  pybind11::arg("arc") = static_cast<int>(DD_FAN_ARC));
}
