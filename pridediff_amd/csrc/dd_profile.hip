// HIP-event timing of gather kernels (see dd_profile.h).
#include "dd_profile.h"

#include <atomic>
#include <mutex>
#include <vector>

namespace dd {
namespace {

struct Pending {
  int cls;
  hipEvent_t start, stop;
};

std::atomic<int> g_enabled{0};
std::mutex g_mu;
std::vector<Pending> g_pending;
double g_total_ms[DD_KERNEL_CLASSES] = {};
long long g_count[DD_KERNEL_CLASSES] = {};

bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &status) != hipSuccess) return true;
  return status != hipStreamCaptureStatusNone;
}

}  // namespace

ProfileScope::ProfileScope(int kernel_class, hipStream_t st)
    : cls_(kernel_class), st_(st) {
  if (!g_enabled.load() || stream_capturing(st)) return;
  if (hipEventCreate(&start_) != hipSuccess) {
    start_ = nullptr;
    return;
  }
  (void)hipEventRecord(start_, st_);
}

ProfileScope::~ProfileScope() {
  if (!start_) return;
  hipEvent_t stop = nullptr;
  if (hipEventCreate(&stop) != hipSuccess) {
    (void)hipEventDestroy(start_);
    return;
  }
  (void)hipEventRecord(stop, st_);
  std::lock_guard<std::mutex> lock(g_mu);
  g_pending.push_back({cls_, start_, stop});
}

}  // namespace dd

extern "C" {

void dd_profile_enable(int on) { dd::g_enabled.store(on ? 1 : 0); }

int dd_profile_read(int kernel_class, double* total_ms, long long* launches) {
  using namespace dd;
  if (kernel_class < 0 || kernel_class >= DD_KERNEL_CLASSES) return DD_BAD_ARG;
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto& p : g_pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(p.stop) == hipSuccess &&
        hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      g_total_ms[p.cls] += ms;
      g_count[p.cls] += 1;
    }
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  g_pending.clear();
  if (total_ms) *total_ms = g_total_ms[kernel_class];
  if (launches) *launches = g_count[kernel_class];
  g_total_ms[kernel_class] = 0.0;
  g_count[kernel_class] = 0;
  return DD_OK;
}

}  // extern "C"
