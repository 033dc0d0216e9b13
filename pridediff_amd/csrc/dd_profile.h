// Optional HIP-event timing of the dominant gather kernels.
//
// When enabled (dd_profile_enable), each launch of a gather kernel is
// bracketed by two events recorded on the launch stream; dd_profile_read
// waits for them and returns the summed kernel time per class.  Nothing is
// recorded while the stream is being captured into a graph.
#pragma once

#include <hip/hip_runtime.h>

#include "dd.h"

namespace dd {

class ProfileScope {
 public:
  ProfileScope(int kernel_class, hipStream_t st);
  ~ProfileScope();

 private:
  int cls_;
  hipStream_t st_;
  hipEvent_t start_ = nullptr;
};

}  // namespace dd
