// 3D distance-driven projectors: not yet implemented in this file's first
// revision; every entry point reports DD_UNSUPPORTED.
#include <hip/hip_runtime.h>

#include "dd3d.h"

extern "C" {

size_t dd3d_workspace_size(int, int, const dd_geom3d*, int, int) { return 1; }

dd_status dd3d_run(int, int, const dd_geom3d*, const float*, float*,
                   const float*, const float*, int, int, void*, size_t,
                   void*) {
  return DD_UNSUPPORTED;
}

}  // extern "C"
