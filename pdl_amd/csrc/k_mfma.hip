// MFMA matmult paths for float ndarrays (gfx950).  Until the split-FP16 and
// FP8 kernels are in place this reports hipErrorNotSupported and the caller
// uses the FP32 SIMT kernel.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

hipError_t launch_mfma_f32(const MmPlan &p, const float *a, const float *b, float *c, int mode, hipStream_t s) {
  (void)p; (void)a; (void)b; (void)c; (void)mode; (void)s;
  return hipErrorNotSupported;
}

}  // namespace pdlhip
