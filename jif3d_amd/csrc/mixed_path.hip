// Mixed-precision path (filled in by the mixed kernels).
#include <hip/hip_runtime.h>

#include "context.hpp"

namespace jgm {

namespace {
__global__ void k_fill_zero(double* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
}  // namespace

hipError_t launch_fill_zero(jgm_context* ctx, double* p, size_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p,
                     n);
  return hipGetLastError();
}

hipError_t build_near_boxes(jgm_context*) { return hipErrorNotSupported; }
hipError_t launch_forward_mixed(jgm_context*, int, const double*, double*, double*) {
  return hipErrorNotSupported;
}
hipError_t launch_adjoint_mixed(jgm_context*, int, const double*, const double*, double*) {
  return hipErrorNotSupported;
}

}  // namespace jgm
