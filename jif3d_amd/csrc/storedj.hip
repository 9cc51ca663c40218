// Stored-J path (filled in by the stored-J kernels).
#include <hip/hip_runtime.h>

#include "context.hpp"

namespace jgm {

hipError_t launch_storedj_build(jgm_context*, int, int) { return hipErrorNotSupported; }
hipError_t launch_storedj_apply(jgm_context*, int, int, const float*, size_t, float*) {
  return hipErrorNotSupported;
}

}  // namespace jgm
