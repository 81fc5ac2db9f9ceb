// Parameter adjoint (forward-over-reverse) of the SUS2-SH EFS path.
#include <hip/hip_runtime.h>

#include "../../include/sus2amd.h"
#include "sus2_runtime.hpp"

namespace sus2 {
void grad_dispatch(const DevModelSet&, Workspace&, int, const Batch&, const double*, const double*, const double*,
                   double*, hipStream_t) {
  throw Error(SUS2_E_UNSUPPORTED, "parameter gradient kernels are not built yet");
}
}  // namespace sus2

extern "C" int sus2_loss_grad(sus2_model*, int, int, int, const int64_t*, const int32_t*, const double*,
                              const double*, const int32_t*, const double*, const double*, const double*, int, double,
                              double, double, double*, double*, void*) {
  return SUS2_E_UNSUPPORTED;
}
