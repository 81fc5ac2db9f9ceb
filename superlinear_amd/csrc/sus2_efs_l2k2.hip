// EFS kernel instances for L=2, K=2.
#include "sus2_efs_impl.hpp"

namespace sus2 {
template void run_efs_shape<float, 2, 2>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
template void run_efs_shape<double, 2, 2>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
}  // namespace sus2
