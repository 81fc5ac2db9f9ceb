// EFS kernel instances for L=5, K=6.
#include "sus2_efs_impl.hpp"

namespace sus2 {
template void run_efs_shape<float, 5, 6>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
template void run_efs_shape<double, 5, 6>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
}  // namespace sus2
