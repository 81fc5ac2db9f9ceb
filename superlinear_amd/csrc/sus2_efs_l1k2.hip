// EFS kernel instances for L=1, K=2.
#include "sus2_efs_impl.hpp"

namespace sus2 {
template void run_efs_shape<float, 1, 2>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
template void run_efs_shape<double, 1, 2>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
}  // namespace sus2
