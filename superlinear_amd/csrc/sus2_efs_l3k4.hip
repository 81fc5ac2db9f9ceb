// EFS kernel instances for L=3, K=4.
#include "sus2_efs_impl.hpp"

namespace sus2 {
template void run_efs_shape<float, 3, 4>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
template void run_efs_shape<double, 3, 4>(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
}  // namespace sus2
