// EFS kernel instances for L=5, K=6, B=10, warp sharing mode 2.
#include "sus2_efs_impl.hpp"

namespace sus2 {
template void run_efs_shape<Cfg<float, 5, 6, 10, 2>>(const DevModelSet&, Workspace&, const Batch&,
                                                          const EfsOut&, hipStream_t);
template void run_efs_shape<Cfg<double, 5, 6, 10, 2>>(const DevModelSet&, Workspace&, const Batch&,
                                                           const EfsOut&, hipStream_t);
}  // namespace sus2
