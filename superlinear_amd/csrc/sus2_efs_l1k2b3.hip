// EFS and parameter-gradient kernel instances for L=1, K=2, B=3, warp sharing mode 3.
#include "sus2_grad_impl.hpp"

namespace sus2 {
template void run_efs_shape<Cfg<float, 1, 2, 3, 3>>(const DevModelSet&, Workspace&, const Batch&,
                                                          const EfsOut&, hipStream_t);
template void run_efs_shape<Cfg<double, 1, 2, 3, 3>>(const DevModelSet&, Workspace&, const Batch&,
                                                           const EfsOut&, hipStream_t);
template void run_grad_shape<Cfg<float, 1, 2, 3, 3>>(const DevModelSet&, Workspace&, const Batch&, const GradWeights&,
                                              double*, double, bool, bool, hipStream_t);
template void run_grad_shape<Cfg<double, 1, 2, 3, 3>>(const DevModelSet&, Workspace&, const Batch&, const GradWeights&,
                                              double*, double, bool, bool, hipStream_t);
template void pair_stage_shape<Cfg<float, 1, 2, 3, 3>>(int, const DevModelSet&, Workspace&, const PairView&,
                                                 hipStream_t);
template void pair_stage_shape<Cfg<double, 1, 2, 3, 3>>(int, const DevModelSet&, Workspace&, const PairView&,
                                                 hipStream_t);
}  // namespace sus2
