// EFS and parameter-gradient kernel instances for L=5, K=6, B=10, warp sharing mode 2.
#include "sus2_grad_impl.hpp"

namespace sus2 {
template void run_efs_shape<Cfg<float, 5, 6, 10, 2>>(const DevModelSet&, Workspace&, const Batch&,
                                                          const EfsOut&, hipStream_t);
template void run_efs_shape<Cfg<double, 5, 6, 10, 2>>(const DevModelSet&, Workspace&, const Batch&,
                                                           const EfsOut&, hipStream_t);
template void run_grad_shape<Cfg<float, 5, 6, 10, 2>>(const DevModelSet&, Workspace&, const Batch&, const GradWeights&,
                                              double*, double, bool, bool, hipStream_t);
template void run_grad_shape<Cfg<double, 5, 6, 10, 2>>(const DevModelSet&, Workspace&, const Batch&, const GradWeights&,
                                              double*, double, bool, bool, hipStream_t);
template void pair_stage_shape<Cfg<float, 5, 6, 10, 2>>(int, const DevModelSet&, Workspace&, const PairView&,
                                                 hipStream_t);
template void pair_stage_shape<Cfg<double, 5, 6, 10, 2>>(int, const DevModelSet&, Workspace&, const PairView&,
                                                 hipStream_t);
}  // namespace sus2
