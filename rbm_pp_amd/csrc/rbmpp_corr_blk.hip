// The block-form estimator with the correlation run's pair pass
// (k_meas_blk<true>, rbmpp_meas_blk_impl.h), in a translation unit of its own
// so that it does not share register allocation with the plain estimator.
#include "rbmpp_meas_blk_impl.h"

void launch_meas_blk_corr(const LocalArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_meas_blk<true>, dim3(grid), dim3(BNT), 0, st, a);
}
