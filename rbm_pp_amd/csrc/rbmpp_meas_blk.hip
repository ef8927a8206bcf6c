// The block-form estimator (k_meas_blk<false>); the kernel itself is in
// rbmpp_meas_blk_impl.h.
#include "rbmpp_meas_blk_impl.h"

void launch_meas_blk(const LocalArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_meas_blk<false>, dim3(grid), dim3(BNT), 0, st, a);
}
