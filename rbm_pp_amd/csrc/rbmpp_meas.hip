// The per-configuration estimator for N <= 104 (k_meas<false>); the kernel
// itself is in rbmpp_meas_impl.h.
#include "rbmpp_meas_impl.h"

void launch_meas(const LocalArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_meas<false>, dim3(grid), dim3(MNT), 0, st, a);
}
