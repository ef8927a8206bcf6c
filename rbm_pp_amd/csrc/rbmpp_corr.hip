// The estimator with the correlation run's pair pass (k_meas<true>,
// rbmpp_meas_impl.h) and the fixed-order sum of its per-workgroup partials.
#include "rbmpp_meas_impl.h"

namespace {

// out[k] = sum over chunks c (in order) of part[c * stride + k], k < len
__global__ void k_sum_strided(int len, int nchunk, int stride, const double* __restrict__ part, double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  double t = 0.0;
  for (int c = 0; c < nchunk; c++) t += part[(size_t)c * stride + k];
  out[k] = t;
}

}  // namespace

void launch_meas_corr(const LocalArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_meas<true>, dim3(grid), dim3(MNT), 0, st, a);
}

void launch_sum_strided(int len, int nchunk, int stride, const double* part, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_strided, dim3((len + 255) / 256), dim3(256), 0, st, len, nchunk, stride, part, out);
}
